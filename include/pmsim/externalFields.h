// External fields.  The reference passes a std::function<Vec3(Vec3)> to
// PMMethod.  A GPU cannot call a host closure, so the drop-in recognises by
// type the callables it evaluates on the device:
//   - SphRadDecr (sphRadDecrField, source/externalFields.cpp:4-16, the
//     field of the reference's galaxy demos), in the interpolation kernel;
//   - ZeroField, and an empty std::function: no field.
// Any other callable (the reference's demos pass lambdas, e.g. one wrapping
// sphRadDecrField) is evaluated on the host for every particle after each
// force evaluation and added on the device (pmgpu_add_accelerations): exact,
// at the cost of one position read-back per step.
#pragma once

#include "vec3.h"

Vec3 sphRadDecrField(Vec3 pos, Vec3 center, float R, float M, float G);
float sphRadDecrFieldPotential(Vec3 pos, Vec3 center, float R, float M, float G);

struct ZeroField {
  Vec3 operator()(Vec3) const { return Vec3::zero(); }
};

struct SphRadDecr {
  Vec3 center;
  float R, M, G;
  Vec3 operator()(Vec3 pos) const { return sphRadDecrField(pos, center, R, M, G); }
};

struct SphRadDecrPotential {
  Vec3 center;
  float R, M, G;
  float operator()(Vec3 pos) const { return sphRadDecrFieldPotential(pos, center, R, M, G); }
};
