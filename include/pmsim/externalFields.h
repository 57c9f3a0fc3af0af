// External fields.  The reference passes a std::function<Vec3(Vec3)> to
// PMMethod; a GPU cannot call a host closure, so the drop-in recognises the
// callables it can evaluate on the device:
//   - SphRadDecr (sphRadDecrField, source/externalFields.cpp:4-16, the
//     field of the reference's galaxy demos) and ZeroField, by type;
//   - any other callable that returns exactly zero on every probe point is
//     taken as a zero field (the reference's demos pass such lambdas);
// anything else makes the PMMethod constructor throw std::invalid_argument.
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
