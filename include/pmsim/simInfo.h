// Diagnostics of the reference (include/simInfo.h), host versions over a
// std::vector<Particle>.  PMMethod::run / P3MMethod::run use the device's
// FP64 reductions (pmgpu_diagnostics) instead.
#pragma once

#include <functional>
#include <vector>

#include "grid.h"
#include "particle.h"
#include "vec3.h"

class SimInfo {
 public:
  static float potentialEnergy(const Grid& grid, const std::vector<Particle>& particles,
                               std::function<float(Vec3)> externalPotential, float H, float DT,
                               float G);
  static float kineticEnergy(const std::vector<Particle>& particles);
  static Vec3 totalMomentum(const std::vector<Particle>& particles);
  static Vec3 totalAngularMomentum(const std::vector<Particle>& particles);
  void setInitialMomentum(const std::vector<Particle>& particles);
  void setInitialMomentum(Vec3 momentum) { expectedMomentum = momentum; }
  Vec3 updateExpectedMomentum(Vec3 externalForce, float DT);

 private:
  Vec3 expectedMomentum;
};
