// SimInfo host reductions (reference source/simInfo.cpp:50-127).
#include "pmsim/simInfo.h"
#include "pmsim/unitConversions.h"

float SimInfo::potentialEnergy(const Grid& grid, const std::vector<Particle>& particles,
                               std::function<float(Vec3)> externalPotential, float H, float DT, float G) {
  double internal = 0.0;
  for (int i = 0; i < grid.getLength(); ++i) {
    auto [x, y, z] = grid.indexTripleFromFlat(i);
    internal += (double)densityToOriginalUnits(grid.getDensity(x, y, z), DT, G) *
                potentialToOriginalUnits(grid.getPotential(x, y, z), H, DT);
  }
  double external = 0.0;
  for (const auto& p : particles) external += (double)p.mass * externalPotential(p.position);
  return (float)(0.5 * H * H * H * internal + external);
}

float SimInfo::kineticEnergy(const std::vector<Particle>& particles) {
  double ke = 0.0;
  for (const auto& p : particles) ke += 0.5 * p.mass * p.velocity.getMagnitudeSquared();
  return (float)ke;
}

Vec3 SimInfo::totalMomentum(const std::vector<Particle>& particles) {
  Vec3 m;
  for (const auto& p : particles) m += p.mass * p.integerStepVelocity;
  return m;
}

Vec3 SimInfo::totalAngularMomentum(const std::vector<Particle>& particles) {
  Vec3 l;
  for (const auto& p : particles) l += p.mass * p.position.cross(p.integerStepVelocity);
  return l;
}

void SimInfo::setInitialMomentum(const std::vector<Particle>& particles) {
  expectedMomentum = totalMomentum(particles);
}

Vec3 SimInfo::updateExpectedMomentum(Vec3 externalForce, float DT) {
  expectedMomentum += DT * externalForce;
  return expectedMomentum;
}
