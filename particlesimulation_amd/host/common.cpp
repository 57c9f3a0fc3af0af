// Free functions of the drop-in host API: unit conversions
// (reference source/unitConversions.cpp:23-71), leapfrog on host vectors
// (source/leapfrog.cpp:5-24), external fields (source/externalFields.cpp) and
// the Green's functions (source/greensFunctions.cpp, evaluated by libpmgpu).
#include <cmath>
#include <complex>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "pmgpu.h"
#include "pmsim/externalFields.h"
#include "pmsim/greensFunctions.h"
#include "pmsim/leapfrog.h"
#include "pmsim/unitConversions.h"

void stateToCodeUnits(std::vector<Vec3>& state, float H, float DT) {
  const size_t n = state.size() / 2;
  for (size_t i = 0; i < n; ++i) state[i] = positionToCodeUntits(state[i], H);
  for (size_t i = n; i < state.size(); ++i) state[i] = velocityToCodeUntits(state[i], H, DT);
}

void stateToOriginalUnits(std::vector<Vec3>& state, float H, float DT) {
  const size_t n = state.size() / 2;
  for (size_t i = 0; i < n; ++i) state[i] = positionToOriginalUnits(state[i], H);
  for (size_t i = n; i < state.size(); ++i) state[i] = velocityToOriginalUnits(state[i], H, DT);
}

void stateToCodeUnits(std::vector<Particle>& particles, float H, float DT) {
  for (auto& p : particles) {
    p.position = positionToCodeUntits(p.position, H);
    p.velocity = velocityToCodeUntits(p.velocity, H, DT);
  }
}

void stateToOriginalUnits(std::vector<Particle>& particles, float H, float DT) {
  for (auto& p : particles) {
    p.position = positionToOriginalUnits(p.position, H);
    p.velocity = velocityToOriginalUnits(p.velocity, H, DT);
  }
}

void velocitiesToCodeUnits(std::vector<Vec3>& v, float H, float DT) {
  for (auto& x : v) x = velocityToCodeUntits(x, H, DT);
}

void velocitiesToOriginalUnits(std::vector<Vec3>& v, float H, float DT) {
  for (auto& x : v) x = velocityToOriginalUnits(x, H, DT);
}

void integerStepVelocitiesToOriginalUnits(std::vector<Particle>& particles, float H, float DT) {
  for (auto& p : particles) p.integerStepVelocity = velocityToOriginalUnits(p.integerStepVelocity, H, DT);
}

void integerStepVelocitiesToCodeUnits(std::vector<Particle>& particles, float H, float DT) {
  for (auto& p : particles) p.integerStepVelocity = velocityToCodeUntits(p.integerStepVelocity, H, DT);
}

void massToCodeUnits(std::vector<Particle>& particles, float H, float DT, float G) {
  for (auto& p : particles) p.mass = massToCodeUnits(p.mass, H, DT, G);
}

void massToOriginalUnits(std::vector<Particle>& particles, float H, float DT, float G) {
  for (auto& p : particles) p.mass = massToOriginalUnits(p.mass, H, DT, G);
}

void setHalfStepVelocities(std::vector<Particle>& particles, float dt) {
  for (auto& p : particles) p.velocity += 0.5f * dt * p.acceleration;
}

void setIntegerStepVelocities(std::vector<Particle>& particles, float dt) {
  for (auto& p : particles) p.integerStepVelocity = p.velocity + 0.5f * dt * p.acceleration;
}

void updateVelocities(std::vector<Particle>& particles, float dt) {
  for (auto& p : particles) p.velocity += dt * p.acceleration;
}

void updatePositions(std::vector<Particle>& particles, float dt) {
  for (auto& p : particles) p.position += dt * p.velocity;
}

Vec3 sphRadDecrField(Vec3 pos, Vec3 center, float R, float M, float G) {
  const Vec3 d = pos - center;
  const float r = d.getMagnitude();
  const float g = r > R ? -G * M / (r * r) : -(G * M / (R * R * R)) * r * (4 - 3 * r / R);
  return g * (d / r);
}

float sphRadDecrFieldPotential(Vec3 pos, Vec3 center, float R, float M, float G) {
  const float r = (pos - center).getMagnitude();
  if (r > R) return -G * M / r;
  const float u = r / R;
  return G * M / R * (-2 + u * u * (2 - u));
}

static std::complex<float> green(int gf, InterpolationScheme is, FiniteDiffScheme fds, float a, int kx,
                                 int ky, int kz, std::tuple<int, int, int> dims) {
  const double v = pmgpu_greens_value(gf, (int)is, (int)fds, (double)a, kx, ky, kz, std::get<0>(dims),
                                      std::get<1>(dims), std::get<2>(dims));
  if (std::isnan(v)) throw std::invalid_argument("not implemented");
  return std::complex<float>((float)v, 0.0f);
}

std::complex<float> GreenOptimal(InterpolationScheme is, int kx, int ky, int kz,
                                 std::tuple<int, int, int> dims, float a, CloudShape cs,
                                 FiniteDiffScheme fds) {
  return green(cs == CloudShape::S1 ? PMGPU_S1_OPTIMAL : PMGPU_S2_OPTIMAL, is, fds, a, kx, ky, kz, dims);
}

std::complex<float> GreenDiscreteLaplacian(int kx, int ky, int kz, std::tuple<int, int, int> dims) {
  return green(PMGPU_DISCRETE_LAPLACIAN, InterpolationScheme::CIC, FiniteDiffScheme::TWO_POINT, 0.0f,
               kx, ky, kz, dims);
}

std::complex<float> GreenPoorMan(int kx, int ky, int kz, std::tuple<int, int, int> dims) {
  return green(PMGPU_POOR_MAN, InterpolationScheme::CIC, FiniteDiffScheme::TWO_POINT, 0.0f, kx, ky, kz,
               dims);
}
