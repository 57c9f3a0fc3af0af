// Leapfrog pieces of the reference (include/leapfrog.h:6-12) on a host
// std::vector<Particle>.  PMMethod::run and P3MMethod::run do not use these:
// they drive the device kernels (pmgpu_leapfrog / the fused pmgpu_run).
#pragma once

#include <vector>

#include "particle.h"

void setHalfStepVelocities(std::vector<Particle>& particles, float dt = 1.0f);
void setIntegerStepVelocities(std::vector<Particle>& particles, float dt = 1.0f);
void updateVelocities(std::vector<Particle>& particles, float dt = 1.0f);
void updatePositions(std::vector<Particle>& particles, float dt = 1.0f);
