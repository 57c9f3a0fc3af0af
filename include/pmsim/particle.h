// Particle record of the drop-in host API: the reference's AoS layout
// (include/particle.h:5-15), 220 bytes, so that code written against
// std::vector<Particle> keeps working.  The GPU keeps its own SoA copy.
#pragma once

#include <array>

#include "vec3.h"

struct Particle {
  Vec3 position;
  Vec3 velocity;
  Vec3 acceleration;
  float mass;
  Vec3 integerStepVelocity;
  Vec3 shortRangeForce;
  std::array<Vec3, 13> shortRangeFromNeighbor;

  Particle(Vec3 position_, Vec3 velocity_, float mass_)
      : position(position_), velocity(velocity_), mass(mass_), integerStepVelocity(velocity_) {}
};

static_assert(sizeof(Particle) == 220, "Particle must keep the reference's 220-byte layout");
