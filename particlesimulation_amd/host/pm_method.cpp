// PMMethod on the GPU (reference source/pmMethod.cpp), driving libpmgpu.
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

#include "pmgpu.h"
#include "pmsim/externalFields.h"
#include "pmsim/leapfrog.h"
#include "pmsim/pmMethod.h"
#include "pmsim/simInfo.h"
#include "pmsim/unitConversions.h"

namespace {

void check(int rc, pmgpu_ctx* ctx, const char* what) {
  if (rc == PMGPU_OK) return;
  const std::string msg = std::string(what) + ": " + pmgpu_last_error(ctx);
  if (rc == PMGPU_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

}  // namespace

PMMethod::PMMethod(const std::vector<Vec3>& state, const std::vector<float>& masses,
                   const std::tuple<float, float, float> effectiveBoxSize,
                   const std::function<Vec3(Vec3)> externalField_,
                   const std::function<float(Vec3)> externalPotential_, const float H_, const float DT_,
                   const float G_, const InterpolationScheme is_, const FiniteDiffScheme fds_,
                   const GreensFunction gFunc_, const float particleDiameter_, Grid& grid_)
    : grid(grid_),
      effectiveBoxSizeX(std::get<0>(effectiveBoxSize)),
      effectiveBoxSizeY(std::get<1>(effectiveBoxSize)),
      effectiveBoxSizeZ(std::get<2>(effectiveBoxSize)),
      N(int(masses.size())),
      externalField(externalField_),
      externalPotential(externalPotential_),
      H(H_),
      DT(DT_),
      G(G_),
      is(is_),
      fds(fds_),
      gFunc(gFunc_),
      particleDiameter(lengthToCodeUnits(particleDiameter_, H_)) {
  if (state.size() < 2 * masses.size()) throw std::invalid_argument("state must hold positions then velocities");
  particles.reserve(N);
  for (int i = 0; i < N; ++i) particles.emplace_back(state[i], state[N + i], masses[i]);

  pmgpu_pm_params p{};
  auto [gx, gy, gz] = grid.getGridPoints();
  p.grid[0] = gx;
  p.grid[1] = gy;
  p.grid[2] = gz;
  p.interpolation = (int32_t)is;
  p.finite_diff = (int32_t)fds;
  p.greens_function = (int32_t)gFunc;
  p.particle_diameter = particleDiameter;
  p.box[0] = effectiveBoxSizeX / H;
  p.box[1] = effectiveBoxSizeY / H;
  p.box[2] = effectiveBoxSizeZ / H;
  p.H = H;
  p.DT = DT;
  if (const SphRadDecr* f = externalField.target<SphRadDecr>()) {
    p.ext_kind = PMGPU_EXT_SPH_RAD_DECR;
    float prm[6] = {f->center.x, f->center.y, f->center.z, f->R, f->M, f->G};
    std::memcpy(p.ext_params, prm, sizeof(prm));
  } else {
    // ZeroField and an empty function add nothing; any other callable (the
    // reference's demos pass lambdas) is evaluated on the host after every
    // force evaluation, as the reference evaluates it per particle
    p.ext_kind = PMGPU_EXT_NONE;
    hostField = externalField && !externalField.target<ZeroField>();
  }
  check(pmgpu_create(&p, &ctx), nullptr, "pmgpu_create");
  grid.attach(ctx);
  upload();
}

PMMethod::~PMMethod() {
  if (ctx) pmgpu_destroy(ctx);
}

uint64_t PMMethod::mirrorHash() const {
  // FNV-1a over the particle records
  const unsigned char* b = reinterpret_cast<const unsigned char*>(particles.data());
  const size_t n = particles.size() * sizeof(Particle);
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

void PMMethod::upload() {
  const float mscale = inCodeUnits ? 1.0f : massToCodeUnits(1.0f, H, DT, G);
  std::vector<float> pos(3 * (size_t)N), vel(3 * (size_t)N), m(N), acc(3 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    const Particle& q = particles[i];
    Vec3 x = inCodeUnits ? q.position : positionToCodeUntits(q.position, H);
    Vec3 v = inCodeUnits ? q.velocity : velocityToCodeUntits(q.velocity, H, DT);
    Vec3 a = inCodeUnits ? q.acceleration : accelerationToCodeUnits(q.acceleration, H, DT);
    pos[3 * i] = x.x, pos[3 * i + 1] = x.y, pos[3 * i + 2] = x.z;
    vel[3 * i] = v.x, vel[3 * i + 1] = v.y, vel[3 * i + 2] = v.z;
    acc[3 * i] = a.x, acc[3 * i + 1] = a.y, acc[3 * i + 2] = a.z;
    m[i] = inCodeUnits ? q.mass : q.mass * mscale;
  }
  check(pmgpu_set_particles(ctx, N, pos.data(), vel.data(), m.data()), ctx, "pmgpu_set_particles");
  check(pmgpu_update_particles(ctx, nullptr, nullptr, acc.data()), ctx, "pmgpu_update_particles");
  mirrorStale = false;
  mirrorSeen = mirrorHash();
}

void PMMethod::uploadIfModified() {
  if (!mirrorStale && mirrorHash() != mirrorSeen) upload();
}

void PMMethod::recordShortRange() {
  check(pmgpu_p3m_record_short_range(ctx, 1), ctx, "pmgpu_p3m_record_short_range");
  srRecorded = true;
}

void PMMethod::downloadMirror() {
  std::vector<float> pos(3 * (size_t)N), vel(3 * (size_t)N), acc(3 * (size_t)N), sr;
  check(pmgpu_get_particles(ctx, pos.data(), vel.data(), acc.data()), ctx, "pmgpu_get_particles");
  if (srRecorded) {
    sr.resize(3 * (size_t)N);
    check(pmgpu_p3m_get_short_range(ctx, sr.data()), ctx, "pmgpu_p3m_get_short_range");
  }
  for (int i = 0; i < N; ++i) {
    Particle& q = particles[i];
    Vec3 x(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    Vec3 v(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    Vec3 a(acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]);
    q.position = inCodeUnits ? x : positionToOriginalUnits(x, H);
    q.velocity = inCodeUnits ? v : velocityToOriginalUnits(v, H, DT);
    q.acceleration = inCodeUnits ? a : accelerationToOriginalUnits(a, H, DT);
    if (srRecorded) {
      // the reference's shortRangeForce + shortRangeFromNeighbor hold the
      // short-range force in code units (source/p3mMethod.cpp:269-272); the
      // GPU pass has no per-neighbour slots, so their sum is kept here
      const float mc = inCodeUnits ? q.mass : massToCodeUnits(q.mass, H, DT, G);
      q.shortRangeForce = mc * Vec3(sr[3 * i], sr[3 * i + 1], sr[3 * i + 2]);
      q.shortRangeFromNeighbor.fill(Vec3::zero());
    }
  }
  mirrorStale = false;
  mirrorSeen = mirrorHash();
}

void PMMethod::setCodeUnits(bool code) {
  if (code == inCodeUnits) return;
  if (mirrorStale) {
    // positions, velocities and accelerations come from the device in the new
    // units at the next download; masses are host-only and convert here
    for (auto& q : particles) q.mass = code ? massToCodeUnits(q.mass, H, DT, G) : massToOriginalUnits(q.mass, H, DT, G);
    inCodeUnits = code;
    return;
  }
  uploadIfModified();
  for (auto& q : particles) {
    if (code) {
      q.position = positionToCodeUntits(q.position, H);
      q.velocity = velocityToCodeUntits(q.velocity, H, DT);
      q.acceleration = accelerationToCodeUnits(q.acceleration, H, DT);
      q.mass = massToCodeUnits(q.mass, H, DT, G);
      q.integerStepVelocity = velocityToCodeUntits(q.integerStepVelocity, H, DT);
    } else {
      q.position = positionToOriginalUnits(q.position, H);
      q.velocity = velocityToOriginalUnits(q.velocity, H, DT);
      q.acceleration = accelerationToOriginalUnits(q.acceleration, H, DT);
      q.mass = massToOriginalUnits(q.mass, H, DT, G);
      q.integerStepVelocity = velocityToOriginalUnits(q.integerStepVelocity, H, DT);
    }
  }
  inCodeUnits = code;
  mirrorSeen = mirrorHash();
}

std::vector<Particle>& PMMethod::getParticles() {
  if (mirrorStale) downloadMirror();
  return particles;
}

void PMMethod::initGreensFunction() {
  check(pmgpu_init_greens(ctx), ctx, "pmgpu_init_greens");
  greensReady = true;
  grid.markDeviceNewer();
}

void PMMethod::pmMethodStep() {
  uploadIfModified();
  if (!greensReady) initGreensFunction();
  check(pmgpu_pm_step(ctx), ctx, "pmgpu_pm_step");
  applyHostField();
  mirrorStale = true;
  grid.markDeviceNewer();
}

// The external-field term of updateAccelerations (source/pmMethod.cpp:
// 384-390) for a callable the device cannot evaluate: the positions come
// back in code units, the field is evaluated at positionToOriginalUnits and
// its accelerationToCodeUnits is added on the device.
void PMMethod::applyHostField() {
  if (!hostField) return;
  std::vector<float> pos(3 * (size_t)N), add(3 * (size_t)N);
  check(pmgpu_get_particles(ctx, pos.data(), nullptr, nullptr), ctx, "pmgpu_get_particles");
  for (int i = 0; i < N; ++i) {
    const Vec3 x(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    const Vec3 a = accelerationToCodeUnits(externalField(positionToOriginalUnits(x, H)), H, DT);
    add[3 * i] = a.x;
    add[3 * i + 1] = a.y;
    add[3 * i + 2] = a.z;
  }
  check(pmgpu_add_accelerations(ctx, add.data()), ctx, "pmgpu_add_accelerations");
}

bool PMMethod::escapedComputationalBox() {
  uploadIfModified();
  int32_t e = 0;
  check(pmgpu_escaped(ctx, &e), ctx, "pmgpu_escaped");
  return e != 0;
}

Vec3 PMMethod::totalExternalForceOrigUnits() {
  const auto& ps = getParticles();
  Vec3 total = Vec3::zero();
  for (const auto& p : ps) {
    const float m = inCodeUnits ? massToOriginalUnits(p.mass, H, DT, G) : p.mass;
    const Vec3 x = inCodeUnits ? positionToOriginalUnits(p.position, H) : p.position;
    total += m * externalField(x);
  }
  return total;
}

std::string PMMethod::run(StateRecorder& stateRecorder, const int simLength, bool collectDiagnostics,
                          bool recordField) {
  return runLoop(stateRecorder, simLength, collectDiagnostics, recordField, false);
}

// The loop of PMMethod::run (source/pmMethod.cpp:62-135) and P3MMethod::run
// (source/p3mMethod.cpp:59-166): every step drifts, records positions (the
// reference records every step), tests the box, computes the force and
// kicks.  Diagnostics come from the device's FP64 reductions.
std::string PMMethod::runLoop(StateRecorder& rec, int simLength, bool collectDiagnostics, bool recordField,
                              bool p3m) {
  SimInfo simInfo;
  uploadIfModified();
  if (collectDiagnostics) {
    setCodeUnits(false);
    simInfo.setInitialMomentum(getParticles());
  }
  setCodeUnits(true);  // stateToCodeUnits + massToCodeUnits
  initGreensFunction();
  auto force = [&]() {
    check(pmgpu_pm_step(ctx), ctx, "pmgpu_pm_step");
    applyHostField();
    if (p3m) check(pmgpu_p3m_short_range(ctx), ctx, "pmgpu_p3m_short_range");
  };
  force();
  check(pmgpu_leapfrog(ctx, PMGPU_HALF_KICK, 1.0f), ctx, "pmgpu_leapfrog");
  mirrorStale = true;

  const float cm = massToCodeUnits(1.0f, H, DT, G);  // code mass per original mass
  std::vector<float> pos(3 * (size_t)N);
  for (int t = 0; t <= simLength; ++t) {
    std::cout << "progress: " << float(t) / simLength << '\r';
    std::cout.flush();
    if (recordField) rec.recordField(getParticles(), H, DT);  // code-unit accelerations
    check(pmgpu_leapfrog(ctx, PMGPU_DRIFT, 1.0f), ctx, "pmgpu_leapfrog");
    if (collectDiagnostics) check(pmgpu_leapfrog(ctx, PMGPU_INTEGER_STEP, 1.0f), ctx, "pmgpu_leapfrog");
    mirrorStale = true;
    setCodeUnits(false);  // stateToOriginalUnits: positions are recorded in original units
    rec.recordPositions(getParticles());
    if (collectDiagnostics) {
      double d[PMGPU_DIAG_COUNT];
      check(pmgpu_diagnostics(ctx, d), ctx, "pmgpu_diagnostics");
      const double vs = H / DT;  // code velocity -> original
      auto vec = [](const double* q, double s) { return Vec3((float)(q[0] * s), (float)(q[1] * s), (float)(q[2] * s)); };
      const Vec3 expected = simInfo.updateExpectedMomentum(totalExternalForceOrigUnits(), DT);
      rec.recordExpectedMomentum(expected);
      rec.recordTotalMomentum(vec(d + 1, vs / cm));
      rec.recordTotalAngularMomentum(vec(d + 4, H * vs / cm));
      double ext = 0.0;
      for (const auto& q : particles) ext += (double)q.mass * externalPotential(q.position);
      const double fourPiG = 4.0 * std::numbers::pi * G;
      const double pe = 0.5 * (double)H * H * H * d[7] * ((double)H * H) / ((double)DT * DT * DT * DT * fourPiG) + ext;
      const double ke = d[0] * vs * vs / cm;
      rec.recordEnergy((float)pe, (float)ke);
    }
    if (escapedComputationalBox()) {
      std::cout << "Particle moved outside the computational box.\n";
      break;
    }
    setCodeUnits(true);
    force();
    check(pmgpu_leapfrog(ctx, PMGPU_KICK, 1.0f), ctx, "pmgpu_leapfrog");
    mirrorStale = true;
    grid.markDeviceNewer();
  }
  return rec.flush();
}
