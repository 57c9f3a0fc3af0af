// Ensembles of independent simulations stepped by one HIP graph launch.
// There is no counterpart in the reference (it runs one simulation per
// process, source/main.cpp).
//
// Each member is an ordinary context (its own particles, grids, tile sort and
// tuning).  One step of every member -- force (deposit, FFT x G, field,
// interpolation, short range for P3M members) with the kick and the next
// drift fused into the interpolation -- is captured once into a graph, two
// steps per graph (each step's tile sort moves the particles to the other of
// two buffers, so two steps leave them where a replay expects them).  The
// capture forks onto `lanes` streams (members assigned largest-first to the
// least loaded lane), so the graph holds that many independent chains the
// runtime can run concurrently; the members' kernels are small (4k-262k
// particles, 32^3-128^3 grids), so a single launch of the whole ensemble
// replaces thousands of host launches per step.
//
// The autotuner times the knobs of pmgpu_set_tuning on the largest member of
// each class (grid, scheme, log2 of the particle count) and applies the
// fastest to the class; results do not depend on the knobs.
#include "pmgpu_internal.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <map>
#include <tuple>

using namespace pmgpu;

struct pmgpu_batch {
  std::vector<pmgpu_ctx*> sims;
  std::vector<int> use_p3m;
  hipStream_t stream = nullptr;
  std::vector<hipStream_t> lanes;
  std::vector<hipEvent_t> events;
  hipGraphExec_t exec = nullptr;
  std::vector<hipGraphExec_t> sim_exec;
  bool prepared = false;
  std::string err;
};

namespace pmgpu {
namespace {

constexpr int STEPS_PER_LAUNCH = 2;

// One run-loop iteration of a member: force, then with mode = ACC|KICK|DRIFT
// the kick of this step and the drift of the next (source/pmMethod.cpp:79-122).
void member_step(pmgpu_ctx* c, bool p3m, int mode, float kick) {
  tile_sort(c);
  deposit(c);
  fft_forward_green_inverse(c);
  if (p3m) {
    interpolate(c, MODE_ACC, 0.0f, c->iter);
    p3m_short_range(c, mode, kick, c->iter);
  } else {
    interpolate(c, mode, kick, c->iter);
  }
}

// Force evaluations only (no kick, no drift): the state is unchanged apart
// from the particles' storage order.
float time_member(pmgpu_ctx* c, bool p3m, int reps) {
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  member_step(c, p3m, MODE_ACC, 0.0f);
  HIP_CHECK(hipEventRecord(a, c->stream));
  for (int r = 0; r < reps; ++r) member_step(c, p3m, MODE_ACC, 0.0f);
  HIP_CHECK(hipEventRecord(b, c->stream));
  HIP_CHECK(hipEventSynchronize(b));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

void apply_knobs(pmgpu_ctx* c, const int32_t k[6]) {
  c->dep_chunk = k[0];
  c->int_chunk = k[1];
  c->blocks_per_cu = k[2];
  c->interp_kernel = k[4];
  c->sr_chunking = k[5];
  if (k[3] != c->sr_div || c->sr_divy != k[3] || c->sr_divz != k[3] || c->sr_divx_base != k[3] ||
      c->sr_xfine_max != 4)
    sr_set_division(c, k[3]);
}

// Coordinate descent over the knobs (the short-range division only for P3M
// members), starting from the defaults; the interpolation variant 0 (auto)
// is replaced by an explicit 1 or 2.
void autotune_member(pmgpu_ctx* c, bool p3m, int32_t best[6]) {
  const std::vector<int> cand[6] = {{1024, 2048, 4096, 8192}, {512, 1024, 2048, 4096}, {2, 4, 8, 16},
                                    {2, 3, 4, 5}, {1, 2}, {1, 2}};
  best[0] = c->dep_chunk;
  best[1] = c->int_chunk;
  best[2] = c->blocks_per_cu;
  best[3] = c->sr_div;
  best[4] = c->interp_kernel;
  best[5] = c->sr_chunking;
  float best_ms = time_member(c, p3m, 3);
  std::vector<int> order = {4, 2, 0, 1};
  if (p3m) order.insert(order.begin(), {3, 5});
  for (int k : order) {
    for (int v : cand[k]) {
      if (v == best[k]) continue;
      int32_t trial[6] = {best[0], best[1], best[2], best[3], best[4], best[5]};
      trial[k] = v;
      apply_knobs(c, trial);
      const float ms = time_member(c, p3m, 3);
      if (ms < 0.97f * best_ms) {
        best_ms = ms;
        std::copy(trial, trial + 6, best);
      }
    }
  }
  apply_knobs(c, best);
}

}  // namespace
}  // namespace pmgpu

#define BATCH_BEGIN(b) \
  if (!(b)) return PMGPU_ERR_INVALID_ARGUMENT; \
  try {
#define BATCH_END(b)               \
  }                                \
  catch (const Error& e) {         \
    (b)->err = e.msg;              \
    return e.code;                 \
  }                                \
  catch (const std::exception& e) { \
    (b)->err = e.what();           \
    return PMGPU_ERR_DEVICE;       \
  }                                \
  return PMGPU_OK;

extern "C" {

int pmgpu_batch_create(int32_t nsim, const pmgpu_pm_params* params, const pmgpu_p3m_params* p3m,
                       const int32_t* use_p3m, int32_t lanes, pmgpu_batch** out) {
  if (!out || nsim <= 0 || !params || lanes < 1 || lanes > 16) return PMGPU_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  pmgpu_batch* b = new pmgpu_batch();
  for (int i = 0; i < nsim; ++i) {
    pmgpu_ctx* c = nullptr;
    int rc = pmgpu_create(&params[i], &c);
    if (rc == PMGPU_OK && use_p3m && use_p3m[i]) {
      if (!p3m) rc = PMGPU_ERR_INVALID_ARGUMENT;
      else rc = pmgpu_p3m_configure(c, &p3m[i]);
    }
    if (rc != PMGPU_OK) {
      if (c) pmgpu_destroy(c);
      pmgpu_batch_destroy(b);
      return rc;
    }
    b->sims.push_back(c);
    b->use_p3m.push_back(use_p3m ? use_p3m[i] : 0);
  }
  try {
    HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (int l = 0; l < lanes; ++l) {
      hipStream_t s;
      HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      b->lanes.push_back(s);
    }
    for (int l = 0; l <= lanes; ++l) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->events.push_back(e);
    }
  } catch (const Error& e) {
    pmgpu_batch_destroy(b);
    return e.code;
  }
  *out = b;
  return PMGPU_OK;
}

void pmgpu_batch_destroy(pmgpu_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  for (auto e : b->sim_exec)
    if (e) (void)hipGraphExecDestroy(e);
  for (auto c : b->sims) pmgpu_destroy(c);
  for (auto e : b->events) (void)hipEventDestroy(e);
  for (auto s : b->lanes) (void)hipStreamDestroy(s);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* pmgpu_batch_last_error(const pmgpu_batch* b) { return b ? b->err.c_str() : "null batch"; }

int pmgpu_batch_size(const pmgpu_batch* b) { return b ? (int)b->sims.size() : -1; }

int pmgpu_batch_set_particles(pmgpu_batch* b, int32_t sim, int64_t n, const float* pos, const float* vel,
                              const float* mass) {
  BATCH_BEGIN(b)
  if (sim < 0 || sim >= (int)b->sims.size()) fail(PMGPU_ERR_INVALID_ARGUMENT, "member index out of range");
  if (b->prepared) fail(PMGPU_ERR_NOT_READY, "the ensemble is already prepared");
  const int rc = pmgpu_set_particles(b->sims[sim], n, pos, vel, mass);
  if (rc != PMGPU_OK) fail(rc, b->sims[sim]->err);
  BATCH_END(b)
}

// Green's functions, the initial force with the half kick and the first drift
// (the run loop's set-up, source/pmMethod.cpp:71-78), optional autotuning,
// and the capture of the ensemble step graph and of one graph per member.
int pmgpu_batch_prepare(pmgpu_batch* b, int32_t autotune) {
  BATCH_BEGIN(b)
  if (b->prepared) fail(PMGPU_ERR_NOT_READY, "the ensemble is already prepared");
  const int n = (int)b->sims.size();
  for (int i = 0; i < n; ++i) {
    pmgpu_ctx* c = b->sims[i];
    if (c->n <= 0) fail(PMGPU_ERR_NOT_READY, "a member has no particles");
    int rc = pmgpu_init_greens(c);
    if (rc != PMGPU_OK) fail(rc, c->err);
  }
  if (autotune) {
    // classes: (nx, ny, nz, scheme, p3m, floor(log2 n)); tune the largest member
    std::map<std::tuple<int, int, int, int, int, int>, int> rep;
    for (int i = 0; i < n; ++i) {
      pmgpu_ctx* c = b->sims[i];
      auto key = std::make_tuple(c->nx, c->ny, c->nz, c->prm.interpolation, b->use_p3m[i],
                                 (int)std::floor(std::log2((double)c->n)));
      auto it = rep.find(key);
      if (it == rep.end() || b->sims[it->second]->n < c->n) rep[key] = i;
    }
    std::map<std::tuple<int, int, int, int, int, int>, std::array<int32_t, 6>> best;
    for (auto& kv : rep) {
      pmgpu_ctx* c = b->sims[kv.second];
      c->iter = INT_MIN;
      int32_t k[6];
      autotune_member(c, b->use_p3m[kv.second] != 0, k);
      best[kv.first] = {k[0], k[1], k[2], k[3], k[4], k[5]};
    }
    for (int i = 0; i < n; ++i) {
      pmgpu_ctx* c = b->sims[i];
      auto key = std::make_tuple(c->nx, c->ny, c->nz, c->prm.interpolation, b->use_p3m[i],
                                 (int)std::floor(std::log2((double)c->n)));
      apply_knobs(c, best[key].data());
    }
  }
  // initial force, half kick and first drift, on each member's own stream
  for (int i = 0; i < n; ++i) {
    pmgpu_ctx* c = b->sims[i];
    c->iter = INT_MIN;  // untagged: escapes are recorded, nothing is gated
    member_step(c, b->use_p3m[i] != 0, MODE_ACC | MODE_KICK | MODE_DRIFT, 0.5f);
  }
  for (int i = 0; i < n; ++i) HIP_CHECK(hipStreamSynchronize(b->sims[i]->stream));

  // the ensemble graph: fork onto the lanes, largest members first onto the
  // least loaded lane, join
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  auto cost = [&](int i) { return (double)b->sims[i]->n * (b->use_p3m[i] ? 8.0 : 1.0) + (double)b->sims[i]->M; };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
  const int L = (int)b->lanes.size();
  std::vector<double> load(L, 0.0);
  std::vector<int> lane_of(n);
  for (int i : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lane_of[i] = l;
    load[l] += cost(i);
  }
  std::vector<hipStream_t> own(n);
  hipGraph_t graph = nullptr;
  HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
  try {
    HIP_CHECK(hipEventRecord(b->events[L], b->stream));
    for (int l = 0; l < L; ++l) HIP_CHECK(hipStreamWaitEvent(b->lanes[l], b->events[L], 0));
    for (int i : order) {
      pmgpu_ctx* c = b->sims[i];
      own[i] = c->stream;
      c->stream = b->lanes[lane_of[i]];
      for (int k = 0; k < STEPS_PER_LAUNCH; ++k)
        member_step(c, b->use_p3m[i] != 0, MODE_ACC | MODE_KICK | MODE_DRIFT, 1.0f);
      c->stream = own[i];
    }
    for (int l = 0; l < L; ++l) {
      HIP_CHECK(hipEventRecord(b->events[l], b->lanes[l]));
      HIP_CHECK(hipStreamWaitEvent(b->stream, b->events[l], 0));
    }
  } catch (...) {
    for (int i = 0; i < n; ++i)
      if (own[i]) b->sims[i]->stream = own[i];
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(b->stream, &graph));
  const hipError_t ie = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_CHECK(ie);
  // one graph per member, for its step latency
  b->sim_exec.assign(n, nullptr);
  for (int i = 0; i < n; ++i) {
    pmgpu_ctx* c = b->sims[i];
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    try {
      for (int k = 0; k < STEPS_PER_LAUNCH; ++k)
        member_step(c, b->use_p3m[i] != 0, MODE_ACC | MODE_KICK | MODE_DRIFT, 1.0f);
    } catch (...) {
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(c->stream, &g));
    const hipError_t e2 = hipGraphInstantiate(&b->sim_exec[i], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_CHECK(e2);
  }
  b->prepared = true;
  BATCH_END(b)
}

// `steps` run-loop iterations of every member (an even number: one graph
// launch per two steps).
int pmgpu_batch_step(pmgpu_batch* b, int32_t steps) {
  BATCH_BEGIN(b)
  if (!b->prepared) fail(PMGPU_ERR_NOT_READY, "call pmgpu_batch_prepare first");
  if (steps < 0 || steps % STEPS_PER_LAUNCH) fail(PMGPU_ERR_INVALID_ARGUMENT, "steps must be even and >= 0");
  for (int s = 0; s < steps / STEPS_PER_LAUNCH; ++s) HIP_CHECK(hipGraphLaunch(b->exec, b->stream));
  BATCH_END(b)
}

int pmgpu_batch_synchronize(pmgpu_batch* b) {
  BATCH_BEGIN(b)
  HIP_CHECK(hipStreamSynchronize(b->stream));
  BATCH_END(b)
}

// Wall time of `steps` (even) ensemble steps, in milliseconds, between two
// events on the ensemble stream.
int pmgpu_batch_time(pmgpu_batch* b, int32_t steps, double* ms) {
  BATCH_BEGIN(b)
  if (!b->prepared) fail(PMGPU_ERR_NOT_READY, "call pmgpu_batch_prepare first");
  if (!ms || steps <= 0 || steps % STEPS_PER_LAUNCH) fail(PMGPU_ERR_INVALID_ARGUMENT, "bad arguments");
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  HIP_CHECK(hipEventRecord(e0, b->stream));
  for (int s = 0; s < steps / STEPS_PER_LAUNCH; ++s) HIP_CHECK(hipGraphLaunch(b->exec, b->stream));
  HIP_CHECK(hipEventRecord(e1, b->stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float t = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t;
  BATCH_END(b)
}

// Step latency of each member alone: `reps` launches of its own graph,
// mean milliseconds per step into ms[member].  Advances every member by
// 2 (reps + 1) steps.
int pmgpu_batch_member_latency(pmgpu_batch* b, int32_t reps, double* ms) {
  BATCH_BEGIN(b)
  if (!b->prepared) fail(PMGPU_ERR_NOT_READY, "call pmgpu_batch_prepare first");
  if (!ms || reps <= 0) fail(PMGPU_ERR_INVALID_ARGUMENT, "bad arguments");
  HIP_CHECK(hipStreamSynchronize(b->stream));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  for (size_t i = 0; i < b->sims.size(); ++i) {
    hipStream_t s = b->sims[i]->stream;
    HIP_CHECK(hipGraphLaunch(b->sim_exec[i], s));  // warm
    HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) HIP_CHECK(hipGraphLaunch(b->sim_exec[i], s));
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    float t = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
    ms[i] = (double)t / (reps * STEPS_PER_LAUNCH);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  BATCH_END(b)
}

int pmgpu_batch_get_particles(pmgpu_batch* b, int32_t sim, float* pos, float* vel, float* acc) {
  BATCH_BEGIN(b)
  if (sim < 0 || sim >= (int)b->sims.size()) fail(PMGPU_ERR_INVALID_ARGUMENT, "member index out of range");
  HIP_CHECK(hipStreamSynchronize(b->stream));
  const int rc = pmgpu_get_particles(b->sims[sim], pos, vel, acc);
  if (rc != PMGPU_OK) fail(rc, b->sims[sim]->err);
  BATCH_END(b)
}

int pmgpu_batch_get_tuning(pmgpu_batch* b, int32_t sim, int32_t* knobs) {
  BATCH_BEGIN(b)
  if (sim < 0 || sim >= (int)b->sims.size() || !knobs) fail(PMGPU_ERR_INVALID_ARGUMENT, "bad arguments");
  const int rc = pmgpu_get_tuning(b->sims[sim], knobs);
  if (rc != PMGPU_OK) fail(rc, b->sims[sim]->err);
  BATCH_END(b)
}

}  // extern "C"
