// C ABI of libpmgpu (include/pmgpu.h): context lifetime, particle upload and
// download, the force-step sequence and the leapfrog run loop.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pmgpu_internal.h"

namespace pmgpu {

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(PMGPU_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
  fail(PMGPU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
static void dalloc(T** p, int64_t count) {
  if (*p) {
    HIP_CHECK(hipFree(*p));
    *p = nullptr;
  }
  if (count <= 0) count = 1;
  HIP_CHECK(hipMalloc((void**)p, sizeof(T) * (size_t)count));
  HIP_CHECK(hipMemset(*p, 0, sizeof(T) * (size_t)count));
  // the context stream is non-blocking: it does not wait for null-stream
  // work, so the clear must finish before any kernel can touch the buffer
  HIP_CHECK(hipDeviceSynchronize());
}

template <typename T>
static void dfree(T** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

int stage_begin(pmgpu_ctx* c, const char* name) {
  if (!c->profiling) return -1;
  if (c->profiling == 2 && std::strcmp(name, "sr_pairs") != 0) return -1;
  int s = -1;
  for (size_t i = 0; i < c->stages.size(); ++i)
    if (std::strcmp(c->stages[i].name, name) == 0) s = (int)i;
  if (s < 0) {
    c->stages.push_back(Stage{name});
    s = (int)c->stages.size() - 1;
  }
  pmgpu_ctx::Pending p;
  p.stage = s;
  for (hipEvent_t* e : {&p.a, &p.b}) {
    if (!c->ev_pool.empty()) {
      *e = c->ev_pool.back();
      c->ev_pool.pop_back();
    } else {
      HIP_CHECK(hipEventCreate(e));
    }
  }
  HIP_CHECK(hipEventRecord(p.a, c->stream));
  c->pending.push_back(p);
  return (int)c->pending.size() - 1;
}

void stage_end(pmgpu_ctx* c, int token) {
  if (token < 0) return;
  HIP_CHECK(hipEventRecord(c->pending[token].b, c->stream));
}

static void profile_collect(pmgpu_ctx* c) {
  if (c->pending.empty()) return;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  for (auto& p : c->pending) {
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    c->stages[p.stage].total_ms += ms;
    c->stages[p.stage].launches += 1;
    c->ev_pool.push_back(p.a);
    c->ev_pool.push_back(p.b);
  }
  c->pending.clear();
}


static void validate(const pmgpu_pm_params* p) {
  if (!p) fail(PMGPU_ERR_INVALID_ARGUMENT, "params is NULL");
  for (int i = 0; i < 3; ++i)
    if (p->grid[i] < 8 || p->grid[i] > 1024)
      fail(PMGPU_ERR_UNSUPPORTED, "grid dimensions must be in [8, 1024]");
  if (p->interpolation < 0 || p->interpolation > PMGPU_TSC_CENTRED)
    fail(PMGPU_ERR_INVALID_ARGUMENT, "Unkown interpolation scheme");
  if (p->finite_diff < 0 || p->finite_diff > 1)
    fail(PMGPU_ERR_INVALID_ARGUMENT, "Unknown finite difference type.");
  if (p->greens_function < 0 || p->greens_function > 3)
    fail(PMGPU_ERR_INVALID_ARGUMENT, "not implemented");
  if ((p->greens_function == PMGPU_S1_OPTIMAL || p->greens_function == PMGPU_S2_OPTIMAL) &&
      !(p->particle_diameter > 0.0f))
    fail(PMGPU_ERR_INVALID_ARGUMENT, "optimal Green's function needs particle_diameter > 0");
  if (p->ext_kind != PMGPU_EXT_NONE && p->ext_kind != PMGPU_EXT_SPH_RAD_DECR)
    fail(PMGPU_ERR_INVALID_ARGUMENT, "unknown external field kind");
  if (p->ext_kind == PMGPU_EXT_SPH_RAD_DECR && !(p->H > 0.0f && p->DT > 0.0f))
    fail(PMGPU_ERR_INVALID_ARGUMENT, "external field needs H > 0 and DT > 0");
}

static void twiddles(float2** dst, int L, double denom_len) {
  std::vector<float2> t(L);
  for (int m = 0; m < L; ++m) {
    double a = -2.0 * M_PI * (double)m / denom_len;
    t[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  dalloc(dst, L);
  HIP_CHECK(hipMemcpy(*dst, t.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
}

static void ensure_scan_scratch(pmgpu_ctx* c, int64_t n) {
  size_t need = scan_scratch_elems(n);
  if (need > c->scan_tmp_elems) {
    dalloc(&c->scan_tmp, (int64_t)need);
    c->scan_tmp_elems = need;
  }
}

static void alloc_particle_arrays(pmgpu_ctx* c, int64_t n) {
  for (int b = 0; b < 2; ++b) {
    dalloc(&c->pm[b], n);
    dalloc(&c->vel[b], n);
  }
  dalloc(&c->acc, n);
  dalloc(&c->key, n);
  dalloc(&c->rank, n);
  dfree(&c->vint);
  dfree(&c->ext_stage);
  if (c->acc_sr) dalloc(&c->acc_sr, n);  // recording stays on for a new particle set
  c->fx_slots = 8 * n / DEPOSIT_CHUNK_MIN + 2;  // enough for any tunable chunk size
  dalloc(&c->fx_accum, c->fx_slots * TILE_CELLS);
  dalloc(&c->slot_tile, c->fx_slots);
  // work-item owner tables, sized for the smallest tunable work items
  dalloc(&c->dep_owner, (int64_t)c->ntiles + 8 * n / DEPOSIT_CHUNK_MIN + 2);
  dalloc(&c->int_owner, (int64_t)c->ntiles + n / INTERP_CHUNK_MIN + 2);
  if (c->p3m_ready) {
    dalloc(&c->sr_owner, (int64_t)c->ncell_sr + n / 64 + 2);
    dalloc(&c->sr_key, n);
    dalloc(&c->sr_rank, n);
    dalloc(&c->sr_pm, n + SR_JGROUP);
    dalloc(&c->sr_idx, n);
  }
}

// Cells the short-range grid may use: about 128 per particle (every step
// clears and scans them, which must stay small beside the pair work), at
// most SR_CELLS_MAX; 16M before the particles are known.
static double sr_cell_budget(const pmgpu_ctx* c) {
  if (c->n <= 0) return 16.0 * 1024 * 1024;
  return std::min((double)SR_CELLS_MAX, std::max(65536.0, 128.0 * (double)c->n));
}

// The short-range cell grid over the particle box: y sides >= cutoff /
// sr_divy, z sides >= cutoff / sr_divz, x sides >= cutoff / (sr_divx_base
// sr_xfine).  Finer x cells trim each neighbour row closer to the cutoff
// sphere (a row is contiguous in memory, so its length costs nothing but
// candidates).  sr_xfine is the largest of sr_xfine_max (4 or 8), its
// halvings and 1 that keeps the grid within the cell budget.
void sr_cells_setup(pmgpu_ctx* c) {
  const double rc = (double)c->p3m.cutoff * (1.0 + 1e-5);
  // cell counts in FP64 first, so that a tiny cutoff cannot overflow an int
  double d[3];
  d[1] = std::max(1.0, std::floor((double)c->p3m.box[1] / (rc / c->sr_divy)));
  d[2] = std::max(1.0, std::floor((double)c->p3m.box[2] / (rc / c->sr_divz)));
  for (c->sr_xfine = c->sr_xfine_max;; c->sr_xfine /= 2) {
    d[0] = std::max(1.0, std::floor((double)c->p3m.box[0] / (rc / (c->sr_divx_base * c->sr_xfine))));
    if (d[0] * d[1] * d[2] <= sr_cell_budget(c) || c->sr_xfine == 1) break;
  }
  if (d[0] * d[1] * d[2] > (double)SR_CELLS_MAX)
    fail(PMGPU_ERR_UNSUPPORTED, "cutoff too small for the box (too many short-range cells)");
  const int m[3] = {(int)d[0], (int)d[1], (int)d[2]};
  c->mcx = m[0];
  c->mcy = m[1];
  c->mcz = m[2];
  c->hcx = c->p3m.box[0] / m[0];
  c->hcy = c->p3m.box[1] / m[1];
  c->hcz = c->p3m.box[2] / m[2];
  c->ncell_sr = m[0] * m[1] * m[2];
  dalloc(&c->sr_count, c->ncell_sr + 1);
  dalloc(&c->sr_start, c->ncell_sr + 1);
  dalloc(&c->sr_nchunk, c->ncell_sr + 1);
  dalloc(&c->sr_chunk_start, c->ncell_sr + 1);
  ensure_scan_scratch(c, c->ncell_sr);
  if (c->n > 0) dalloc(&c->sr_owner, (int64_t)c->ncell_sr + c->n / 64 + 2);
}

// The division p3m_configure picks when none is set: the finest whose
// isotropic cell grid stays modest (finer cells visit fewer candidate pairs;
// the pair set is the same for every division).
static void sr_auto_division(pmgpu_ctx* c) {
  auto cells_at = [&](int d) {
    double cells = 1.0;
    for (int i = 0; i < 3; ++i)
      cells *= std::floor((double)c->p3m.box[i] / ((double)c->p3m.cutoff / d * (1.0 + 1e-5)));
    return cells;
  };
  // cells of side >= cutoff (the reference's chaining mesh) only when even
  // half that side gives too many cells
  int div = cells_at(2) <= (double)SR_CELLS_MAX ? 2 : SR_CELL_DIV;
  const double budget = std::min(16.0 * 1024 * 1024, sr_cell_budget(c));
  for (int d = SR_CELL_DIV_MAX; d > 2; --d) {
    if (cells_at(d) <= budget) {
      div = d;
      break;
    }
  }
  c->sr_div = div;
}

void sr_set_division(pmgpu_ctx* c, int div) {
  c->sr_div = c->sr_divy = c->sr_divz = c->sr_divx_base = div;
  c->sr_xfine_max = 4;
  if (c->p3m_ready) {
    HIP_CHECK(hipStreamSynchronize(c->stream));
    sr_cells_setup(c);
  }
}

// The common mass and the particles off it (original ids ascending), if
// there are at most SR_ODD_MAX of them: a majority vote over the mass bits
// (Boyer-Moore) and a count.  A mass held by all but SR_ODD_MAX particles
// is the majority whenever n > 2 SR_ODD_MAX; below that one may be missed,
// which leaves the pair loop's per-pair mass multiply on.
static void set_mass_classes(pmgpu_ctx* c, const float* mass, int64_t n) {
  c->sr_nodd = -1;
  if (n <= 0) return;
  auto bits = [&](int64_t i) {
    uint32_t b;
    std::memcpy(&b, &mass[i], sizeof b);
    return b;
  };
  uint32_t cand = 0;
  int64_t votes = 0;
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t b = bits(i);
    if (votes == 0) {
      cand = b;
      votes = 1;
    } else {
      votes += b == cand ? 1 : -1;
    }
  }
  std::vector<int32_t> odd;
  for (int64_t i = 0; i < n; ++i) {
    if (bits(i) == cand) continue;
    if ((int64_t)odd.size() == SR_ODD_MAX) return;
    odd.push_back((int32_t)i);
  }
  if (!c->sr_odd_ids) {
    dalloc(&c->sr_odd_ids, SR_ODD_MAX);
    dalloc(&c->sr_odd_pos, SR_ODD_MAX);
  }
  if (!odd.empty())
    HIP_CHECK(hipMemcpy(c->sr_odd_ids, odd.data(), sizeof(int32_t) * odd.size(), hipMemcpyHostToDevice));
  std::memcpy(&c->sr_m0, &cand, sizeof cand);
  c->sr_nodd = (int)odd.size();
}

static void set_shift(pmgpu_ctx* c) {
  // every cell sum of the deposit is bounded by 1.5 * sum|m| (the largest
  // weight product of any scheme is 1.125^3 < 1.5); keep it below 2^62
  if (c->mass_abs_sum > 0.0) {
    int s = (int)std::floor(62.0 - std::log2(1.5 * c->mass_abs_sum)) - 1;
    c->fx_shift = std::max(-120, std::min(120, s));
  } else {
    c->fx_shift = 0;
  }
}

static void reset_escape(pmgpu_ctx* c) {
  int32_t v = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(c->escape, &v, sizeof(v), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

// Retires the captured run-loop graphs (their kernel arguments may change).
static void graphs_stale(pmgpu_ctx* c) { ++c->graph_gen; }

static void graphs_free(pmgpu_ctx* c) {
  for (auto& row : c->run_exec)
    for (auto& e : row)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  c->run_exec_gen[0] = c->run_exec_gen[1] = -1;
}

static void require_particles(pmgpu_ctx* c) {
  if (c->n <= 0) fail(PMGPU_ERR_NOT_READY, "no particles: call pmgpu_set_particles first");
}

// The force step of PMMethod::pmMethodStep on the current state.
static void pm_force(pmgpu_ctx* c, int interp_mode, float kick) {
  tile_sort(c);
  deposit(c);
  fft_forward_green_inverse(c);
  interpolate(c, interp_mode, kick, c->iter);
}

// The P3M force of one run-loop iteration with the short-range cell sort on
// the second stream, beside the deposit, transforms and interpolation: the
// sort needs only the particles' order after the tile sort, and the pair
// pass, which adds to the interpolated accelerations, waits for both.
static void p3m_force_overlapped(pmgpu_ctx* c, int mode, float kick) {
  if (!c->side) {
    HIP_CHECK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  const size_t need = std::max(scan_scratch_elems(c->ncell_sr), scan_scratch_elems(c->ncell_sr + 1));
  if (need > c->scan_tmp_side_elems) {
    dalloc(&c->scan_tmp_side, (int64_t)need);
    c->scan_tmp_side_elems = need;
  }
  tile_sort(c);
  HIP_CHECK(hipEventRecord(c->ev_fork, c->stream));
  HIP_CHECK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
  {
    // the sort's kernels go to the side stream with the side scratch
    struct Swap {
      pmgpu_ctx* c;
      explicit Swap(pmgpu_ctx* c_) : c(c_) { flip(); }
      ~Swap() { flip(); }
      void flip() {
        std::swap(c->stream, c->side);
        std::swap(c->scan_tmp, c->scan_tmp_side);
        std::swap(c->scan_tmp_elems, c->scan_tmp_side_elems);
      }
    } swap(c);
    p3m_sort(c, c->iter);
    HIP_CHECK(hipEventRecord(c->ev_join, c->stream));
  }
  deposit(c);
  fft_forward_green_inverse(c);
  interpolate(c, MODE_ACC, 0.0f, c->iter);
  HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  p3m_pairs(c, MODE_ACC | mode, kick, c->iter);
}

}  // namespace pmgpu

using namespace pmgpu;

static thread_local std::string g_create_error;

#define PMGPU_API_BEGIN(ctx) \
  if (!(ctx)) return PMGPU_ERR_INVALID_ARGUMENT; \
  try {
#define PMGPU_API_END(ctx)        \
  }                               \
  catch (const Error& e) {        \
    (ctx)->err = e.msg;           \
    return e.code;                \
  }                               \
  catch (const std::exception& e) { \
    (ctx)->err = e.what();        \
    return PMGPU_ERR_DEVICE;      \
  }                               \
  return PMGPU_OK;

// escape[1]: the iteration of kernels replayed from a graph (ITER_DEVICE)
__global__ void k_iter_set(int32_t* __restrict__ escape, int v) { escape[1] = v; }
__global__ void k_iter_advance(int32_t* __restrict__ escape) { escape[1] += 1; }

// The graph of two run-loop iterations that both drift into the next, each
// followed by the advance of escape[1], for the particle buffers as they are
// now (c->cur); built for both buffer parities at once (a capture enqueues
// nothing, and two iterations leave the parity as it was) and kept until
// graph_gen moves.  Returns null when the capture or instantiation fails:
// the caller then enqueues the iterations directly.
template <typename Force>
static hipGraphExec_t run_graph(pmgpu_ctx* c, bool p3m, Force& force) {
  const int f = p3m ? 1 : 0;
  if (c->run_exec_gen[f] == c->graph_gen && c->run_exec[f][c->cur]) return c->run_exec[f][c->cur];
  for (auto& e : c->run_exec[f])
    if (e) {
      (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
  c->run_exec_gen[f] = c->graph_gen;
  const int cur0 = c->cur, iter0 = c->iter;
  for (int parity = 0; parity < 2; ++parity) {
    c->cur = parity;
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess) break;
    bool ok = true;
    try {
      c->iter = ITER_DEVICE;
      for (int k = 0; k < 2; ++k) {
        force(MODE_KICK | MODE_DRIFT, 1.0f);
        hipLaunchKernelGGL(k_iter_advance, dim3(1), dim3(1), 0, c->stream, c->escape);
      }
    } catch (...) {
      ok = false;
    }
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (ok && ec == hipSuccess && graph)
      ok = hipGraphInstantiate(&c->run_exec[f][parity], graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (!ok) c->run_exec[f][parity] = nullptr;
  }
  c->cur = cur0;
  c->iter = iter0;
  return c->run_exec[f][c->cur];
}

extern "C" {

int pmgpu_create(const pmgpu_pm_params* params, pmgpu_ctx** out) {
  if (!out) return PMGPU_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  pmgpu_ctx* c = nullptr;
  try {
    validate(params);
    c = new pmgpu_ctx();
    c->prm = *params;
    if (c->prm.interpolation == PMGPU_TSC_CENTRED) {
      c->prm.interpolation = PMGPU_TSC;
      c->tsc_centred = 1;
    }
    c->nx = params->grid[0];
    c->ny = params->grid[1];
    c->nz = params->grid[2];
    c->nxh = c->nx / 2 + 1;
    c->M = (int64_t)c->nx * c->ny * c->nz;
    // rows of the half spectrum start on 128-byte boundaries, so the y and z
    // passes' column tiles (16 or 32 complex values) are whole cache lines
    c->sp = (c->nxh + 15) / 16 * 16;
    c->K = (int64_t)c->sp * c->ny * c->nz;
    c->general = general_grid(c->nx, c->ny, c->nz);
    // the tiled deposit and interpolation need sides that the tiles divide
    c->ntx = c->general ? 0 : c->nx / TILE_X;
    c->nty = c->general ? 0 : c->ny / TILE_Y;
    c->ntz = c->general ? 0 : c->nz / TILE_Z;
    c->ntiles = c->ntx * c->nty * c->ntz;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    dalloc(&c->rho, c->M);
    dalloc(&c->phi, c->M);
    dalloc(&c->spec, c->K);
    dalloc(&c->greens, c->K);
    twiddles(&c->tw_x, c->nx / 2, c->nx / 2);
    twiddles(&c->tw_y, c->ny, c->ny);
    twiddles(&c->tw_z, c->nz, c->nz);
    twiddles(&c->tw_r2c, c->nx / 2 + 1, c->nx);
    if (c->general) {
      twiddles(&c->tw_gx, c->nx, c->nx);
      dalloc(&c->fx_grid, c->M);
    }
    const int64_t nkeys = (int64_t)c->ntiles * NCLASS;
    dalloc(&c->cls_count, nkeys + 1);
    dalloc(&c->cls_start, nkeys + 1);
    dalloc(&c->tile_nchunk, c->ntiles + 1);
    dalloc(&c->tile_chunk_start, c->ntiles + 1);
    dalloc(&c->tile_slot, c->ntiles + 1);
    dalloc(&c->int_nchunk, c->ntiles + 1);
    dalloc(&c->int_chunk_start, c->ntiles + 1);
    dalloc(&c->counters, 32);  // [16, 24): per-XCD work counters of the short-range pass
    dalloc(&c->escape, 2);
    ensure_scan_scratch(c, nkeys);
    std::vector<int4> tab = deposit_range_table(c->prm.interpolation);
    c->dep_table_len = (int)tab.size();
    dalloc(&c->dep_table, std::max<int64_t>(1, (int64_t)tab.size()));
    if (!tab.empty())
      HIP_CHECK(hipMemcpy(c->dep_table, tab.data(), sizeof(int4) * tab.size(), hipMemcpyHostToDevice));
    reset_escape(c);
  } catch (const Error& e) {
    g_create_error = e.msg;
    if (c) pmgpu_destroy(c);
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    if (c) pmgpu_destroy(c);
    return PMGPU_ERR_DEVICE;
  }
  *out = c;
  return PMGPU_OK;
}

void pmgpu_destroy(pmgpu_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  graphs_free(c);
  for (int b = 0; b < 2; ++b) {
    dfree(&c->pm[b]);
    dfree(&c->vel[b]);
  }
  dfree(&c->acc); dfree(&c->vint); dfree(&c->ext_stage); dfree(&c->acc_sr); dfree(&c->key); dfree(&c->rank);
  dfree(&c->cls_count); dfree(&c->cls_start); dfree(&c->tile_nchunk); dfree(&c->tile_chunk_start);
  dfree(&c->tile_slot); dfree(&c->slot_tile); dfree(&c->int_nchunk); dfree(&c->int_chunk_start);
  dfree(&c->dep_table); dfree(&c->counters); dfree(&c->fx_accum); dfree(&c->scan_tmp);
  dfree(&c->dep_owner); dfree(&c->int_owner); dfree(&c->sr_owner);
  dfree(&c->rho); dfree(&c->spec); dfree(&c->phi); dfree(&c->greens); dfree(&c->greens_z); dfree(&c->field);
  dfree(&c->efield);
  dfree(&c->tw_x); dfree(&c->tw_y); dfree(&c->tw_z); dfree(&c->tw_r2c); dfree(&c->tw_gx); dfree(&c->fx_grid);
  dfree(&c->sr_table); dfree(&c->sr_count); dfree(&c->sr_start); dfree(&c->sr_key);
  dfree(&c->sr_rank); dfree(&c->sr_pm); dfree(&c->sr_idx); dfree(&c->sr_nchunk);
  dfree(&c->sr_odd_ids); dfree(&c->sr_odd_pos);
  dfree(&c->sr_chunk_start); dfree(&c->pair_count); dfree(&c->escape); dfree(&c->diag_scratch);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  dfree(&c->scan_tmp_side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* pmgpu_last_error(const pmgpu_ctx* c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

int pmgpu_synchronize(pmgpu_ctx* c) {
  PMGPU_API_BEGIN(c)
  HIP_CHECK(hipStreamSynchronize(c->stream));
  PMGPU_API_END(c)
}

int64_t pmgpu_num_particles(const pmgpu_ctx* c) { return c ? c->n : -1; }

int pmgpu_set_particles(pmgpu_ctx* c, int64_t n, const float* pos, const float* vel,
                        const float* mass) {
  PMGPU_API_BEGIN(c)
  graphs_stale(c);
  if (n < 0 || (n > 0 && (!pos || !mass)))
    fail(PMGPU_ERR_INVALID_ARGUMENT, "pos and mass are required");
  if (n > (int64_t)INT_MAX / 2) fail(PMGPU_ERR_UNSUPPORTED, "too many particles");
  HIP_CHECK(hipStreamSynchronize(c->stream));
  const int64_t old_n = c->n;
  alloc_particle_arrays(c, n);
  c->n = n;
  std::vector<float4> hp(n), hv(n);
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    hp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass[i]);
    float vx = vel ? vel[3 * i] : 0.f, vy = vel ? vel[3 * i + 1] : 0.f, vz = vel ? vel[3 * i + 2] : 0.f;
    int32_t id = (int32_t)i;
    float w;
    std::memcpy(&w, &id, sizeof(w));
    hv[i] = make_float4(vx, vy, vz, w);
    s += std::fabs((double)mass[i]);
  }
  c->cur = 0;
  if (n > 0) {
    HIP_CHECK(hipMemcpy(c->pm[0], hp.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->vel[0], hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
  }
  c->mass_abs_sum = s;
  set_mass_classes(c, mass, n);
  set_shift(c);
  c->sorted_valid = false;
  c->run_iter = 0;
  reset_escape(c);
  // the accurate pair table's layout depends on the masses' scale
  if (c->p3m_ready) p3m_setup(c);
  // the short-range cells depend on the particles: their budget on the
  // count, and (unless the division is set) their shape on the positions
  if (c->p3m_ready && n != old_n) {
    if (!c->sr_div_set) {
      sr_auto_division(c);
      c->sr_divy = c->sr_divz = c->sr_divx_base = c->sr_div;
      c->sr_xfine_max = 4;
    }
    sr_cells_setup(c);
    if (!c->sr_div_set && n > 0) sr_choose_cells(c);
  }
  PMGPU_API_END(c)
}

int pmgpu_get_particles(pmgpu_ctx* c, float* pos, float* vel, float* acc) {
  PMGPU_API_BEGIN(c)
  const int64_t n = c->n;
  if (n == 0) return PMGPU_OK;
  std::vector<float4> hp(n), hv(n), ha;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  HIP_CHECK(hipMemcpy(hv.data(), c->vel[c->cur], sizeof(float4) * n, hipMemcpyDeviceToHost));
  if (pos) HIP_CHECK(hipMemcpy(hp.data(), c->pm[c->cur], sizeof(float4) * n, hipMemcpyDeviceToHost));
  if (acc) {
    ha.resize(n);
    HIP_CHECK(hipMemcpy(ha.data(), c->acc, sizeof(float4) * n, hipMemcpyDeviceToHost));
  }
  for (int64_t s = 0; s < n; ++s) {
    int32_t id;
    std::memcpy(&id, &hv[s].w, sizeof(id));
    if (id < 0 || id >= n) fail(PMGPU_ERR_DEVICE, "corrupted particle index");
    if (pos) {
      pos[3 * id] = hp[s].x;
      pos[3 * id + 1] = hp[s].y;
      pos[3 * id + 2] = hp[s].z;
    }
    if (vel) {
      vel[3 * id] = hv[s].x;
      vel[3 * id + 1] = hv[s].y;
      vel[3 * id + 2] = hv[s].z;
    }
    if (acc) {
      acc[3 * id] = ha[s].x;
      acc[3 * id + 1] = ha[s].y;
      acc[3 * id + 2] = ha[s].z;
    }
  }
  PMGPU_API_END(c)
}

int pmgpu_update_particles(pmgpu_ctx* c, const float* pos, const float* vel, const float* acc) {
  PMGPU_API_BEGIN(c)
  graphs_stale(c);
  const int64_t n = c->n;
  if (n == 0) return PMGPU_OK;
  std::vector<float4> hp(n), hv(n), ha(n);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  HIP_CHECK(hipMemcpy(hv.data(), c->vel[c->cur], sizeof(float4) * n, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(hp.data(), c->pm[c->cur], sizeof(float4) * n, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(ha.data(), c->acc, sizeof(float4) * n, hipMemcpyDeviceToHost));
  for (int64_t s = 0; s < n; ++s) {
    int32_t id;
    std::memcpy(&id, &hv[s].w, sizeof(id));
    if (pos) {
      hp[s].x = pos[3 * id];
      hp[s].y = pos[3 * id + 1];
      hp[s].z = pos[3 * id + 2];
    }
    if (vel) {
      hv[s].x = vel[3 * id];
      hv[s].y = vel[3 * id + 1];
      hv[s].z = vel[3 * id + 2];
    }
    if (acc) ha[s] = make_float4(acc[3 * id], acc[3 * id + 1], acc[3 * id + 2], 0.f);
  }
  HIP_CHECK(hipMemcpy(c->pm[c->cur], hp.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->vel[c->cur], hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->acc, ha.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
  if (pos) c->sorted_valid = false;
  c->run_iter = 0;
  reset_escape(c);
  PMGPU_API_END(c)
}

int pmgpu_init_greens(pmgpu_ctx* c) {
  PMGPU_API_BEGIN(c)
  graphs_stale(c);
  greens_compute(c, nullptr);
  if (!c->mlp_weights.empty()) mlp_apply_to_greens(c);  // G_sym * (1 + y(k))
  if (!c->general) {
    if (!c->greens_z) dalloc(&c->greens_z, c->K);  // (Nx/2 + 1) Ny Nz <= K values
    greens_block_for_z(c);
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->greens_ready = true;
  PMGPU_API_END(c)
}

int pmgpu_pm_step(pmgpu_ctx* c) {
  PMGPU_API_BEGIN(c)
  require_particles(c);
  if (!c->greens_ready) fail(PMGPU_ERR_NOT_READY, "call pmgpu_init_greens first");
  c->iter = INT_MIN;
  pm_force(c, MODE_ACC, 0.0f);
  PMGPU_API_END(c)
}

int pmgpu_p3m_configure(pmgpu_ctx* c, const pmgpu_p3m_params* p) {
  PMGPU_API_BEGIN(c)
  graphs_stale(c);
  if (!p) fail(PMGPU_ERR_INVALID_ARGUMENT, "params is NULL");
  if (!(p->cutoff > 0.0f)) fail(PMGPU_ERR_INVALID_ARGUMENT, "cutoff must be > 0");
  if (p->cloud_shape != PMGPU_S1 && p->cloud_shape != PMGPU_S2)
    fail(PMGPU_ERR_INVALID_ARGUMENT, "not implemented");
  if (p->use_table < 0 || p->use_table > 2) fail(PMGPU_ERR_INVALID_ARGUMENT, "use_table must be 0, 1 or 2");
  if (!(p->particle_diameter > 0.0f)) fail(PMGPU_ERR_INVALID_ARGUMENT, "particle_diameter must be > 0");
  for (int i = 0; i < 3; ++i)
    if (!(p->box[i] > 0.0f)) fail(PMGPU_ERR_INVALID_ARGUMENT, "box must be > 0");
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->p3m = *p;
  if (!c->sr_div_set) sr_auto_division(c);
  c->sr_divy = c->sr_divz = c->sr_divx_base = c->sr_div;
  c->sr_xfine_max = 4;
  sr_cells_setup(c);
  dalloc(&c->sr_table, SR_TABLE_EXT);  // the largest of the tables
  dalloc(&c->pair_count, 3);
  c->p3m_ready = true;
  if (c->n > 0) {
    dalloc(&c->sr_key, c->n);
    dalloc(&c->sr_rank, c->n);
    dalloc(&c->sr_pm, c->n + SR_JGROUP);
    dalloc(&c->sr_idx, c->n);
  }
  p3m_setup(c);
  if (!c->sr_div_set && c->n > 0) sr_choose_cells(c);
  PMGPU_API_END(c)
}

int pmgpu_p3m_short_range(pmgpu_ctx* c) {
  PMGPU_API_BEGIN(c)
  require_particles(c);
  if (!c->p3m_ready) fail(PMGPU_ERR_NOT_READY, "call pmgpu_p3m_configure first");
  c->iter = INT_MIN;
  p3m_short_range(c, MODE_ACC, 0.0f, c->iter);
  PMGPU_API_END(c)
}

int pmgpu_p3m_count_pairs(pmgpu_ctx* c, int64_t* pairs, int64_t* tested) {
  PMGPU_API_BEGIN(c)
  require_particles(c);
  if (!c->p3m_ready) fail(PMGPU_ERR_NOT_READY, "call pmgpu_p3m_configure first");
  if (!pairs) fail(PMGPU_ERR_INVALID_ARGUMENT, "pairs is NULL");
  p3m_count_pairs(c, pairs, tested, nullptr);
  PMGPU_API_END(c)
}

int pmgpu_p3m_pair_stats(pmgpu_ctx* c, int64_t* stats) {
  PMGPU_API_BEGIN(c)
  require_particles(c);
  if (!c->p3m_ready) fail(PMGPU_ERR_NOT_READY, "call pmgpu_p3m_configure first");
  if (!stats) fail(PMGPU_ERR_INVALID_ARGUMENT, "stats is NULL");
  p3m_count_pairs(c, &stats[0], &stats[1], &stats[2]);
  PMGPU_API_END(c)
}

int pmgpu_p3m_cell_geometry(pmgpu_ctx* c, int32_t* geo) {
  PMGPU_API_BEGIN(c)
  if (!c->p3m_ready) fail(PMGPU_ERR_NOT_READY, "call pmgpu_p3m_configure first");
  if (!geo) fail(PMGPU_ERR_INVALID_ARGUMENT, "geo is NULL");
  geo[0] = c->sr_divx_base * c->sr_xfine;
  geo[1] = c->sr_divy;
  geo[2] = c->sr_divz;
  geo[3] = (int32_t)c->ncell_sr;
  PMGPU_API_END(c)
}

int pmgpu_leapfrog(pmgpu_ctx* c, int32_t op, float dt) {
  PMGPU_API_BEGIN(c)
  require_particles(c);
  if (op < PMGPU_HALF_KICK || op > PMGPU_DRIFT) fail(PMGPU_ERR_INVALID_ARGUMENT, "unknown leapfrog op");
  if (op == PMGPU_INTEGER_STEP && !c->vint) dalloc(&c->vint, c->n);
  leapfrog_op(c, op, dt, INT_MIN);
  if (op == PMGPU_DRIFT) c->sorted_valid = false;
  PMGPU_API_END(c)
}

int pmgpu_p3m_record_short_range(pmgpu_ctx* c, int32_t on) {
  PMGPU_API_BEGIN(c)
  graphs_stale(c);
  if (on && !c->acc_sr) dalloc(&c->acc_sr, std::max<int64_t>(c->n, 1));
  if (!on) dfree(&c->acc_sr);
  PMGPU_API_END(c)
}

int pmgpu_p3m_get_short_range(pmgpu_ctx* c, float* acc) {
  PMGPU_API_BEGIN(c)
  if (!acc) fail(PMGPU_ERR_INVALID_ARGUMENT, "acc is NULL");
  if (!c->acc_sr) fail(PMGPU_ERR_NOT_READY, "call pmgpu_p3m_record_short_range first");
  const int64_t n = c->n;
  if (n == 0) return PMGPU_OK;
  std::vector<float4> hv(n), ha(n);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  HIP_CHECK(hipMemcpy(hv.data(), c->vel[c->cur], sizeof(float4) * n, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(ha.data(), c->acc_sr, sizeof(float4) * n, hipMemcpyDeviceToHost));
  for (int64_t s = 0; s < n; ++s) {
    int32_t id;
    std::memcpy(&id, &hv[s].w, sizeof(id));
    if (id < 0 || id >= n) fail(PMGPU_ERR_DEVICE, "corrupted particle index");
    acc[3 * id] = ha[s].x;
    acc[3 * id + 1] = ha[s].y;
    acc[3 * id + 2] = ha[s].z;
  }
  PMGPU_API_END(c)
}

int pmgpu_add_accelerations(pmgpu_ctx* c, const float* acc) {
  PMGPU_API_BEGIN(c)
  graphs_stale(c);
  require_particles(c);
  if (!acc) fail(PMGPU_ERR_INVALID_ARGUMENT, "acc is NULL");
  if (!c->ext_stage) dalloc(&c->ext_stage, 3 * c->n);
  // the copy is ordered on the stream after the force that wrote c->acc and
  // returns only once the host array has been read
  HIP_CHECK(hipMemcpyAsync(c->ext_stage, acc, sizeof(float) * 3 * (size_t)c->n, hipMemcpyHostToDevice, c->stream));
  add_accelerations(c, c->ext_stage);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  PMGPU_API_END(c)
}

int pmgpu_get_integer_step_velocities(pmgpu_ctx* c, float* vel) {
  PMGPU_API_BEGIN(c)
  if (!vel) fail(PMGPU_ERR_INVALID_ARGUMENT, "vel is NULL");
  if (!c->vint) fail(PMGPU_ERR_NOT_READY, "call pmgpu_leapfrog(PMGPU_INTEGER_STEP) first");
  const int64_t n = c->n;
  std::vector<float4> h(n);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  HIP_CHECK(hipMemcpy(h.data(), c->vint, sizeof(float4) * n, hipMemcpyDeviceToHost));
  for (int64_t s = 0; s < n; ++s) {
    int32_t id;
    std::memcpy(&id, &h[s].w, sizeof(id));
    vel[3 * id] = h[s].x;
    vel[3 * id + 1] = h[s].y;
    vel[3 * id + 2] = h[s].z;
  }
  PMGPU_API_END(c)
}

int pmgpu_run(pmgpu_ctx* c, int32_t iterations, int32_t initial, int32_t p3m,
              int32_t* iterations_done) {
  PMGPU_API_BEGIN(c)
  require_particles(c);
  if (!c->greens_ready) fail(PMGPU_ERR_NOT_READY, "call pmgpu_init_greens first");
  if (p3m && !c->p3m_ready) fail(PMGPU_ERR_NOT_READY, "call pmgpu_p3m_configure first");
  if (iterations < 0) fail(PMGPU_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
  // run-loop iteration numbers continue across calls; escape[0] holds the
  // first iteration whose drift left the box
  const int base = c->run_iter;
  auto force = [&](int mode, float kick) {
    if (p3m) {
      p3m_force_overlapped(c, mode, kick);
    } else {
      pm_force(c, MODE_ACC | mode, kick);
    }
  };
  if (initial) {
    c->iter = base - 1;
    force(MODE_KICK | (iterations > 0 ? MODE_DRIFT : 0), 0.5f);
  } else if (iterations > 0) {
    leapfrog_op(c, PMGPU_DRIFT, 1.0f, base);
  }
  int it = 0;
  // Iterations that drift into a following one, two at a time, from a
  // captured graph (the per-step launches of a small problem cost more than
  // its kernels); the first is enqueued directly, which also makes any lazy
  // allocation before a capture.  Not while profiling (its events are host
  // state).
  if (!c->profiling && iterations >= 2) {
    c->iter = base;
    force(MODE_KICK | MODE_DRIFT, 1.0f);
    it = 1;
    hipGraphExec_t exec = run_graph(c, p3m != 0, force);
    if (exec && it + 2 < iterations) {
      hipLaunchKernelGGL(k_iter_set, dim3(1), dim3(1), 0, c->stream, c->escape, base + it);
      HIP_CHECK(hipGetLastError());
      for (; it + 2 < iterations; it += 2) HIP_CHECK(hipGraphLaunch(exec, c->stream));
    }
  }
  for (; it < iterations; ++it) {
    c->iter = base + it;
    force(MODE_KICK | (it + 1 < iterations ? MODE_DRIFT : 0), 1.0f);
  }
  c->run_iter = base + iterations;
  c->sorted_valid = false;
  if (iterations_done) {
    int32_t e = INT_MAX;
    HIP_CHECK(hipMemcpyAsync(&e, c->escape, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    int64_t done = (int64_t)e - base;
    *iterations_done = (int32_t)std::max<int64_t>(0, std::min<int64_t>(iterations, done));
  }
  PMGPU_API_END(c)
}

int pmgpu_escaped(pmgpu_ctx* c, int32_t* escaped) {
  PMGPU_API_BEGIN(c)
  if (!escaped) fail(PMGPU_ERR_INVALID_ARGUMENT, "escaped is NULL");
  *escaped = 0;
  if (c->n == 0) return PMGPU_OK;
  zero_fill_i32(c, c->counters + 8, 1);
  escape_check(c, c->counters + 8);
  int32_t h = 0;
  HIP_CHECK(hipMemcpyAsync(&h, c->counters + 8, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  *escaped = h != 0;
  PMGPU_API_END(c)
}

int pmgpu_get_grid(pmgpu_ctx* c, int32_t which, float* out) {
  PMGPU_API_BEGIN(c)
  if (!out) fail(PMGPU_ERR_INVALID_ARGUMENT, "out is NULL");
  HIP_CHECK(hipStreamSynchronize(c->stream));
  switch (which) {
    case PMGPU_GRID_DENSITY:
      HIP_CHECK(hipMemcpy(out, c->rho, sizeof(float) * c->M, hipMemcpyDeviceToHost));
      break;
    case PMGPU_GRID_POTENTIAL:
      HIP_CHECK(hipMemcpy(out, c->phi, sizeof(float) * c->M, hipMemcpyDeviceToHost));
      break;
    case PMGPU_GRID_FIELD:
      if (!c->field) dalloc(&c->field, 3 * c->M);
      field_grid(c);
      HIP_CHECK(hipStreamSynchronize(c->stream));
      HIP_CHECK(hipMemcpy(out, c->field, sizeof(float) * 3 * c->M, hipMemcpyDeviceToHost));
      break;
    case PMGPU_GRID_GREENS: {
      std::vector<float> g((size_t)c->K);
      HIP_CHECK(hipMemcpy(g.data(), c->greens, sizeof(float) * c->K, hipMemcpyDeviceToHost));
      const double m = (double)c->M;
      for (int64_t r = 0; r < (int64_t)c->ny * c->nz; ++r)
        for (int kx = 0; kx < c->nxh; ++kx) out[r * c->nxh + kx] = (float)((double)g[r * c->sp + kx] * m);
      break;
    }
    default:
      fail(PMGPU_ERR_INVALID_ARGUMENT, "unknown grid");
  }
  PMGPU_API_END(c)
}

int pmgpu_set_tuning(pmgpu_ctx* c, const int32_t* knobs) {
  PMGPU_API_BEGIN(c)
  graphs_stale(c);
  if (!knobs) fail(PMGPU_ERR_INVALID_ARGUMENT, "knobs is NULL");
  const int dep = knobs[0], itp = knobs[1], bpc = knobs[2], div = knobs[3], ik = knobs[4], sc = knobs[5];
  if (dep < DEPOSIT_CHUNK_MIN || dep > DEPOSIT_CHUNK_MAX)
    fail(PMGPU_ERR_INVALID_ARGUMENT, "deposit chunk out of range");
  if (itp < INTERP_CHUNK_MIN || itp > INTERP_CHUNK_MAX)
    fail(PMGPU_ERR_INVALID_ARGUMENT, "interpolation chunk out of range");
  if (bpc < 1 || bpc > BLOCKS_PER_CU_MAX) fail(PMGPU_ERR_INVALID_ARGUMENT, "blocks per CU out of range");
  if (div < SR_CELL_DIV || div > SR_CELL_DIV_MAX) fail(PMGPU_ERR_INVALID_ARGUMENT, "short-range cell division out of range");
  if (ik < 0 || ik > 3) fail(PMGPU_ERR_INVALID_ARGUMENT, "interpolation variant must be 0 .. 3");
  c->interp_kernel = ik;
  if (sc < 0 || sc > 2) fail(PMGPU_ERR_INVALID_ARGUMENT, "short-range chunking must be 0, 1 or 2");
  c->sr_chunking = sc;
  c->dep_chunk = dep;
  c->int_chunk = itp;
  c->blocks_per_cu = bpc;
  c->sr_div_set = true;
  if (div != c->sr_div || c->sr_divy != div || c->sr_divz != div || c->sr_divx_base != div ||
      c->sr_xfine_max != 4)
    sr_set_division(c, div);
  PMGPU_API_END(c)
}

int pmgpu_get_tuning(pmgpu_ctx* c, int32_t* knobs) {
  PMGPU_API_BEGIN(c)
  if (!knobs) fail(PMGPU_ERR_INVALID_ARGUMENT, "knobs is NULL");
  knobs[0] = c->dep_chunk;
  knobs[1] = c->int_chunk;
  knobs[2] = c->blocks_per_cu;
  knobs[3] = c->sr_div;
  knobs[4] = c->interp_kernel;
  knobs[5] = c->sr_chunking;
  PMGPU_API_END(c)
}

int pmgpu_profile_enable(pmgpu_ctx* c, int32_t on) {
  PMGPU_API_BEGIN(c)
  graphs_stale(c);
  c->profiling = on == 2 ? 2 : (on != 0 ? 1 : 0);
  PMGPU_API_END(c)
}

int pmgpu_profile_reset(pmgpu_ctx* c) {
  PMGPU_API_BEGIN(c)
  profile_collect(c);
  for (auto& s : c->stages) {
    s.total_ms = 0.0;
    s.launches = 0;
  }
  PMGPU_API_END(c)
}

int pmgpu_profile_read(pmgpu_ctx* c, int32_t max, const char** names, double* total_ms,
                       int64_t* launches, int32_t* count) {
  PMGPU_API_BEGIN(c)
  profile_collect(c);
  int k = 0;
  for (auto& s : c->stages) {
    if (k >= max) break;
    if (names) names[k] = s.name;
    if (total_ms) total_ms[k] = s.total_ms;
    if (launches) launches[k] = s.launches;
    ++k;
  }
  if (count) *count = k;
  PMGPU_API_END(c)
}

}  // extern "C"
