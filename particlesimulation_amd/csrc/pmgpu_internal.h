// Internal declarations shared by the HIP translation units of libpmgpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pmgpu.h"

namespace pmgpu {

// ---------------------------------------------------------------------------
// Run-loop iterations.  Every kernel of a force step takes its iteration and
// returns at once when escape[0] (the first iteration whose drift left the
// box) is not after it, so a whole run is enqueued without host round trips
// and stops where the reference stops.  Kernels replayed from a captured
// graph take ITER_DEVICE instead and read their iteration from escape[1],
// which the graph advances after each iteration it holds.
constexpr int ITER_DEVICE = -2147483647;  // INT_MIN + 1 (INT_MIN: untagged, never gated)
__device__ __forceinline__ int gate_iteration(const int32_t* escape, int iteration) {
  return iteration == ITER_DEVICE ? escape[1] : iteration;
}

// ---------------------------------------------------------------------------
// Mesh tiling used by the deposit and the interpolation.  A tile is
// TILE_X x TILE_Y x TILE_Z cells; particles are sorted by (tile, boundary
// class), the class recording per axis whether the anchor cell sits on the
// tile's low face, inside, or on its high face.  A tile then finds the
// particles of its 26 neighbours that reach into it as a few contiguous
// class ranges.
constexpr int TILE_X = 16;
constexpr int TILE_Y = 16;
constexpr int TILE_Z = 8;
constexpr int TILE_CELLS = TILE_X * TILE_Y * TILE_Z;  // 2048
constexpr int NCLASS = 27;
// Work-item size of the deposit (virtual particles per work item).
// work-item sizes (particles) of the deposit and interpolation passes: defaults
// and limits of the tunable values (pmgpu_set_tuning)
constexpr int DEPOSIT_CHUNK = 1024, DEPOSIT_CHUNK_MIN = 1024, DEPOSIT_CHUNK_MAX = 16384;
constexpr int INTERP_CHUNK = 2048, INTERP_CHUNK_MIN = 256, INTERP_CHUNK_MAX = 16384;
constexpr int BLOCKS_PER_CU = 32, BLOCKS_PER_CU_MAX = 32;
// automatic interpolation variant: the gather variant for grids of fewer
// tiles than this many per CU (too few work items to fill the device), the
// tiled variant otherwise
constexpr int INTERP_GATHER_TILES_PER_CU = 4;

// P3M short range: cell side >= cutoff / div, stencil +-div cells (div tunable)
constexpr int SR_CELL_DIV = 1, SR_CELL_DIV_MAX = 6;
// neighbourhood rows (2 divy + 1)(2 divz + 1) of any cell geometry: the
// isotropic maximum, which also bounds the anisotropic geometries tried
constexpr int SR_ROWS_MAX = (2 * SR_CELL_DIV_MAX + 1) * (2 * SR_CELL_DIV_MAX + 1);
// the pair loop reads j in groups of SR_JGROUP; the sorted particle copy
// carries SR_JGROUP - 1 zeroed entries past the last particle for it
constexpr int SR_JGROUP = 8;
constexpr int SR_ODD_MAX = 64;  // particles off the common mass that the uniform-mass pair loop corrects
constexpr int64_t SR_CELLS_MAX = 128ll * 1024 * 1024;  // short-range cells, x refinement included
constexpr int SR_TABLE = 512;  // padded reference table (the reference uses 500 entries)
constexpr int SR_TABLE_REF = 500;
// The accurate pair table: at most SR_TABLE_LOG bins of s = r^2 (times a
// scale) taken from the float bits of s, 2^(23 - SR_LOG_SHIFT) bins per
// binade, the last of them zero.  Two layouts:
//   absolute (over r^2, unsoftened, a <= rc): the short-range pass scales
//     positions by 2^-q so that the first bin starts at the smallest normal
//     float, whose bits >> SR_LOG_SHIFT are SR_LOG_ABS0; the bin of s sits at
//     entry (bits(s) >> SR_LOG_SHIFT) - SR_LOG_ABS0, one shift of the bits
//     with the constant folded into the LDS address;
//   wrapped (otherwise): bin k sits at k mod SR_LOG_WRAP, a shift and a mask.
// The absolute layout also comes whole (no clamp of s): entry bits(s) >>
// SR_LOG_SHIFT for every s from 0 to the largest r^2 the short-range copy's
// positions allow (clamped to half a box outside the particle box), when that
// fits SR_TABLE_EXT entries.
constexpr int SR_LOG_SHIFT = 16;
constexpr int SR_LOG_WRAP = 4096;
constexpr int SR_TABLE_LOG = SR_LOG_WRAP;
constexpr int SR_LOG_ABS0 = (1 << 23) >> SR_LOG_SHIFT;
constexpr int SR_TABLE_EXT = 6144;

struct Stage {
  const char* name;
  double total_ms = 0.0;
  int64_t launches = 0;
};

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Error carrying a status code; thrown inside the library only, converted to
// a return code at the C ABI.
struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
void check_hip(hipError_t e, const char* what);
#define HIP_CHECK(x) ::pmgpu::check_hip((x), #x)

}  // namespace pmgpu

struct pmgpu_ctx {
  // configuration
  pmgpu_pm_params prm{};  // prm.interpolation holds the assignment order (NGP, CIC or TSC)
  int tsc_centred = 0;     // PMGPU_TSC_CENTRED requested: TSC about the nearest grid point
  int nx = 0, ny = 0, nz = 0, nxh = 0;  // nxh = nx/2 + 1
  int sp = 0;                            // pitch of the half spectrum's rows: nxh rounded up to 16
  bool general = false;                  // a side is not a power of two in [32, 1024]: the untiled path
  int64_t M = 0, K = 0;                  // cells, half-spectrum entries (sp * ny * nz, pitch included)
  int ntx = 0, nty = 0, ntz = 0, ntiles = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // particles (sorted order, double buffered)
  int64_t n = 0;
  float4* pm[2] = {nullptr, nullptr};   // x, y, z, mass
  float4* vel[2] = {nullptr, nullptr};  // vx, vy, vz, original index (int bits)
  float4* acc = nullptr;                // ax, ay, az, -
  float4* vint = nullptr;               // integer-step velocities (lazy)
  float* ext_stage = nullptr;           // host-evaluated accelerations, caller order, 3 n (lazy)
  float4* acc_sr = nullptr;             // the short-range part of acc, when recorded (sorted order)
  int cur = 0;                          // which buffer holds the state
  int32_t* key = nullptr;
  int32_t* rank = nullptr;
  double mass_abs_sum = 0.0;
  // mass classes of the short-range pass: when all but at most SR_ODD_MAX
  // particles have the mass sr_m0, the pair loop sums at that mass and the
  // others (original ids ascending in sr_odd_ids, their sorted positions in
  // sr_odd_pos) are corrected per work item; sr_nodd = -1 otherwise
  float sr_m0 = 0.0f;
  int sr_nodd = -1;
  int32_t* sr_odd_ids = nullptr;        // SR_ODD_MAX
  int32_t* sr_odd_pos = nullptr;        // SR_ODD_MAX
  int fx_shift = 0;                     // fixed-point scale 2^fx_shift
  bool sorted_valid = false;            // tile sort matches current positions

  // tile sort
  int32_t* cls_count = nullptr;         // ntiles*27 (+1)
  int32_t* cls_start = nullptr;         // ntiles*27 + 1
  int32_t* tile_nchunk = nullptr;       // ntiles (+1)
  int32_t* tile_chunk_start = nullptr;  // ntiles + 1
  int32_t* tile_slot = nullptr;         // multi-chunk accumulation slot or -1
  int32_t* slot_tile = nullptr;         // tile of each used slot
  int32_t* int_nchunk = nullptr;        // interpolation work items per tile (+1)
  int32_t* int_chunk_start = nullptr;   // ntiles + 1
  int32_t* dep_owner = nullptr;         // tile of each deposit work item
  int32_t* int_owner = nullptr;         // tile of each interpolation work item
  int32_t* sr_owner = nullptr;          // short-range cell of each work item
  int4* dep_table = nullptr;            // neighbour (dx,dy,dz,class) ranges of the scheme
  int dep_table_len = 0;
  int num_cu = 256;
  int dep_chunk = pmgpu::DEPOSIT_CHUNK;  // tunables (pmgpu_set_tuning)
  int int_chunk = pmgpu::INTERP_CHUNK;
  int blocks_per_cu = pmgpu::BLOCKS_PER_CU;
  int sr_div = pmgpu::SR_CELL_DIV;      // the division knob (pmgpu_set_tuning knobs[3])
  bool sr_div_set = false;              // set explicitly (else chosen at p3m_configure)
  // the short-range cell geometry: cell sides >= cutoff / sr_divy in y,
  // cutoff / sr_divz in z and cutoff / (sr_divx_base sr_xfine) in x; all three
  // equal sr_div unless p3m_configure picked anisotropic cells for the particles
  int sr_divy = pmgpu::SR_CELL_DIV, sr_divz = pmgpu::SR_CELL_DIV, sr_divx_base = pmgpu::SR_CELL_DIV;
  int sr_xfine = 1;                     // x cells are this many times finer than sr_divx_base gives
  int sr_xfine_max = 4;                 // the x refinement sr_cells_setup starts from
  int interp_kernel = 0;                // 0 auto, 1 tiled (LDS halo), 2 gather from an E grid
  int sr_chunking = 0;                  // 0 auto (rows), 1 cells, 2 rows of cells
  float4* efield = nullptr;             // M, the gather variant's E grid
  int32_t* counters = nullptr;          // small device counters
  long long* fx_accum = nullptr;        // multi-chunk tiles x 512 fixed-point sums
  int64_t fx_slots = 0;
  int32_t* scan_tmp = nullptr;
  size_t scan_tmp_elems = 0;
  // a second stream for the short-range cell sort, which runs beside the
  // PM force in pmgpu_run, with its own scan scratch
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int32_t* scan_tmp_side = nullptr;
  size_t scan_tmp_side_elems = 0;

  // grids
  float* rho = nullptr;      // M
  float2* spec = nullptr;    // K
  float* phi = nullptr;      // M
  float* greens = nullptr;   // K (symmetrised, /M)
  // G in the z pass's tile order (per y and column tile, the tile's L x ncols
  // values contiguous), built by pmgpu_init_greens; any rewrite of greens
  // clears greens_z_valid
  float* greens_z = nullptr;
  bool greens_z_valid = false;
  float* field = nullptr;    // 3M, only on request
  float2* tw_x = nullptr;    // twiddles for nx/2
  float2* tw_y = nullptr;
  float2* tw_z = nullptr;
  float2* tw_r2c = nullptr;  // exp(-2 pi i k / nx), k in [0, nx/2]
  float2* tw_gx = nullptr;   // general grids: exp(-2 pi i m / nx), m in [0, nx)
  unsigned long long* fx_grid = nullptr;  // general grids: fixed-point density, M
  bool greens_ready = false;
  std::vector<float> mlp_weights;  // learned correction (empty: analytic)

  // P3M
  bool p3m_ready = false;
  pmgpu_p3m_params p3m{};
  int mcx = 0, mcy = 0, mcz = 0, ncell_sr = 0;
  float hcx = 0, hcy = 0, hcz = 0;
  float2* sr_table = nullptr;           // SR_TABLE_EXT entries (the largest of the tables)
  float sr_inv_d2 = 0.0f;
  float sr_log_scale = 1.0f;             // the accurate table's argument s = r^2 * scale
  uint32_t sr_log_lo = 0, sr_log_hi = 0;  // float bits of its first and last (zero) bin
  int sr_log_mode = 3;                    // 2: absolute layout over r^2 (scaled), 3: wrapped,
                                          // 4: absolute and whole (no clamp)
  // the pass's sorted copy holds positions x sr_pos_scale and masses x
  // sr_mass_scale (powers of two; 1 outside the absolute layout)
  float sr_pos_scale = 1.0f, sr_mass_scale = 1.0f;
  int sr_tmax = 0;
  int32_t* sr_count = nullptr;          // ncell_sr + 1
  int32_t* sr_start = nullptr;          // ncell_sr + 1
  int32_t* sr_key = nullptr;            // n
  int32_t* sr_rank = nullptr;           // n
  float4* sr_pm = nullptr;              // n, SR-cell order
  int32_t* sr_idx = nullptr;            // n, index into the tile-sorted arrays
  int32_t* sr_nchunk = nullptr;         // ncell_sr + 1
  int32_t* sr_chunk_start = nullptr;    // ncell_sr + 1
  unsigned long long* pair_count = nullptr;

  double* diag_scratch = nullptr;       // FP64 diagnostics partial sums

  // run loop state
  int32_t* escape = nullptr;   // [0] = first iteration with an escaped particle, or INT_MAX;
                               // [1] = the iteration of graph-replayed kernels (ITER_DEVICE)
  int iter = 0;                // tag of the kernels being enqueued (INT_MIN: standalone call)
  int run_iter = 0;            // next run-loop iteration number
  // two run-loop iterations captured as one graph, per P3M flag and per
  // particle-buffer parity at the graph's start (pmgpu_run); any call that
  // can change a kernel argument bumps graph_gen, which retires them
  hipGraphExec_t run_exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int64_t run_exec_gen[2] = {-1, -1};
  int64_t graph_gen = 0;

  // profiling: 0 off, 1 every stage, 2 only the short-range pair kernel
  int profiling = 0;
  std::vector<pmgpu::Stage> stages;
  std::vector<hipEvent_t> ev_pool;
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
};

namespace pmgpu {

// profiling helpers (api.hip)
int stage_begin(pmgpu_ctx* c, const char* name);
void stage_end(pmgpu_ctx* c, int token);

// launchers (one per translation unit)
void zero_fill_i32(pmgpu_ctx* c, int32_t* p, int64_t n);
void exclusive_scan_i32(pmgpu_ctx* c, const int32_t* in, int32_t* out, int64_t n);
// owner[w] = t for the work items w in [starts[t], starts[t+1]), t < n (O(1) item lookup)
void expand_owner(pmgpu_ctx* c, const int32_t* starts, int64_t n, int32_t* owner, int iteration);

void tile_sort(pmgpu_ctx* c);
size_t scan_scratch_elems(int64_t n);
std::vector<int4> deposit_range_table(int scheme);
void deposit(pmgpu_ctx* c);
// grids that are not all powers of two in [32, 1024] (csrc/dft.hip)
bool general_grid(int nx, int ny, int nz);
void deposit_general(pmgpu_ctx* c);
void fft_general(pmgpu_ctx* c);
void fft_forward_green_inverse(pmgpu_ctx* c);
// greens -> greens_z (power-of-two grids)
void greens_block_for_z(pmgpu_ctx* c);
// unnormalised forward R2C of a real grid into a half spectrum (in may be rho)
void fft_r2c_forward(pmgpu_ctx* c, const float* in, float2* spec);
// out = unnormalised C2R of (src * g); src is kept, work holds the spectrum
void fft_mul_c2r_inverse(pmgpu_ctx* c, const float2* src, const float* g, float2* work, float* out);
void greens_compute(pmgpu_ctx* c, const float* mlp_dev);
void interpolate(pmgpu_ctx* c, int mode, float kick, int iteration);
// acc[s] += add[3 id(s) ..], id(s) the original index of sorted particle s
void add_accelerations(pmgpu_ctx* c, const float* add);
void field_grid(pmgpu_ctx* c);
void leapfrog_op(pmgpu_ctx* c, int op, float dt, int iteration);
void p3m_setup(pmgpu_ctx* c);
void sr_cells_setup(pmgpu_ctx* c);  // short-range cell grid for c->p3m and the divisions
void sr_set_division(pmgpu_ctx* c, int div);  // isotropic division, cells rebuilt if configured
void sr_choose_cells(pmgpu_ctx* c);           // the cell geometry with the fewest candidates
void p3m_short_range(pmgpu_ctx* c, int mode, float kick, int iteration);
// its two halves: the cell sort and the pair pass
void p3m_sort(pmgpu_ctx* c, int iteration);
void p3m_pairs(pmgpu_ctx* c, int mode, float kick, int iteration);
void p3m_count_pairs(pmgpu_ctx* c, int64_t* accepted, int64_t* tested, int64_t* issued);
// learned correction (mlp.hip)
struct MlpDevice;
MlpDevice* mlp_device_create();
void mlp_device_destroy(MlpDevice* d);
void mlp_pack(MlpDevice* d, const float* params_dev, hipStream_t s);
// out[k] = y(k) on the half spectrum (rows of `pitch` entries), or
// greens[k] * (1 + y(k)) when greens is given
void mlp_forward_grid(MlpDevice* d, int nx, int ny, int nz, int pitch, int scheme, const float* greens,
                      float* out, hipStream_t s);
// grad (+)= scale * d/dparams sum_k up[k] y(k)
void mlp_backward_grid(MlpDevice* d, int nx, int ny, int nz, int pitch, int scheme, const float* up, float scale,
                       bool accumulate, float* grad, hipStream_t s);
void mlp_apply_to_greens(pmgpu_ctx* c);
void escape_check(pmgpu_ctx* c, int32_t* flag_dev);
void sr_table_host(float rc, float a, float eps, int cs, std::vector<float>& table, float& delta2);

// XCD-aware dynamic work distribution.  The hardware places workgroup b on
// XCD b % 8, each XCD with its own L2.  Work items [0, total) are split into
// NXCD contiguous segments with one counter each (cnt[0..NXCD), zeroed before
// the launch); a workgroup takes items from its XCD's segment, so an XCD's
// L2 holds the data of neighbouring items, and then from the other segments.
// Returns -1 when every segment is exhausted.
constexpr int NXCD = 8;
__device__ __forceinline__ int next_item(int32_t* cnt, int total, int xcd) {
  for (int k = 0; k < NXCD; ++k) {
    const int x = (xcd + k) & (NXCD - 1);
    const int b = (int)((int64_t)total * x / NXCD), e = (int)((int64_t)total * (x + 1) / NXCD);
    if (b >= e) continue;
    const int w = b + atomicAdd(&cnt[x], 1);
    if (w < e) return w;
  }
  return -1;
}

// Static XCD-aware assignment for a grid-stride loop over work items
// [0, total) ordered so that neighbouring items share data: workgroup b
// works through the segment of its XCD (b % NXCD), striding over that XCD's
// workgroups, so neighbouring items meet in one L2.  With fewer workgroups
// than XCDs, the plain grid stride.
struct XcdRange {
  int first, end, stride;
};
__device__ __forceinline__ XcdRange xcd_range(int total) {
  if (gridDim.x < (unsigned)NXCD) return {(int)blockIdx.x, total, (int)gridDim.x};
  const int x = (int)(blockIdx.x % NXCD);
  const int b = (int)((int64_t)total * x / NXCD), e = (int)((int64_t)total * (x + 1) / NXCD);
  const int nb = ((int)gridDim.x - x + NXCD - 1) / NXCD;  // workgroups on XCD x
  return {b + (int)(blockIdx.x / NXCD), e, nb};
}

// The lanes of this wavefront whose key equals this lane's (0 for an
// inactive lane): one ballot per distinct key.
__device__ __forceinline__ unsigned long long wave_group(int32_t key, bool active) {
  unsigned long long pending = __ballot(active);
  unsigned long long mine = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int32_t lk = __builtin_amdgcn_readlane(key, leader);
    const unsigned long long same = __ballot(active && key == lk) & pending;
    if (active && key == lk) mine = same;
    pending &= ~same;
  }
  return mine;
}

// Rank of this lane's key among all lanes of the grid that share it: one
// integer atomic per distinct key per wavefront (lanes with equal keys are
// ranked by lane order), so densely populated keys do not serialise 64
// atomics on one address.
__device__ __forceinline__ int32_t wave_rank(int32_t* count, int32_t key, bool active) {
  const int lane = threadIdx.x & 63;
  // group the lanes by key, then one atomic instruction for all groups at
  // once: the leaders' atomics overlap instead of waiting for each other's
  // return
  const unsigned long long mine = wave_group(key, active);
  const int leader = active ? __ffsll((unsigned long long)mine) - 1 : lane;
  int32_t base = 0;
  if (active && lane == leader) base = atomicAdd(&count[key], (int32_t)__popcll(mine));
  base = __shfl(base, leader);
  return active ? base + (int32_t)__popcll(mine & ((1ull << lane) - 1ull)) : 0;
}

// escapedComputationalBox (source/pmMethod.cpp:146-156): inclusive [0, box];
// an axis with box <= 0 is not tested (periodic use of the mesh).
__device__ __forceinline__ bool inside_box(float x, float y, float z, float3 box) {
  return (box.x <= 0.f || (x >= 0.f && x <= box.x)) && (box.y <= 0.f || (y >= 0.f && y <= box.y)) &&
         (box.z <= 0.f || (z >= 0.f && z <= box.z));
}

// interpolate/p3m modes
constexpr int MODE_ACC = 1;     // write acceleration
constexpr int MODE_KICK = 2;    // v += kick * a
constexpr int MODE_DRIFT = 4;   // x += v after the kick (fused next drift)

}  // namespace pmgpu
