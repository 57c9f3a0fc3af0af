// Device-wide exclusive scan, zero fill, and the (tile, boundary class)
// counting sort that orders particles for the deposit and the interpolation.
//
// The sort is a counting sort: per-particle keys with an integer-atomic rank
// inside the key bucket, an exclusive scan of the bucket counts, and a
// scatter.  Ranks inside one bucket depend on atomic arrival order; nothing
// downstream depends on that order (the deposit sums in 64-bit fixed point,
// interpolation is per particle).
#include "pmgpu_internal.h"

namespace pmgpu {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 4096

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Exclusive scan of the 256 thread totals of a block; returns the exclusive
// prefix of this thread and writes the block total to *total.
__device__ int32_t block_exclusive_scan(int32_t v, int32_t* total) {
  __shared__ int32_t wave_tot[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) wave_tot[wave] = inc;
  __syncthreads();
  int32_t wbase = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    int32_t t = wave_tot[w];
    if (w < wave) wbase += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return wbase + inc - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int32_t* __restrict__ in,
                                                              int64_t n,
                                                              int32_t* __restrict__ sums) {
  // striped (coalesced) reads: the order of a sum does not matter
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  int32_t v[SCAN_ITEMS];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = base + (int64_t)k * SCAN_THREADS;
    v[k] = i < n ? in[i] : 0;
  }
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) s += v[k];
  int32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// out[i] = offsets[block] + exclusive prefix inside the block; the last block
// also writes out[n] = grand total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int32_t* __restrict__ in,
                                                             int32_t* __restrict__ out,
                                                             int64_t n,
                                                             const int32_t* __restrict__ offsets) {
  // striped (coalesced) loads into LDS, one padding int per 16 so that a
  // thread's 16 consecutive elements spread over the banks; each thread scans
  // its 16, the block scans the thread totals, and the results leave striped
  __shared__ int32_t tile[SCAN_TILE + SCAN_TILE / SCAN_ITEMS];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int32_t v[SCAN_ITEMS];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = base + (int64_t)k * SCAN_THREADS + threadIdx.x;
    v[k] = i < n ? in[i] : 0;
  }
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int e = k * SCAN_THREADS + threadIdx.x;
    tile[e + e / SCAN_ITEMS] = v[k];
  }
  __syncthreads();
  const int own = threadIdx.x * (SCAN_ITEMS + 1);
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = tile[own + k];
    s += v[k];
  }
  int32_t tot;
  int32_t pre = block_exclusive_scan(s, &tot);  // synchronises the block
  const int32_t off = offsets ? offsets[blockIdx.x] : 0;
  pre += off;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    tile[own + k] = pre;
    pre += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int e = k * SCAN_THREADS + threadIdx.x;
    const int64_t i = base + e;
    if (i < n) out[i] = tile[e + e / SCAN_ITEMS];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = off + tot;
}

inline int grid_for(int64_t n, int threads, int cap = 65536) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// ---------------------------------------------------------------------------
// Tile sort
// ---------------------------------------------------------------------------

__device__ __forceinline__ int wrap_cell(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}

// Anchor cell of one coordinate: round for NGP (source/pmMethod.cpp:207),
// floor for CIC/TSC ((int) truncation for x >= 0, pmMethod.cpp:219, 250),
// round for the textbook (centred) TSC.
__device__ __forceinline__ int anchor_of(float x, int scheme, int centred) {
  float f = (scheme == PMGPU_NGP || centred) ? roundf(x) : floorf(x);
  f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);  // NaN -> -1e9, keeps the int conversion defined
  return (int)f;
}

__device__ __forceinline__ int face_class(int local, int extent) {
  return local == 0 ? 0 : (local == extent - 1 ? 2 : 1);
}

// K particles per thread (a workgroup takes 256 K consecutive ones, its
// threads striding by 256): their loads are in flight together, and so are
// their rank atomics, instead of one load and one returning atomic after the
// other per thread (16.7M particles: 0.322 ms for the whole sort at K = 4
// against 0.347 at K = 1).  Small particle sets keep K = 1, which fills more
// of the GPU.
template <int K>
__global__ __launch_bounds__(256) void k_tile_keys(const float4* __restrict__ pm, int64_t n, int nx, int ny,
                                                   int nz, int ntx, int nty, int scheme, int centred,
                                                   int32_t* __restrict__ key, int32_t* __restrict__ rank,
                                                   int32_t* __restrict__ count, const int32_t* __restrict__ escape,
                                                   int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int64_t i0 = (int64_t)blockIdx.x * (256 * K) + threadIdx.x;
  const int lane = threadIdx.x & 63;
  float4 p[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int64_t i = i0 + 256 * q;
    p[q] = i < n ? pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  int32_t k[K];
  unsigned long long mine[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const bool active = i0 + 256 * q < n;
    int ax = wrap_cell(anchor_of(p[q].x, scheme, centred), nx);
    int ay = wrap_cell(anchor_of(p[q].y, scheme, centred), ny);
    int az = wrap_cell(anchor_of(p[q].z, scheme, centred), nz);
    int tx = ax / TILE_X, ty = ay / TILE_Y, tz = az / TILE_Z;
    int cls = face_class(ax - tx * TILE_X, TILE_X) + 3 * face_class(ay - ty * TILE_Y, TILE_Y) +
              9 * face_class(az - tz * TILE_Z, TILE_Z);
    k[q] = (tx + ntx * (ty + nty * tz)) * NCLASS + cls;
    mine[q] = wave_group(k[q], active);
  }
  // the group leaders' atomics of all K particles, then their results
  int32_t base[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    base[q] = 0;
    if (mine[q] && lane == __ffsll(mine[q]) - 1) base[q] = atomicAdd(&count[k[q]], (int32_t)__popcll(mine[q]));
  }
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int64_t i = i0 + 256 * q;
    const int leader = mine[q] ? __ffsll(mine[q]) - 1 : lane;
    const int32_t b = __shfl(base[q], leader);
    if (i < n) {
      key[i] = k[q];
      rank[i] = b + (int32_t)__popcll(mine[q] & ((1ull << lane) - 1ull));
    }
  }
}

__global__ void k_tile_scatter(const float4* __restrict__ pm_in, const float4* __restrict__ vel_in,
                               int64_t n, const int32_t* __restrict__ key,
                               const int32_t* __restrict__ rank, const int32_t* __restrict__ start,
                               float4* __restrict__ pm_out, float4* __restrict__ vel_out,
                               const int32_t* __restrict__ escape, int iteration) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // After an escape the keys were not computed: copy in place so that the
  // buffer swap on the host keeps the state.
  iteration = gate_iteration(escape, iteration);
  int64_t d = escape[0] <= iteration ? i : (int64_t)start[key[i]] + rank[i];
  pm_out[d] = pm_in[i];
  vel_out[d] = vel_in[i];
}

}  // namespace

void zero_fill_i32(pmgpu_ctx* c, int32_t* p, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->stream, p, n, 0);
  HIP_CHECK(hipGetLastError());
}

// Scratch needed by exclusive_scan_i32 for n elements (ints).
static size_t scan_scratch(int64_t n) {
  size_t s = 0;
  while (n > SCAN_TILE) {
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    s += 2 * (size_t)(nb + 1);
    n = nb;
  }
  return s;
}

static void scan_rec(pmgpu_ctx* c, const int32_t* in, int32_t* out, int64_t n, int32_t* tmp) {
  int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SCAN_THREADS), 0, c->stream, in, out, n,
                       (const int32_t*)nullptr);
    HIP_CHECK(hipGetLastError());
    return;
  }
  int32_t* sums = tmp;
  int32_t* offs = tmp + (nb + 1);
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, c->stream, in, n,
                     sums);
  HIP_CHECK(hipGetLastError());
  scan_rec(c, sums, offs, nb, tmp + 2 * (nb + 1));
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, c->stream, in, out,
                     n, (const int32_t*)offs);
  HIP_CHECK(hipGetLastError());
}

void exclusive_scan_i32(pmgpu_ctx* c, const int32_t* in, int32_t* out, int64_t n) {
  size_t need = scan_scratch(n);
  if (need > c->scan_tmp_elems)
    fail(PMGPU_ERR_NOT_READY, "internal: scan scratch too small (" + std::to_string(need) + ")");
  scan_rec(c, in, out, n, c->scan_tmp);
}

size_t scan_scratch_elems(int64_t n) { return scan_scratch(n); }

namespace {
// owner[w] = t for every work item w in [starts[t], starts[t+1])
__global__ void k_expand_owner(const int32_t* __restrict__ starts, int64_t n, int32_t* __restrict__ owner,
                               const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  for (int32_t w = starts[t]; w < starts[t + 1]; ++w) owner[w] = (int32_t)t;
}
}  // namespace

void expand_owner(pmgpu_ctx* c, const int32_t* starts, int64_t n, int32_t* owner, int iteration) {
  hipLaunchKernelGGL(k_expand_owner, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, starts, n, owner,
                     c->escape, iteration);
  HIP_CHECK(hipGetLastError());
}

void tile_sort(pmgpu_ctx* c) {
  if (c->general) return;  // the untiled path keeps the particles in place
  const int64_t nkeys = (int64_t)c->ntiles * NCLASS;
  int tok = stage_begin(c, "tile_sort");
  zero_fill_i32(c, c->cls_count, nkeys);
  unsigned g = (unsigned)((c->n + 255) / 256);
  // four keys per thread once that still gives every CU four workgroups
  const bool four = c->n >= (int64_t)1024 * c->num_cu * 4;
  const unsigned gk = four ? (unsigned)((c->n + 1023) / 1024) : g;
  auto keys = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(gk), dim3(256), 0, c->stream, c->pm[c->cur], c->n, c->nx, c->ny, c->nz, c->ntx,
                       c->nty, c->prm.interpolation, c->tsc_centred, c->key, c->rank, c->cls_count, c->escape,
                       c->iter);
  };
  if (four) keys(k_tile_keys<4>);
  else keys(k_tile_keys<1>);
  HIP_CHECK(hipGetLastError());
  exclusive_scan_i32(c, c->cls_count, c->cls_start, nkeys);
  int nxt = c->cur ^ 1;
  hipLaunchKernelGGL(k_tile_scatter, dim3(g), dim3(256), 0, c->stream, c->pm[c->cur],
                     c->vel[c->cur], c->n, c->key, c->rank, c->cls_start, c->pm[nxt], c->vel[nxt],
                     c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
  c->cur = nxt;
  c->sorted_valid = true;
}

}  // namespace pmgpu
