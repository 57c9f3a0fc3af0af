// Mass deposition (PMMethod::spreadMass, source/pmMethod.cpp:200-277) as a
// deterministic segmented reduction over the tile-sorted particles.
//
// Each tile (TILE_X x TILE_Y x TILE_Z cells) gathers every particle whose
// assignment stencil reaches into it: its own particles (one contiguous
// range) plus the boundary-class ranges of its 26 neighbours.  That virtual
// list is cut into work items of dep_chunk particles (tunable).  A work item sums
// its contributions in LDS as 64-bit fixed point (integer adds are
// associative, so the sum does not depend on thread or atomic order), then
//  - a tile with one work item writes its TILE_CELLS cells straight to the density
//    grid (every cell is written exactly once: no clearing pass, no float
//    atomics),
//  - a tile with several work items adds its partial sums into a per-tile
//    fixed-point slot with integer atomics; a finalize pass converts them.
// The contribution of one particle to one cell is formed in FP32 in the
// reference's operation order ((m/8)*fx*fy*fz for TSC, m*tx*ty*tz for CIC)
// and quantised at 2^-shift with shift chosen so that no cell sum can
// overflow 2^62.
#include "pmgpu_internal.h"

namespace pmgpu {

namespace {

constexpr int DEP_THREADS = 256;
constexpr int MAX_RANGES = 101;  // 3 own ranges + 98 neighbour class ranges (TSC)

// Slots of the device counter block (ctx->counters).
constexpr int CNT_SLOTS = 0;

__device__ __forceinline__ int wrapc(int a, int n) {
  if ((unsigned)a < (unsigned)n) return a;  // the common case: no division
  a %= n;
  return a < 0 ? a + n : a;
}

__device__ __forceinline__ int anchor_of(float x, int scheme, int centred) {
  float f = (scheme == PMGPU_NGP || centred) ? roundf(x) : floorf(x);
  f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
  return (int)f;
}

// Per-axis stencil: first offset relative to the anchor and weights.
// TSC: PMMethod::TSCAssignmentFunc (pmMethod.cpp:187-198) values f(d,-1),
// f(d,0), f(d,1); the 1/8 factor is applied to the mass.
__device__ __forceinline__ void axis_weights(int scheme, float d, float w[3]) {
  if (scheme == PMGPU_TSC) {
    w[0] = (0.5f - d) * (0.5f - d);
    w[1] = 1.5f - 2.0f * d * d;
    w[2] = (0.5f + d) * (0.5f + d);
  } else if (scheme == PMGPU_CIC) {
    w[0] = 1.0f - d;
    w[1] = d;
    w[2] = 0.0f;
  } else {
    w[0] = 1.0f;
    w[1] = 0.0f;
    w[2] = 0.0f;
  }
}

// Ranges of tile t's virtual particle list and their lengths.  Entries 0-2
// are the tile's own particles: 0 the interior class (13: the anchor at
// least one cell from every face, so every stencil cell of any scheme lies in
// the tile), 1 classes 0-12, 2 classes 14-26; entries 3.. come from the
// host-built neighbour table.
constexpr int OWN_RANGES = 3, INTERIOR_CLASS = 13;
__device__ __forceinline__ void tile_range(int r, int t, int ntx, int nty, int ntz,
                                           const int4* __restrict__ table,
                                           const int32_t* __restrict__ cls_start, int32_t* start,
                                           int32_t* len) {
  if (r < OWN_RANGES) {
    const int lo = r == 0 ? INTERIOR_CLASS : (r == 1 ? 0 : INTERIOR_CLASS + 1);
    const int hi = r == 0 ? INTERIOR_CLASS + 1 : (r == 1 ? INTERIOR_CLASS : NCLASS);
    int32_t s = cls_start[t * NCLASS + lo];
    *start = s;
    *len = cls_start[t * NCLASS + hi] - s;
    return;
  }
  int4 e = table[r - OWN_RANGES];
  int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  int ux = wrapc(tx + e.x, ntx), uy = wrapc(ty + e.y, nty), uz = wrapc(tz + e.z, ntz);
  int k = (ux + ntx * (uy + nty * uz)) * NCLASS + e.w;
  int32_t s = cls_start[k];
  *start = s;
  *len = cls_start[k + 1] - s;
}

// Per tile: number of deposit work items (>= 1: every tile writes its cells)
// and of interpolation work items (own particles only), plus a fixed-point
// slot for tiles split over several deposit work items.
__global__ void k_deposit_plan(int ntiles, int ntx, int nty, int ntz, int nranges,
                               const int4* __restrict__ table,
                               const int32_t* __restrict__ cls_start,
                               int32_t* __restrict__ dep_nchunk, int32_t* __restrict__ int_nchunk,
                               int32_t* __restrict__ slot, int32_t* __restrict__ slot_tile,
                               int32_t* __restrict__ counters, int64_t max_slots, int dep_chunk,
                               int int_chunk, const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  // one wave per tile, its lanes over the ranges
  const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (t >= ntiles) return;
  long long total = 0;
  for (int r = lane; r < nranges; r += 64) {
    int32_t s, l;
    tile_range(r, t, ntx, nty, ntz, table, cls_start, &s, &l);
    total += l;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) total += __shfl_down(total, o, 64);
  if (lane != 0) return;
  const int32_t own = cls_start[t * NCLASS + NCLASS] - cls_start[t * NCLASS];  // the tile's own particles
  int nch = (int)((total + dep_chunk - 1) / dep_chunk);
  if (nch < 1) nch = 1;
  dep_nchunk[t] = nch;
  int_nchunk[t] = (own + int_chunk - 1) / int_chunk;
  int s = -1;
  if (nch > 1) {
    s = atomicAdd(&counters[CNT_SLOTS], 1);
    if (s >= max_slots) s = -2;  // cannot happen with the host's bound; flagged below
    else slot_tile[s] = t;
  }
  slot[t] = s;
}

// Fixed-point value of a float contribution, round(x * 2^shift) with ties to
// even, as a two's-complement 64-bit integer, in FP32 and integer operations
// only (scale = 2^shift).  t = x * scale is exact, and |t| < 2^62 for every
// contribution (the host keeps 1.5 * sum|m| * 2^shift below 2^62).  Then
// hi = trunc(t / 2^31) is exact (at most 24 significant bits), the remainder
// t - hi 2^31 is exact in FP32 (|rem| < 2^31, a multiple of t's ulp), and
// hi 2^31 is even, so hi 2^31 + rint(rem) is t rounded to nearest even.
// (The deposit passes t already scaled: it folds the scale into the mass.)
__device__ __forceinline__ unsigned long long to_fixed_scaled(float t) {
  const int hi = (int)(t * 0x1p-31f);
  const float rem = fmaf((float)hi, -0x1p31f, t);
  const int lo = (int)rintf(rem);
  return (unsigned long long)(((long long)hi << 31) + (long long)lo);
}

// Eight waves per SIMD (at most 64 VGPRs; eight workgroups' LDS fits a CU):
// a work item's particles are few per thread, so their load latency is
// hidden by other waves, not within one (512^3, 16.7M particles: 0.66 ms
// against 0.75 ms at six waves).
template <int SCHEME>
__global__ __launch_bounds__(DEP_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void k_deposit(
    const float4* __restrict__ pm, int nx, int ny, int nz, int ntx, int nty, int ntz,
    int ntiles, int nranges, const int4* __restrict__ table,
    const int32_t* __restrict__ cls_start, const int32_t* __restrict__ chunk_start,
    const int32_t* __restrict__ owner, const int32_t* __restrict__ slot, long long* __restrict__ fx_accum,
    int shift, float* __restrict__ rho, int dep_chunk, int centred, const int32_t* __restrict__ escape,
    int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  __shared__ unsigned long long acc[TILE_CELLS];
  __shared__ int32_t rstart[MAX_RANGES], rpre[MAX_RANGES + 1];
  __shared__ int s_tile, s_chunk, s_nchunk;

  const int total = chunk_start[ntiles];
  const float scale = ldexpf(1.0f, shift);
  constexpr int LO = SCHEME == PMGPU_TSC ? -1 : 0;
  constexpr int CNT = SCHEME == PMGPU_TSC ? 3 : (SCHEME == PMGPU_CIC ? 2 : 1);

  const XcdRange wr = xcd_range(total);
  for (int w = wr.first; w < wr.end; w += wr.stride) {
    if (threadIdx.x == 0) {
      const int t = owner[w];
      s_tile = t;
      s_chunk = w - chunk_start[t];
      s_nchunk = chunk_start[t + 1] - chunk_start[t];
    }
    for (int i = threadIdx.x; i < TILE_CELLS; i += DEP_THREADS) acc[i] = 0ull;
    __syncthreads();
    const int t = s_tile;
    if (threadIdx.x < nranges) {
      int32_t s, l;
      tile_range(threadIdx.x, t, ntx, nty, ntz, table, cls_start, &s, &l);
      rstart[threadIdx.x] = s;
      rpre[threadIdx.x + 1] = l;
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // inclusive scan of the range lengths by the first wave
      const int lane = threadIdx.x;
      int carry = 0;
      for (int base = 0; base < nranges; base += 64) {
        const int r = base + lane;
        int v = r < nranges ? rpre[r + 1] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int u = __shfl_up(v, o, 64);
          if (lane >= o) v += u;
        }
        if (r < nranges) rpre[r + 1] = carry + v;
        carry += __shfl(v, 63, 64);
      }
      if (lane == 0) rpre[0] = 0;
    }
    __syncthreads();
    const int ox = (t % ntx) * TILE_X, oy = ((t / ntx) % nty) * TILE_Y,
              oz = (t / (ntx * nty)) * TILE_Z;
    const int vbeg = s_chunk * dep_chunk;
    const int vend = min(vbeg + dep_chunk, rpre[nranges]);
    // range of this thread's first virtual index; later ones advance linearly
    int a = 0;
    {
      const int v0 = vbeg + (int)threadIdx.x;
      int b = nranges;
      while (b - a > 1) {
        const int mid = (a + b) >> 1;
        if (rpre[mid] <= v0) a = mid;
        else b = mid;
      }
    }
    // the next particle's load is issued before this one's contributions
    float4 pnext = make_float4(0.f, 0.f, 0.f, 0.f);
    {
      const int v = vbeg + (int)threadIdx.x;
      if (v < vend) {
        while (rpre[a + 1] <= v) ++a;
        pnext = pm[rstart[a] + (v - rpre[a])];
      }
    }
    for (int v = vbeg + (int)threadIdx.x; v < vend; v += DEP_THREADS) {
      const float4 p = pnext;
      const int vn = v + DEP_THREADS;
      if (vn < vend) {
        while (rpre[a + 1] <= vn) ++a;
        pnext = pm[rstart[a] + (vn - rpre[a])];
      }
      const int ax = anchor_of(p.x, SCHEME, centred), ay = anchor_of(p.y, SCHEME, centred), az = anchor_of(p.z, SCHEME, centred);
      float wx[3], wy[3], wz[3];
      axis_weights(SCHEME, p.x - (float)ax, wx);
      axis_weights(SCHEME, p.y - (float)ay, wy);
      axis_weights(SCHEME, p.z - (float)az, wz);
      // the mass times 2^shift: every product below is the unscaled one times
      // a power of two, so its rounding is the same
      const float m = (SCHEME == PMGPU_TSC ? p.w / 8.0f : p.w) * scale;
      if (v < rpre[1]) {
        // interior class: no stencil cell leaves the tile, so no wrap per cell
        // and no mask; the cells sit at fixed offsets from the first
        const int base = (wrapc(ax, nx) - ox + LO) + TILE_X * ((wrapc(ay, ny) - oy + LO) + TILE_Y * (wrapc(az, nz) - oz + LO));
#pragma unroll
        for (int i = 0; i < CNT; ++i) {
          const float t1 = m * wx[i];
#pragma unroll
          for (int j = 0; j < CNT; ++j) {
            const float t2 = t1 * wy[j];
#pragma unroll
            for (int k = 0; k < CNT; ++k)
              atomicAdd(&acc[base + i + TILE_X * (j + TILE_Y * k)], to_fixed_scaled(t2 * wz[k]));
          }
        }
        continue;
      }
      // local coordinates of the stencil cells relative to the tile, after
      // periodic wrap, and whether each lies in the tile
      int lx[3], ly[3], lz[3];
      bool okx[3], oky[3], okz[3];
#pragma unroll
      for (int i = 0; i < CNT; ++i) {
        lx[i] = wrapc(ax + LO + i, nx) - ox;
        ly[i] = wrapc(ay + LO + i, ny) - oy;
        lz[i] = wrapc(az + LO + i, nz) - oz;
        okx[i] = (unsigned)lx[i] < (unsigned)TILE_X;
        oky[i] = (unsigned)ly[i] < (unsigned)TILE_Y;
        okz[i] = (unsigned)lz[i] < (unsigned)TILE_Z;
      }
#pragma unroll
      for (int i = 0; i < CNT; ++i) {
        if (!okx[i]) continue;
        const float t1 = m * wx[i];
#pragma unroll
        for (int j = 0; j < CNT; ++j) {
          if (!oky[j]) continue;
          const float t2 = t1 * wy[j];
#pragma unroll
          for (int k = 0; k < CNT; ++k) {
            if (!okz[k]) continue;
            const float t3 = t2 * wz[k];
            atomicAdd(&acc[lx[i] + TILE_X * (ly[j] + TILE_Y * lz[k])], to_fixed_scaled(t3));
          }
        }
      }
    }
    __syncthreads();
    if (s_nchunk == 1) {
      for (int i = threadIdx.x; i < TILE_CELLS; i += DEP_THREADS) {
        const int lx = i % TILE_X, ly = (i / TILE_X) % TILE_Y, lz = i / (TILE_X * TILE_Y);
        // grids of sides up to 1024 have fewer than 2^31 cells, and their
        // sides and plane sizes fit 24 bits: full-rate 24-bit multiplies
        const int cell = (ox + lx) + (int)__umul24(nx, (oy + ly) + (int)__umul24(ny, oz + lz));
        rho[cell] = ldexpf((float)(long long)acc[i], -shift);
      }
    } else {
      const int s = slot[t];
      if (s >= 0) {
        for (int i = threadIdx.x; i < TILE_CELLS; i += DEP_THREADS) {
          unsigned long long v = acc[i];
          if (v) atomicAdd((unsigned long long*)&fx_accum[(int64_t)s * TILE_CELLS + i], v);
        }
      }
    }
    __syncthreads();
  }
}

__global__ void k_deposit_finalize(int nx, int ny, int ntx, int nty, long long* __restrict__ fx_accum,
                                   const int32_t* __restrict__ slot_tile,
                                   const int32_t* __restrict__ counters, int64_t max_slots,
                                   int shift, float* __restrict__ rho,
                                   const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int nslots = (int)min((int64_t)counters[CNT_SLOTS], max_slots);
  for (int s = blockIdx.x; s < nslots; s += gridDim.x) {
    const int t = slot_tile[s];
    const int ox = (t % ntx) * TILE_X, oy = ((t / ntx) % nty) * TILE_Y,
              oz = (t / (ntx * nty)) * TILE_Z;
    for (int i = threadIdx.x; i < TILE_CELLS; i += blockDim.x) {
      int lx = i % TILE_X, ly = (i / TILE_X) % TILE_Y, lz = i / (TILE_X * TILE_Y);
      long long* q = &fx_accum[(int64_t)s * TILE_CELLS + i];
      rho[(int64_t)(ox + lx) + (int64_t)nx * ((oy + ly) + (int64_t)ny * (oz + lz))] =
          ldexpf((float)*q, -shift);
      *q = 0;
    }
  }
}

}  // namespace

// Neighbour ranges for a scheme: (dx, dy, dz, class) of every neighbour tile
// class whose particles can reach the tile.  TSC reaches -1..+1 cells, CIC
// 0..+1 (so only the lower neighbour's upper face matters), NGP none.
std::vector<int4> deposit_range_table(int scheme) {
  std::vector<int4> out;
  int dlo = scheme == PMGPU_NGP ? 0 : -1;
  int dhi = scheme == PMGPU_TSC ? 1 : 0;
  for (int dz = dlo; dz <= dhi; ++dz)
    for (int dy = dlo; dy <= dhi; ++dy)
      for (int dx = dlo; dx <= dhi; ++dx) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        int d[3] = {dx, dy, dz};
        // classes compatible per axis: d=-1 -> upper face (2), d=+1 -> lower face (0), 0 -> any
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
          if (d[a] == -1) lo[a] = hi[a] = 2;
          else if (d[a] == 1) lo[a] = hi[a] = 0;
          else { lo[a] = 0; hi[a] = 2; }
        }
        for (int cz = lo[2]; cz <= hi[2]; ++cz)
          for (int cy = lo[1]; cy <= hi[1]; ++cy)
            for (int cx = lo[0]; cx <= hi[0]; ++cx)
              out.push_back(make_int4(dx, dy, dz, cx + 3 * cy + 9 * cz));
      }
  return out;
}

void deposit(pmgpu_ctx* c) {
  if (c->general) return deposit_general(c);
  const int nranges = OWN_RANGES + c->dep_table_len;
  int tok = stage_begin(c, "deposit_plan");
  zero_fill_i32(c, c->counters, 4);
  hipLaunchKernelGGL(k_deposit_plan, dim3((unsigned)(((int64_t)c->ntiles * 64 + 255) / 256)), dim3(256), 0, c->stream,
                     c->ntiles, c->ntx, c->nty, c->ntz, nranges, c->dep_table, c->cls_start,
                     c->tile_nchunk, c->int_nchunk, c->tile_slot, c->slot_tile, c->counters,
                     c->fx_slots, c->dep_chunk, c->int_chunk, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  exclusive_scan_i32(c, c->tile_nchunk, c->tile_chunk_start, c->ntiles);
  exclusive_scan_i32(c, c->int_nchunk, c->int_chunk_start, c->ntiles);
  expand_owner(c, c->tile_chunk_start, c->ntiles, c->dep_owner, c->iter);
  expand_owner(c, c->int_chunk_start, c->ntiles, c->int_owner, c->iter);
  stage_end(c, tok);

  tok = stage_begin(c, "deposit");
  int64_t bound = (int64_t)c->ntiles + 8 * c->n / c->dep_chunk + 1;
  // four times the interpolation's workgroups: each takes few work items, and
  // a workgroup that finishes early frees its slot for one not yet started,
  // which evens out the static assignment (0.61 ms against 0.66 at 1x on
  // 512^3; 16x and one item per workgroup are slower again)
  int grid = (int)std::min<int64_t>(bound, (int64_t)c->num_cu * c->blocks_per_cu * 4);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(DEP_THREADS), 0, c->stream, c->pm[c->cur], c->nx, c->ny, c->nz,
                       c->ntx, c->nty, c->ntz, c->ntiles, nranges, c->dep_table, c->cls_start,
                       c->tile_chunk_start, c->dep_owner, c->tile_slot, c->fx_accum, c->fx_shift, c->rho,
                       c->dep_chunk, c->tsc_centred, c->escape, c->iter);
  };
  if (c->prm.interpolation == PMGPU_TSC) go(k_deposit<PMGPU_TSC>);
  else if (c->prm.interpolation == PMGPU_CIC) go(k_deposit<PMGPU_CIC>);
  else go(k_deposit<PMGPU_NGP>);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_deposit_finalize, dim3(256), dim3(256), 0, c->stream, c->nx, c->ny, c->ntx,
                     c->nty, c->fx_accum, c->slot_tile, c->counters, c->fx_slots, c->fx_shift,
                     c->rho, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
}

}  // namespace pmgpu
