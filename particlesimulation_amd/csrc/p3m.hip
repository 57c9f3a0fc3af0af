// P3M short-range pass (P3MMethod::calculateShortRangeForces +
// correctAccelerations, source/p3mMethod.cpp:50-57, 168-322).
//
// The reference walks a chaining mesh of cells >= cutoff with linked lists,
// a 13-cell half shell and per-neighbour accumulation slots on CPU threads.
// Here:
//  - particles are counting-sorted by a short-range cell of sides >=
//    cutoff/div per axis, into a compact float4 (x, y, z, m) copy: cubes of
//    division div in y and z with x up to 4 times finer (div 1..6, chosen
//    from the cell count or tuned), or, when the division is not tuned, the
//    shape among a few anisotropic ones (finer across a thin particle set's
//    plane, one cutoff deep through it) that tests the fewest candidates
//    (sr_choose_cells); non-periodic, as the chaining mesh; out-of-box
//    particles clamp to the edge cells, which never loses a pair because
//    clamping is 1-Lipschitz;
//  - one work item = 64 consecutive i-particles of one cell, or of one
//    (y, z) row of cells (spanning x cells cx0..cx1, so sparse cells do not
//    leave lanes idle; tunable);
//    its j-stream is the contiguous (dz, dy) rows of the neighbourhood, each
//    trimmed to the x cells that can hold a pair within the cutoff (corner
//    rows drop out for div >= 3), their ranges (at most SR_ROWS_MAX)
//    computed in parallel into LDS.  The SR_WAVES
//    waves of a workgroup take interleaved 64-particle blocks of that stream
//    and read each j through the scalar unit (uniform address), so the inner
//    loop is pure FP32 VALU with scalar operands;
//  - the per-wave partial sums are added in LDS in fixed wave order; the pair
//    force comes from a table in LDS or from the direct formula.
// Newton's third law is not used: every ordered pair is evaluated once on
// the side of its i, which removes all write conflicts.  The mass m_i that
// the reference multiplies into the pair force and divides out again in
// correctAccelerations cancels and is never applied.
#include "pmgpu_internal.h"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <array>
#include <vector>

namespace pmgpu {

namespace {

constexpr int SR_THREADS = 512;
constexpr int SR_WAVES = SR_THREADS / 64;  // waves sharing a work item's j-stream
constexpr int SR_ICHUNK = 64;
constexpr int CNT_WORK = 16;  // ctx->counters slots of the per-XCD work counters (NXCD)
constexpr float kGc = 0.079577471545947668f;  // 1/(4 pi), the code-unit coupling

__device__ __forceinline__ int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

__device__ __forceinline__ int sr_cell_1d(float x, float inv_h, int m) {
  float f = floorf(x * inv_h);
  f = fminf(fmaxf(f, -1.0f), (float)m);  // NaN -> -1
  return clampi((int)f, 0, m - 1);
}

__global__ void k_sr_keys(const float4* __restrict__ pm, int64_t n, float ihx, float ihy, float ihz,
                          int mx, int my, int mz, int32_t* __restrict__ key,
                          int32_t* __restrict__ rank, int32_t* __restrict__ count,
                          const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < n;
  float4 p = active ? pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  int k = sr_cell_1d(p.x, ihx, mx) + mx * (sr_cell_1d(p.y, ihy, my) + my * sr_cell_1d(p.z, ihz, mz));
  const int32_t r = wave_rank(count, k, active);
  if (active) {
    key[i] = k;
    rank[i] = r;
  }
}

// Also records the sorted positions of the particles off the common mass m0
// (original ids ascending in odd_ids, nodd of them; nodd <= 0: none kept).
// The particle arrays are in the tile sort's order: the original id of the
// particle at i is the int in vel[i].w.  The sorted copy holds positions
// times pscale and masses times mscale (powers of two: exact).
__global__ void k_sr_scatter(const float4* __restrict__ pm, const float4* __restrict__ vel, int64_t n,
                             const int32_t* __restrict__ key,
                             const int32_t* __restrict__ rank, const int32_t* __restrict__ start,
                             float4* __restrict__ sr_pm, int32_t* __restrict__ sr_idx, float pscale,
                             float mscale, int clamp, float3 clo, float3 chi, float m0, int nodd,
                             const int32_t* __restrict__ odd_ids, int32_t* __restrict__ odd_pos,
                             const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t d = (int64_t)start[key[i]] + rank[i];
  float4 p = pm[i];
  if (clamp) {
    // for the whole table: positions in [clo, chi], which bounds every r^2
    p.x = fminf(fmaxf(p.x, clo.x), chi.x);
    p.y = fminf(fmaxf(p.y, clo.y), chi.y);
    p.z = fminf(fmaxf(p.z, clo.z), chi.z);
  }
  sr_pm[d] = make_float4(p.x * pscale, p.y * pscale, p.z * pscale, p.w * mscale);
  sr_idx[d] = (int32_t)i;
  if (nodd > 0 && __float_as_uint(p.w) != __float_as_uint(m0)) {
    const int32_t id = __float_as_int(vel[i].w);
    int lo = 0, hi = nodd - 1;  // odd_ids holds id
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (odd_ids[mid] < id) lo = mid + 1;
      else hi = mid;
    }
    // set_particles lists every particle off m0, so the search finds the id;
    // the guard keeps a mass changed behind the list's back from writing
    // another particle's slot
    if (odd_ids[lo] == id) odd_pos[lo] = (int32_t)d;
  }
}

// Work items per unit in chunks of SR_ICHUNK particles; a unit is a cell
// (stride 1) or a (y, z) row of cells (stride mx: its particles, x cells in
// order, so sparse cells share a chunk instead of idling lanes).
__global__ void k_sr_plan(const int32_t* __restrict__ start, int stride, int nunit, int32_t* __restrict__ nchunk,
                          const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nunit) return;
  const int count = start[(r + 1) * stride] - start[r * stride];
  nchunk[r] = (count + SR_ICHUNK - 1) / SR_ICHUNK;
}

// P3MMethod::referenceForceS1 / S2 (p3mMethod.cpp:194-218).
__device__ __forceinline__ float reference_force(int cs, float r, float a) {
  if (cs == PMGPU_S1) {
    if (r >= a) return kGc / (r * r);
    float q = r / a;
    return kGc / (a * a) * (8.0f * q - 9.0f * q * q + 2.0f * q * q * q * q);
  }
  float u = 2.0f * r / a;
  float u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u, u6 = u3 * u3;
  if (u <= 1.0f)
    return kGc / (35.0f * a * a) * (224.0f * u - 224.0f * u3 + 70.0f * u4 + 48.0f * u5 - 21.0f * u6);
  if (u <= 2.0f)
    return kGc / (35.0f * a * a) *
           (12.0f / u2 - 224.0f + 896.0f * u - 840.0f * u2 + 224.0f * u3 + 70.0f * u4 - 48.0f * u5 +
            7.0f * u6);
  return kGc / (r * r);
}

struct SrParams {
  int mx, my, mz, rowspan;
  int divy, divz;    // y and z reach in cells (cell sides >= cutoff / div)
  int divx;          // x reach in x cells
  float ihx;         // 1 / cell side in x, as the cell keys use it
  float hx, hy, hz;  // cell sides
  float rc2m;        // rc^2 (1 + 1e-4): row trimming bound with a rounding margin
  float inv_d2, rc2;
  float a, eps2;
  int cs;
  float tmax_f;         // last index of the reference table; entries past the cutoff are zero
  float log_scale;      // the accurate table's argument is s = r^2 log_scale (MODE 3)
  uint32_t log_lo, log_hi;  // float bits of its first bin and of its zero bin
};

// Pair factor F(r^2) such that a_i += m_j F (x_i - x_j).
//   MODE 0: shortRangeForce (p3mMethod.cpp:220-238), direct formula;
//   MODE 1: shortRangeForceFromTable (p3mMethod.cpp:240-245) on the
//           reference's 500-entry table, zero-padded past the cutoff so the
//           r < rc test is the table range itself;
//   MODE 2, 3: the accurate table of F itself over s = r^2 (MODE 2) or
//           s = r^2 / rc^2 (MODE 3), binned by the float bits of s:
//           bits >> 16 selects one of 128 bins per binade (log-spaced, so
//           the r^-3 and r^-1 singularities at small r are resolved to a
//           fixed relative accuracy), and F is linear in s within the bin:
//           F = B + A s, one FMA on the clamped s.  s is clamped to the
//           floats of bits [log_lo, log_hi], 31 to 32 binades: closer pairs
//           take the value at the start of the first bin, and the last bin
//           is zero (A = B = 0).
//   MODE 2 serves the unsoftened force with a <= rc, which vanishes for
//           r >= a, so the zero bin can start at the bin boundary at or
//           above a^2 with no scale.  Positions arrive scaled by 2^-q, which
//           puts the first bin at the smallest normal float: the entry of s
//           is (bits >> 16) - SR_LOG_ABS0, one shift, the constant folded
//           into the LDS address (the caller passes the table's start minus
//           SR_LOG_ABS0 entries).  The table is scaled by 2^-t (B) and
//           2^(2q - t) (A) to keep A finite there; masses arrive scaled by
//           2^(q + t), which undoes both: every product is the unscaled one
//           times a power of two, the same floats up to that factor.
//   MODE 3 puts the cutoff s = 1 on a bin boundary with one multiply (or
//           runs over r^2 itself, scale 1); its LDS copy holds bin k at
//           k mod SR_LOG_WRAP, a shift and a mask.
//   MODE 4 is MODE 2 without the clamp: the short-range copy's positions are
//           clamped to half a box outside the particle box, which bounds
//           r^2, and the table holds an entry for every float from 0 to that
//           bound (below the first bin the first bin's value at its start,
//           past the last bin zero), at entry bits >> 16: one shift and the
//           FMA (11 VALU instructions per pair).  Used when that is at most
//           SR_TABLE_EXT entries (a particle box up to about 73 diameters).
template <int MODE>
__device__ __forceinline__ float pair_factor(float r2, float mj, const float2* table,
                                             const SrParams& P, uint32_t log_hi_v) {
  if constexpr (MODE == 0) {
    const float r = sqrtf(r2);
    const float R = reference_force(P.cs, r, P.a);
    const float F = (-kGc / (r2 + P.eps2) + R) / r;
    return (r2 < P.rc2 && r2 > 0.0f) ? F * mj : 0.0f;
  } else if constexpr (MODE == 1) {
    // t = min(floor(ksi), tmax); a NaN r^2 also lands on the zero entry
    // tmax.  Past the cutoff every entry is (0, 0), so the fraction of an
    // out-of-range ksi is irrelevant.
    const float ksi = r2 * P.inv_d2;
    const int t = (int)fminf(ksi, P.tmax_f);
    const float2 tv = table[t];
    return fmaf(__builtin_amdgcn_fractf(ksi), tv.y, tv.x) * mj;
  } else if constexpr (MODE == 2) {
    // clamp in the float domain (s >= 0, where float order is bit order)
    const float sc = __builtin_amdgcn_fmed3f(r2, __uint_as_float(P.log_lo), __uint_as_float(log_hi_v));
    const float2 tv = table[__float_as_uint(sc) >> SR_LOG_SHIFT];
    return fmaf(tv.y, sc, tv.x) * mj;
  } else if constexpr (MODE == 4) {
    // the whole table: every r^2 the clamped positions allow has its entry
    const float2 tv = table[__float_as_uint(r2) >> SR_LOG_SHIFT];
    return fmaf(tv.y, r2, tv.x) * mj;
  } else {
    const float s = r2 * P.log_scale;
    // clamp in the float domain (s >= 0, where float order is bit order)
    const float sc = __builtin_amdgcn_fmed3f(s, __uint_as_float(P.log_lo), __uint_as_float(log_hi_v));
    const float2 tv = table[(__float_as_uint(sc) >> SR_LOG_SHIFT) & (SR_LOG_WRAP - 1)];
    return fmaf(tv.y, sc, tv.x) * mj;
  }
}

// Row (dy, dz) of the neighbourhood and its x half-width kx: cells whose
// smallest possible separation from the i cell, (|offset| - 1) * side per
// axis, is below the cutoff (with a rounding margin).  The bound holds for
// particles clamped into edge cells too (clamping only shrinks offsets), so
// trimming never drops a pair within the cutoff.
__device__ __forceinline__ bool row_extent(const SrParams& P, int dy, int dz, int* kx) {
  const float gy = (float)max(abs(dy) - 1, 0) * P.hy, gz = (float)max(abs(dz) - 1, 0) * P.hz;
  const float rem = P.rc2m - gy * gy - gz * gz;
  if (!(rem > 0.0f)) return false;
  int k = P.divx;
  while (k > 1) {
    const float gx = (float)(k - 1) * P.hx;
    if (gx * gx < rem) break;
    --k;
  }
  *kx = k;
  return true;
}

// UNIFORM: every j enters the pair loop at the common mass m0 (no multiply
// per pair); the sum is scaled by m0 once, and the nodd particles off m0
// (sorted positions in odd_pos) add their (m_j - m0) share per i.
template <int MODE, bool UNIFORM>
__global__ __launch_bounds__(SR_THREADS) void k_sr(
    const float4* __restrict__ sr_pm, const int32_t* __restrict__ sr_start,
    const int32_t* __restrict__ chunk_start, const int32_t* __restrict__ owner, int nunit,
    const int32_t* __restrict__ sr_idx,
    const float2* __restrict__ table_g, SrParams P, float4* __restrict__ pm,
    float4* __restrict__ vel, float4* __restrict__ acc, int mode, float kick, float3 box,
    float m0, int nodd, const int32_t* __restrict__ odd_pos,
    int32_t* __restrict__ counters, float4* __restrict__ acc_sr, int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  // The partial sums of half of the waves at a time.
  constexpr int HALF = SR_WAVES / 2;
  static_assert(SR_WAVES % 2 == 0, "the partial sums are added in halves");
  constexpr int TS = MODE == 4 ? SR_TABLE_EXT : (MODE >= 2 ? SR_TABLE_LOG : (MODE == 1 ? SR_TABLE : 1));
  // One LDS block, the table last: the absolute layout (MODE 2) reads it
  // through a pointer SR_LOG_ABS0 entries before its start, which stays
  // inside the block, so the entry offset folds into the LDS address.
  constexpr int PART_B = HALF * 64 * (int)sizeof(float3);
  // the j-ranges of the work item's (dz, dy) rows, filled in parallel
  static_assert(SR_ROWS_MAX <= SR_THREADS, "one thread per neighbourhood row");
  constexpr int ROWS_B = SR_ROWS_MAX * (int)sizeof(int2);
  constexpr int TAB_OFF = (PART_B + ROWS_B + (int)sizeof(int) + 15) / 16 * 16;
  static_assert(TAB_OFF >= SR_LOG_ABS0 * (int)sizeof(float2), "the absolute table's base stays in the block");
  __shared__ __align__(16) unsigned char lds[TAB_OFF + TS * sizeof(float2)];
  float3* part = reinterpret_cast<float3*>(lds);
  int2* rows = reinterpret_cast<int2*>(lds + PART_B);
  int& s_w = *reinterpret_cast<int*>(lds + PART_B + ROWS_B);
  float2* table = reinterpret_cast<float2*>(lds + TAB_OFF);
  const float2* table_rd = reinterpret_cast<const float2*>(
      lds + TAB_OFF - (MODE == 2 ? SR_LOG_ABS0 * (int)sizeof(float2) : 0));
  const int total = chunk_start[nunit];
  // the first work item before the table: workgroups that find none (small
  // particle sets leave most of the grid idle) exit without loading it
  if (threadIdx.x == 0) s_w = next_item(counters + CNT_WORK, total, blockIdx.x & (NXCD - 1));
  __syncthreads();
  if (s_w < 0) return;
  if constexpr (MODE == 1)
    for (int i = threadIdx.x; i < SR_TABLE; i += SR_THREADS) table[i] = table_g[i];
  if constexpr (MODE == 2)
    for (int i = threadIdx.x; i < SR_TABLE_LOG; i += SR_THREADS) table[i] = table_g[i];
  if constexpr (MODE == 4)
    for (int i = threadIdx.x; i < SR_TABLE_EXT; i += SR_THREADS) table[i] = table_g[i];
  if constexpr (MODE == 3) {
    const uint32_t k0 = P.log_lo >> SR_LOG_SHIFT;
    for (int i = threadIdx.x; i < SR_TABLE_LOG; i += SR_THREADS) table[(k0 + i) % SR_LOG_WRAP] = table_g[i];
  }
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // the upper clamp bound in a VGPR, so the clamp is one v_med3_f32 (GFX9
  // VALU instructions read at most one scalar operand)
  uint32_t log_hi_v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(log_hi_v) : "s"(P.log_hi));
  for (bool fetched = true;; fetched = false) {
    __syncthreads();
    if (!fetched && threadIdx.x == 0) s_w = next_item(counters + CNT_WORK, total, blockIdx.x & (NXCD - 1));
    __syncthreads();
    const int w = s_w;
    if (w < 0) break;
    // a work item: 64 consecutive particles of one unit -- a cell, or a
    // (y, z) row of cells whose particles may span x cells [cx0, cx1]
    // (sorted by x cell within the row)
    const int u = owner[w];
    const int k = w - chunk_start[u];
    int cy, cz, rs, re, cx0, cx1;
    if (P.rowspan) {
      cy = u % P.my;
      cz = u / P.my;
      rs = sr_start[u * P.mx];
      re = sr_start[(u + 1) * P.mx];
    } else {
      cx0 = cx1 = u % P.mx;
      cy = (u / P.mx) % P.my;
      cz = u / (P.mx * P.my);
      rs = sr_start[u];
      re = sr_start[u + 1];
    }
    const int i0 = rs + k * SR_ICHUNK;
    const int i = i0 + lane;
    const bool valid = i < re;
    if (P.rowspan) {
      cx0 = sr_cell_1d(sr_pm[i0].x, P.ihx, P.mx);
      cx1 = sr_cell_1d(sr_pm[min(i0 + SR_ICHUNK, re) - 1].x, P.ihx, P.mx);
    }
    const float4 pi = valid ? sr_pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    // row t = (dz + divz) (2 divy + 1) + dy + divy: its j-range, trimmed to
    // the x cells that can hold a pair (empty outside the cell grid)
    const int sidey = 2 * P.divy + 1, nrows = sidey * (2 * P.divz + 1);
    if (threadIdx.x < nrows) {
      const int dz = (int)threadIdx.x / sidey - P.divz, dy = (int)threadIdx.x % sidey - P.divy;
      const int zc = cz + dz, yc = cy + dy;
      int2 r = make_int2(0, 0);
      int kx;
      if (zc >= 0 && zc < P.mz && yc >= 0 && yc < P.my && row_extent(P, dy, dz, &kx)) {
        const int xlo = max(cx0 - kx, 0), xhi = min(cx1 + kx, P.mx - 1);
        const int row = P.mx * (yc + P.my * zc);
        r = make_int2(sr_start[row + xlo], sr_start[row + xhi + 1]);
      }
      rows[threadIdx.x] = r;
    }
    __syncthreads();
    float ax = 0.f, ay = 0.f, az = 0.f;
    int blk = 0;
    for (int t = 0; t < nrows; ++t) {
      {
        const int2 r = rows[t];
        const int js = __builtin_amdgcn_readfirstlane(r.x), je = __builtin_amdgcn_readfirstlane(r.y);
        for (int jb = js; jb < je; jb += 64, ++blk) {
          if (blk % SR_WAVES != wave) continue;
          const int jend = min(jb + 64, je);
          // whole groups of SR_JGROUP j; the last, partial group of a block
          // reads a whole group too, and its j past jend (particles of other
          // rows, or the zeroed pad after the last particle) enter with zero
          // mass, so they add exactly nothing
          int j0 = jb;
          for (; j0 + SR_JGROUP <= jend; j0 += SR_JGROUP) {
#pragma unroll
            for (int k = 0; k < SR_JGROUP; ++k) {
              const float4 pj = sr_pm[j0 + k];  // uniform address: scalar load
              const float dx = pi.x - pj.x, dy_ = pi.y - pj.y, dz_ = pi.z - pj.z;
              const float r2 = fmaf(dz_, dz_, fmaf(dy_, dy_, dx * dx));
              const float F = pair_factor<MODE>(r2, UNIFORM ? 1.0f : pj.w, table_rd, P, log_hi_v);
              ax = fmaf(F, dx, ax);
              ay = fmaf(F, dy_, ay);
              az = fmaf(F, dz_, az);
            }
          }
          if (j0 < jend) {
#pragma unroll
            for (int k = 0; k < SR_JGROUP; ++k) {
              const float4 pj = sr_pm[j0 + k];
              const float mj = j0 + k < jend ? (UNIFORM ? 1.0f : pj.w) : 0.0f;
              const float dx = pi.x - pj.x, dy_ = pi.y - pj.y, dz_ = pi.z - pj.z;
              const float r2 = fmaf(dz_, dz_, fmaf(dy_, dy_, dx * dx));
              const float F = pair_factor<MODE>(r2, mj, table_rd, P, log_hi_v);
              ax = fmaf(F, dx, ax);
              ay = fmaf(F, dy_, ay);
              az = fmaf(F, dz_, az);
            }
          }
        }
      }
    }
    // waves HALF.. hand their sums to waves 0..HALF-1, which hand theirs to
    // wave 0: a fixed order, whatever the timing
    if (wave >= HALF) part[threadIdx.x - HALF * 64] = make_float3(ax, ay, az);
    __syncthreads();
    if (wave < HALF) {
      const float3 o = part[threadIdx.x];
      ax += o.x;
      ay += o.y;
      az += o.z;
    }
    __syncthreads();
    if (wave > 0 && wave < HALF) part[threadIdx.x - 64] = make_float3(ax, ay, az);
    __syncthreads();
    if (wave == 0 && valid) {
      float3 s = make_float3(ax, ay, az);
      for (int q = 1; q < HALF; ++q) {
        const float3 o = part[lane + 64 * (q - 1)];
        s.x += o.x;
        s.y += o.y;
        s.z += o.z;
      }
      if constexpr (UNIFORM) {
        s.x *= m0;
        s.y *= m0;
        s.z *= m0;
        for (int q = 0; q < nodd; ++q) {
          const float4 pj = sr_pm[odd_pos[q]];
          const float dx = pi.x - pj.x, dy_ = pi.y - pj.y, dz_ = pi.z - pj.z;
          const float r2 = fmaf(dz_, dz_, fmaf(dy_, dy_, dx * dx));
          const float F = pair_factor<MODE>(r2, pj.w - m0, table_rd, P, log_hi_v);
          s.x = fmaf(F, dx, s.x);
          s.y = fmaf(F, dy_, s.y);
          s.z = fmaf(F, dz_, s.z);
        }
      }
      const int idx = sr_idx[i];
      if (acc_sr) acc_sr[idx] = make_float4(s.x, s.y, s.z, 0.0f);  // recorded on request
      float4 a = acc[idx];
      a.x += s.x;
      a.y += s.y;
      a.z += s.z;
      if (mode & MODE_ACC) acc[idx] = a;
      if (mode & (MODE_KICK | MODE_DRIFT)) {
        float4 v = vel[idx];
        float4 p = pm[idx];
        if (mode & MODE_KICK) {
          v.x += kick * a.x;
          v.y += kick * a.y;
          v.z += kick * a.z;
        }
        if (mode & MODE_DRIFT) {
          p.x += v.x;
          p.y += v.y;
          p.z += v.z;
          if (!inside_box(p.x, p.y, p.z, box)) atomicMin(escape, iteration + 1);
          pm[idx] = p;
        }
        vel[idx] = v;
      }
    }
  }
}

// Ordered pairs with 0 < r^2 < rc^2 (accepted), all (i, j) candidates the
// pass visits (tested) and the lane-slots it issues for them (issued: 64 per
// j, idle lanes of partly filled i chunks included), over k_sr's traversal.
// PAIRS = false counts tested and issued only (no pair loop), to compare
// cell geometries.
template <bool PAIRS>
__global__ __launch_bounds__(SR_THREADS) void k_sr_count(
    const float4* __restrict__ sr_pm, const int32_t* __restrict__ sr_start,
    const int32_t* __restrict__ chunk_start, const int32_t* __restrict__ owner, int nunit, SrParams P,
    int32_t* __restrict__ counters, unsigned long long* __restrict__ out) {
  __shared__ int s_w;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int total = chunk_start[nunit];
  unsigned long long accepted = 0, tested = 0, issued = 0;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_w = next_item(counters + CNT_WORK, total, blockIdx.x & (NXCD - 1));
    __syncthreads();
    const int w = s_w;
    if (w < 0) break;
    // a work item: 64 consecutive particles of one unit -- a cell, or a
    // (y, z) row of cells whose particles may span x cells [cx0, cx1]
    // (sorted by x cell within the row)
    const int u = owner[w];
    const int k = w - chunk_start[u];
    int cy, cz, rs, re, cx0, cx1;
    if (P.rowspan) {
      cy = u % P.my;
      cz = u / P.my;
      rs = sr_start[u * P.mx];
      re = sr_start[(u + 1) * P.mx];
    } else {
      cx0 = cx1 = u % P.mx;
      cy = (u / P.mx) % P.my;
      cz = u / (P.mx * P.my);
      rs = sr_start[u];
      re = sr_start[u + 1];
    }
    const int i0 = rs + k * SR_ICHUNK;
    const int i = i0 + lane;
    const bool valid = i < re;
    if (P.rowspan) {
      cx0 = sr_cell_1d(sr_pm[i0].x, P.ihx, P.mx);
      cx1 = sr_cell_1d(sr_pm[min(i0 + SR_ICHUNK, re) - 1].x, P.ihx, P.mx);
    }
    const float4 pi = valid ? sr_pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    int blk = 0, t = -1;
    for (int dz = -P.divz; dz <= P.divz; ++dz) {
      const int zc = cz + dz;
      for (int dy = -P.divy; dy <= P.divy; ++dy) {
        ++t;
        const int yc = cy + dy;
        if (zc < 0 || zc >= P.mz || yc < 0 || yc >= P.my) continue;
        int kx;
        if (!row_extent(P, dy, dz, &kx)) continue;
        const int xlo = max(cx0 - kx, 0), xhi = min(cx1 + kx, P.mx - 1);
        const int row = P.mx * (yc + P.my * zc);
        const int js = sr_start[row + xlo], je = sr_start[row + xhi + 1];
        if constexpr (!PAIRS) {
          // one wave per row: the whole j-range, once
          if (t % SR_WAVES == wave) {
            if (valid) tested += (unsigned long long)(je - js);
            issued += (unsigned long long)(je - js);
          }
          continue;
        }
        for (int jb = js; jb < je; jb += 64, ++blk) {
          if (blk % SR_WAVES != wave) continue;
          const int jend = min(jb + 64, je);
          if (valid) tested += (unsigned long long)(jend - jb);
          issued += (unsigned long long)(jend - jb);
          unsigned int in_block = 0;  // at most 64 per lane
#pragma unroll 8
          for (int j = jb; j < jend; ++j) {
            const float4 pj = sr_pm[j];
            const float dx = pi.x - pj.x, dy_ = pi.y - pj.y, dz_ = pi.z - pj.z;
            const float r2 = fmaf(dz_, dz_, fmaf(dy_, dy_, dx * dx));
            in_block += (r2 < P.rc2 && r2 > 0.0f) ? 1u : 0u;
          }
          if (valid) accepted += in_block;
        }
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    accepted += __shfl_down(accepted, o, 64);
    tested += __shfl_down(tested, o, 64);
    issued += __shfl_down(issued, o, 64);
  }
  if (lane == 0) {
    atomicAdd(&out[0], accepted);
    atomicAdd(&out[1], tested);
    atomicAdd(&out[2], issued);
  }
}

double ref_force_d(int cs, double r, double a) {
  const double G = 1.0 / (4.0 * M_PI);
  if (cs == PMGPU_S1) {
    if (r >= a) return G / (r * r);
    double q = r / a;
    return G / (a * a) * (8.0 * q - 9.0 * q * q + 2.0 * q * q * q * q);
  }
  double u = 2.0 * r / a, u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u, u6 = u3 * u3;
  if (u <= 1.0) return G / (35.0 * a * a) * (224.0 * u - 224.0 * u3 + 70.0 * u4 + 48.0 * u5 - 21.0 * u6);
  if (u <= 2.0)
    return G / (35.0 * a * a) *
           (12.0 / u2 - 224.0 + 896.0 * u - 840.0 * u2 + 224.0 * u3 + 70.0 * u4 - 48.0 * u5 + 7.0 * u6);
  return G / (r * r);
}

}  // namespace

// P3MMethod::initSRForceTable (p3mMethod.cpp:275-294) in FP32 on the host,
// laid out as (T[t], T[t+1]-T[t]) pairs, zero past the last interval.
void sr_table_host(float rc, float a, float eps, int cs, std::vector<float>& table, float& d2) {
  const int cnt = SR_TABLE_REF;
  d2 = rc * rc / (float)(cnt - 1);
  const float G = 1.0f / (4.0f * 3.14159265358979323846f);
  std::vector<float> T(cnt);
  for (int i = 0; i < cnt; ++i) {
    float r2 = (float)i * d2;
    float r = std::sqrt(r2);
    float R;
    if (cs == PMGPU_S1) {
      if (r >= a) R = G / (r * r);
      else {
        float q = r / a;
        R = G / (a * a) * (8.0f * r / a - 9.0f * r * r / (a * a) + 2.0f * q * q * q * q);
      }
    } else {
      float u = 2.0f * r / a;
      float u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u, u6 = u3 * u3;
      if (u <= 1.0f)
        R = G / (35.0f * a * a) * (224.0f * u - 224.0f * u3 + 70.0f * u4 + 48.0f * u5 - 21.0f * u6);
      else if (u <= 2.0f)
        R = G / (35.0f * a * a) *
            (12.0f / u2 - 224.0f + 896.0f * u - 840.0f * u2 + 224.0f * u3 + 70.0f * u4 - 48.0f * u5 +
             7.0f * u6);
      else R = G / (r * r);
    }
    R = -R;
    float total = -G / (r * r + eps * eps);
    T[i] = r == 0.0f ? 0.0f : (total - R) / std::sqrt(r * r + eps * eps);
  }
  table.assign(2 * SR_TABLE, 0.0f);
  for (int i = 0; i + 1 < cnt; ++i) {
    table[2 * i] = T[i];
    table[2 * i + 1] = T[i + 1] - T[i];
  }
}

// The accurate table of the direct formula's pair factor
// F(s) = (-G/(r^2 + eps^2) + R(r)) / r (shortRangeForce,
// p3mMethod.cpp:220-238) over s = r^2 scale, binned by float bits: bin i of
// nbins covers the floats s whose bits >> 16 equal (lo >> 16) + i, [s0, s1),
// and holds the chord of F through the bin's ends as (B, A), F = B + A s,
// lowered by half the sum of its largest deviations above and below F (so
// the error is centred); entry nbins is zero.  Built in FP64, rounded to FP32.
static void sr_table_log_bins(float a, float eps, int cs, float scale, uint32_t lo, int nbins,
                              std::vector<float>& table) {
  auto from_bits = [](uint32_t b) {
    float f;
    std::memcpy(&f, &b, sizeof f);
    return (double)f;
  };
  const uint32_t bin = 1u << SR_LOG_SHIFT;
  const double c = (double)scale, G = 1.0 / (4.0 * M_PI), e2 = (double)eps * eps;
  auto F = [&](double sp) {
    const double s = sp / c, r = std::sqrt(s);
    return (-G / (s + e2) + ref_force_d(cs, r, a)) / r;
  };
  table.assign(2 * (nbins + 1), 0.0f);
  for (int i = 0; i < nbins; ++i) {
    const uint32_t b0 = lo + (uint32_t)i * bin;
    const double s0 = from_bits(b0), s1 = from_bits(b0 + bin);
    const double f0 = F(s0), f1 = F(s1);
    double dlo = 0.0, dhi = 0.0;
    for (int t = 1; t < 32; ++t) {
      const double u = t / 32.0, e = f0 + u * (f1 - f0) - F(s0 + u * (s1 - s0));
      dlo = std::min(dlo, e);
      dhi = std::max(dhi, e);
    }
    const double A = (f1 - f0) / (s1 - s0);
    table[2 * i] = (float)(f0 - A * s0 - 0.5 * (dlo + dhi));
    table[2 * i + 1] = (float)A;
  }
}

// The accurate table in the layout the pair loop reads.
//  - Unsoftened with a <= rc (F = 0 for r >= a), over r^2 itself: the zero
//    bin at the bin boundary at or above a^2 (hi), the first bin at the start
//    of the binade 30 or 31 below hi's (lo, an odd exponent field), so with
//    positions scaled by 2^-q the first bin starts at the smallest normal
//    float, entry SR_LOG_ABS0 (MODE 2, absolute layout).  B is scaled by 2^-t
//    and A by 2^(2q - t), t putting the largest |A| in [2^125, 2^126) (A
//    grows as r^-5 towards lo), and masses by 2^(q + t), which undoes both;
//    all scalings are powers of two.  Where the scaled masses could overflow
//    (sum |m| 2^(q + t) >= 2^120) the same range stays unscaled and wrapped
//    (MODE 3, scale 1).
//    The scaled table comes whole (MODE 4) when every r^2 of the clamped
//    positions fits SR_TABLE_EXT entries, else clamped (MODE 2).
//  - Otherwise over r^2 / rc^2, the zero bin at s = 1 (the cutoff) and
//    SR_TABLE_LOG - 1 bins below it, wrapped: bin k at k mod SR_LOG_WRAP
//    (MODE 3).
static void sr_table_log_host(pmgpu_ctx* c, std::vector<float>& table) {
  const float rc = c->p3m.cutoff, a = c->p3m.particle_diameter, eps = c->p3m.softening;
  const int cs = c->p3m.cloud_shape;
  const uint32_t bin = 1u << SR_LOG_SHIFT;
  c->sr_pos_scale = c->sr_mass_scale = 1.0f;
  c->sr_log_mode = 3;
  if (!(eps == 0.0f && a <= rc && a >= 1.0e-6f)) {
    c->sr_log_scale = 1.0f / (rc * rc);
    c->sr_log_hi = 0x3f800000u;  // 1.0f
    c->sr_log_lo = c->sr_log_hi - (uint32_t)(SR_TABLE_LOG - 1) * bin;
    sr_table_log_bins(a, eps, cs, c->sr_log_scale, c->sr_log_lo, SR_TABLE_LOG - 1, table);
    return;
  }
  uint32_t hi;
  const float a2 = a * a;
  std::memcpy(&hi, &a2, sizeof hi);
  hi = (hi + bin - 1) / bin * bin;
  const int e_hi = (int)(hi >> 23);
  const int e_lo = (e_hi - 31) % 2 != 0 ? e_hi - 31 : e_hi - 30;
  const uint32_t lo = (uint32_t)e_lo << 23;
  const int nbins = (int)((hi - lo) / bin);  // <= 31 * 128 + 127
  const int q = (e_lo - 1) / 2;
  c->sr_log_scale = 1.0f;
  std::vector<float> bins;
  sr_table_log_bins(a, eps, cs, 1.0f, lo, nbins, bins);
  float amax = 0.0f;
  for (int i = 0; i < nbins; ++i) amax = std::max(amax, std::fabs(bins[2 * i + 1]));
  const int t = std::isfinite(amax) && amax > 0.0f ? std::ilogb(amax) + 2 * q - 125 : 0;
  const bool scaled = e_lo >= 3 && nbins + 1 <= SR_TABLE_LOG && std::isfinite(amax) && amax > 0.0f &&
                      q + t <= 126 && q + t >= -126 && std::ldexp(c->mass_abs_sum, q + t) < 0x1p120;
  if (!scaled) {
    // wrapped: the same bins from lo, the zero bin repeated up to SR_TABLE_LOG
    c->sr_log_lo = lo;
    c->sr_log_hi = hi;
    table = bins;
    table.resize(2 * SR_TABLE_LOG, 0.0f);
    return;
  }
  c->sr_log_lo = lo - ((uint32_t)(2 * q) << 23);  // 2^-126
  c->sr_log_hi = hi - ((uint32_t)(2 * q) << 23);
  c->sr_pos_scale = std::ldexp(1.0f, -q);
  c->sr_mass_scale = std::ldexp(1.0f, q + t);
  // The whole table (MODE 4) when it fits: the short-range copy's positions
  // are clamped to [-B/2, 3B/2] per axis (B the particle box), so r^2 (scaled)
  // stays below (1 + 1e-5) sum (2 B s)^2, FP32 rounding of the sum included;
  // every float up to that bound has its entry: below the first bin (r^2 <
  // 2^-126, zero included) the first bin's value at its start, which the
  // clamp gave, past the last bin zero.
  double bound = 0.0;
  bool boxed = true;
  for (int i = 0; i < 3; ++i) {
    boxed = boxed && c->p3m.box[i] > 0.0f;
    const double side = 2.0 * (double)c->p3m.box[i] * (double)c->sr_pos_scale;
    bound += side * side;
  }
  const float bound_f = (float)(bound * (1.0 + 1e-5));
  uint32_t bound_bits;
  std::memcpy(&bound_bits, &bound_f, sizeof bound_bits);
  const int whole = (int)(bound_bits >> SR_LOG_SHIFT) + 2;
  if (boxed && std::isfinite(bound_f) && whole <= SR_TABLE_EXT && SR_LOG_ABS0 + nbins < whole) {
    table.assign(2 * SR_TABLE_EXT, 0.0f);
    const float b0 = std::ldexp(bins[0], -t), a0 = std::ldexp(bins[1], 2 * q - t);
    const float first = fmaf(a0, 0x1p-126f, b0);  // the first bin at its start, as the clamp gave it
    for (int i = 0; i < SR_LOG_ABS0; ++i) table[2 * i] = first;
    for (int i = 0; i < nbins; ++i) {
      table[2 * (SR_LOG_ABS0 + i)] = std::ldexp(bins[2 * i], -t);
      table[2 * (SR_LOG_ABS0 + i) + 1] = std::ldexp(bins[2 * i + 1], 2 * q - t);
    }
    c->sr_log_mode = 4;
    return;
  }
  table.assign(2 * SR_TABLE_LOG, 0.0f);
  for (int i = 0; i < nbins; ++i) {
    table[2 * i] = std::ldexp(bins[2 * i], -t);
    table[2 * i + 1] = std::ldexp(bins[2 * i + 1], 2 * q - t);
  }
  c->sr_log_mode = 2;
}

void p3m_setup(pmgpu_ctx* c) {
  std::vector<float> tab;
  float d2 = 0.0f;
  if (c->p3m.use_table == 2) {
    sr_table_log_host(c, tab);
    d2 = 1.0f;
    c->sr_tmax = SR_TABLE_LOG - 1;
  } else {
    sr_table_host(c->p3m.cutoff, c->p3m.particle_diameter, c->p3m.softening, c->p3m.cloud_shape,
                  tab, d2);
    c->sr_tmax = SR_TABLE - 1;
    c->sr_pos_scale = c->sr_mass_scale = 1.0f;
  }
  HIP_CHECK(hipMemcpy(c->sr_table, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
  c->sr_inv_d2 = (float)(1.0 / (double)d2);
}

// Row-span work items (knobs[5] = 2, or automatic) or one cell per unit (1).
static bool sr_rowspan(const pmgpu_ctx* c) { return c->sr_chunking != 1; }
static int sr_units(const pmgpu_ctx* c) { return sr_rowspan(c) ? c->mcy * c->mcz : c->ncell_sr; }

static SrParams sr_params(pmgpu_ctx* c) {
  SrParams P;
  P.mx = c->mcx;
  P.my = c->mcy;
  P.mz = c->mcz;
  P.divy = c->sr_divy;
  P.divz = c->sr_divz;
  P.divx = c->sr_divx_base * c->sr_xfine;
  P.rowspan = sr_rowspan(c);
  // the sorted copy's positions are scaled by sr_pos_scale (a power of two):
  // its cell index and cutoff test scale with it, exactly
  P.ihx = (1.0f / c->hcx) / c->sr_pos_scale;
  P.hx = c->hcx;
  P.hy = c->hcy;
  P.hz = c->hcz;
  P.rc2m = c->p3m.cutoff * c->p3m.cutoff * (1.0f + 1e-4f);
  P.inv_d2 = c->sr_inv_d2;
  P.rc2 = c->p3m.cutoff * c->p3m.cutoff * (c->sr_pos_scale * c->sr_pos_scale);
  P.a = c->p3m.particle_diameter;
  P.eps2 = c->p3m.softening * c->p3m.softening;
  P.cs = c->p3m.cloud_shape;
  P.tmax_f = (float)c->sr_tmax;
  P.log_scale = c->sr_log_scale;
  P.log_lo = c->sr_log_lo;
  P.log_hi = c->sr_log_hi;
  return P;
}

static void sr_sort(pmgpu_ctx* c, int iteration) {
  int tok = stage_begin(c, "sr_sort");
  zero_fill_i32(c, c->sr_count, c->ncell_sr);
  unsigned g = (unsigned)((c->n + 255) / 256);
  hipLaunchKernelGGL(k_sr_keys, dim3(g), dim3(256), 0, c->stream, c->pm[c->cur], c->n,
                     1.0f / c->hcx, 1.0f / c->hcy, 1.0f / c->hcz, c->mcx, c->mcy, c->mcz, c->sr_key,
                     c->sr_rank, c->sr_count, c->escape, iteration);
  HIP_CHECK(hipGetLastError());
  exclusive_scan_i32(c, c->sr_count, c->sr_start, c->ncell_sr);
  hipLaunchKernelGGL(k_sr_scatter, dim3(g), dim3(256), 0, c->stream, c->pm[c->cur], c->vel[c->cur], c->n,
                     c->sr_key,
                     c->sr_rank, c->sr_start, c->sr_pm, c->sr_idx, c->sr_pos_scale, c->sr_mass_scale,
                     (int)(c->sr_log_mode == 4 && c->p3m.use_table == 2),
                     make_float3(-0.5f * c->p3m.box[0], -0.5f * c->p3m.box[1], -0.5f * c->p3m.box[2]),
                     make_float3(1.5f * c->p3m.box[0], 1.5f * c->p3m.box[1], 1.5f * c->p3m.box[2]),
                     c->sr_m0, c->sr_nodd, c->sr_odd_ids,
                     c->sr_odd_pos, c->escape, iteration);
  HIP_CHECK(hipGetLastError());
  const int nunit = sr_units(c), stride = sr_rowspan(c) ? c->mcx : 1;
  hipLaunchKernelGGL(k_sr_plan, dim3((nunit + 255) / 256), dim3(256), 0, c->stream, c->sr_start, stride, nunit,
                     c->sr_nchunk, c->escape, iteration);
  HIP_CHECK(hipGetLastError());
  exclusive_scan_i32(c, c->sr_nchunk, c->sr_chunk_start, nunit);
  expand_owner(c, c->sr_chunk_start, nunit, c->sr_owner, iteration);
  zero_fill_i32(c, c->counters + CNT_WORK, NXCD);
  stage_end(c, tok);
}

template <int MODE, bool UNIFORM>
static void launch_sr_mass(pmgpu_ctx* c, int mode, float kick, int iteration) {
  const int grid = c->num_cu * 8;
  hipLaunchKernelGGL((k_sr<MODE, UNIFORM>), dim3(grid), dim3(SR_THREADS), 0, c->stream, c->sr_pm,
                     c->sr_start, c->sr_chunk_start, c->sr_owner, sr_units(c), c->sr_idx, c->sr_table,
                     sr_params(c), c->pm[c->cur], c->vel[c->cur], c->acc, mode, kick,
                     make_float3(c->prm.box[0], c->prm.box[1], c->prm.box[2]), c->sr_m0 * c->sr_mass_scale,
                     std::max(c->sr_nodd, 0), c->sr_odd_pos, c->counters, c->acc_sr, c->escape, iteration);
  HIP_CHECK(hipGetLastError());
}

template <int MODE>
static void launch_sr(pmgpu_ctx* c, int mode, float kick, int iteration) {
  if ((2 * c->sr_divy + 1) * (2 * c->sr_divz + 1) > SR_ROWS_MAX)
    fail(PMGPU_ERR_INVALID_ARGUMENT, "short-range neighbourhood has too many rows");
  if (c->sr_nodd >= 0) launch_sr_mass<MODE, true>(c, mode, kick, iteration);
  else launch_sr_mass<MODE, false>(c, mode, kick, iteration);
}

void p3m_sort(pmgpu_ctx* c, int iteration) { sr_sort(c, iteration); }

void p3m_short_range(pmgpu_ctx* c, int mode, float kick, int iteration) {
  sr_sort(c, iteration);
  p3m_pairs(c, mode, kick, iteration);
}

void p3m_pairs(pmgpu_ctx* c, int mode, float kick, int iteration) {
  int tok = stage_begin(c, "sr_pairs");
  switch (c->p3m.use_table) {
    case 0: launch_sr<0>(c, mode, kick, iteration); break;
    case 1: launch_sr<1>(c, mode, kick, iteration); break;
    default:
      if (c->sr_log_mode == 4) launch_sr<4>(c, mode, kick, iteration);
      else if (c->sr_log_mode == 2) launch_sr<2>(c, mode, kick, iteration);
      else launch_sr<3>(c, mode, kick, iteration);
  }
  stage_end(c, tok);
}

static void sr_count(pmgpu_ctx* c, bool pairs, unsigned long long h[3]) {
  const int it = -2147483647 - 1;
  sr_sort(c, it);
  zero_fill_i32(c, (int32_t*)c->pair_count, 6);
  if (pairs)
    hipLaunchKernelGGL(k_sr_count<true>, dim3(c->num_cu * 8), dim3(SR_THREADS), 0, c->stream, c->sr_pm,
                       c->sr_start, c->sr_chunk_start, c->sr_owner, sr_units(c), sr_params(c), c->counters,
                       c->pair_count);
  else
    hipLaunchKernelGGL(k_sr_count<false>, dim3(c->num_cu * 8), dim3(SR_THREADS), 0, c->stream, c->sr_pm,
                       c->sr_start, c->sr_chunk_start, c->sr_owner, sr_units(c), sr_params(c), c->counters,
                       c->pair_count);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(h, c->pair_count, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

void p3m_count_pairs(pmgpu_ctx* c, int64_t* accepted, int64_t* tested, int64_t* issued) {
  unsigned long long h[3] = {0, 0, 0};
  sr_count(c, true, h);
  *accepted = (int64_t)h[0];
  if (tested) *tested = (int64_t)h[1];
  if (issued) *issued = (int64_t)h[2];
}

// Anisotropic cells for anisotropic particle sets: in a thin disk, cells one
// cutoff deep across the disk and finer in the plane visit fewer candidates
// than cubes (the candidates of a pair are the cells within the cutoff of
// the i cell, so the excess over the cutoff sphere grows with the cell sides
// along which the particles extend).  Each geometry -- the isotropic
// division d, y or z cells 2d, 3d or 4d times finer with the other one
// cutoff deep, or 2d times finer with the other half a cutoff deep, each with
// x cells 4 or 8 times finer than d (as the cell budget allows) -- is sorted
// and scored by the candidates it tests plus 16 per cell (the per-step
// sort's cost per cell, in pair-evaluation times); the lowest score wins.
// Every geometry visits the same pair set.
void sr_choose_cells(pmgpu_ctx* c) {
  const int d = c->sr_div;
  std::vector<std::array<int, 3>> cand = {{d, d, 4}};
  for (int xf : {4, 8}) {
    if (xf == 8) cand.push_back({d, d, 8});
    for (int k : {2, 3, 4}) {
      cand.push_back({k * d, 1, xf});
      cand.push_back({1, k * d, xf});
    }
    cand.push_back({2 * d, 2, xf});
    cand.push_back({2, 2 * d, xf});
  }
  // cells at x refinement 1, the fewest sr_cells_setup can fall back to
  const double rc = (double)c->p3m.cutoff * (1.0 + 1e-5);
  auto cells_min = [&](int y, int z) {
    return std::max(1.0, std::floor((double)c->p3m.box[0] / (rc / d))) *
           std::max(1.0, std::floor((double)c->p3m.box[1] / (rc / y))) *
           std::max(1.0, std::floor((double)c->p3m.box[2] / (rc / z)));
  };
  std::vector<std::array<int, 3>> seen;  // (y, z, x) divisions scored
  std::array<int, 3> best = cand[0];
  double best_score = -1.0;
  for (const auto& g : cand) {
    if ((2 * g[0] + 1) * (2 * g[1] + 1) > SR_ROWS_MAX) continue;
    if (&g != &cand[0] && cells_min(g[0], g[1]) > (double)SR_CELLS_MAX) continue;
    c->sr_divy = g[0];
    c->sr_divz = g[1];
    c->sr_divx_base = d;
    c->sr_xfine_max = g[2];
    sr_cells_setup(c);
    const std::array<int, 3> got = {g[0], g[1], d * c->sr_xfine};
    if (std::find(seen.begin(), seen.end(), got) != seen.end()) continue;
    seen.push_back(got);
    unsigned long long h[3] = {0, 0, 0};
    sr_count(c, false, h);
    const double score = (double)h[1] + 16.0 * (double)c->ncell_sr;
    if (best_score < 0.0 || score < best_score) {
      best = g;
      best_score = score;
    }
  }
  c->sr_divy = best[0];
  c->sr_divz = best[1];
  c->sr_divx_base = d;
  c->sr_xfine_max = best[2];
  sr_cells_setup(c);
}

}  // namespace pmgpu
