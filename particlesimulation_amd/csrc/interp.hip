// Field and interpolation (PMMethod::findFieldInCells + updateAccelerations,
// source/pmMethod.cpp:279-338, 352-390) fused with the leapfrog kick and the
// next drift (source/leapfrog.cpp:16-24).
//
// One work item = one tile's own particles (a contiguous range of the
// tile-sorted arrays, cut into int_chunk pieces, tunable).  The work item stages
// the potential of the tile plus the stencil and finite-difference halo in
// LDS, forms E = -grad(phi) on the stencil region in LDS (2- or 4-point,
// pmMethod.cpp:355-365), and interpolates with the deposit's weights.  The
// HBM traffic per step is the potential once (halo re-reads hit L2) plus the
// particle state once.
#include "pmgpu_internal.h"

namespace pmgpu {

namespace {

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

struct ExtField {
  int kind;
  float cx, cy, cz, R, M, G, H, DT;
};

// sphRadDecrField (source/externalFields.cpp:4-16) at pos*H, converted with
// accelerationToCodeUnits (include/unitConversions.h:22-24).
__device__ __forceinline__ float3 ext_field(const ExtField& e, float x, float y, float z) {
  if (e.kind != PMGPU_EXT_SPH_RAD_DECR) return make_float3(0.f, 0.f, 0.f);
  float dx = x * e.H - e.cx, dy = y * e.H - e.cy, dz = z * e.H - e.cz;
  float r = sqrtf(dx * dx + dy * dy + dz * dz);
  float g = r > e.R ? -e.G * e.M / (r * r) : -(e.G * e.M / (e.R * e.R * e.R)) * r * (4.0f - 3.0f * r / e.R);
  float s = e.DT * e.DT * g / (r * e.H);
  return make_float3(s * dx, s * dy, s * dz);
}

__device__ __forceinline__ void leapfrog_tail(float4& p, float4& v, float ax, float ay, float az,
                                              int mode, float kick, float3 box,
                                              int32_t* __restrict__ escape, int iteration) {
  if (mode & MODE_KICK) {
    v.x += kick * ax;
    v.y += kick * ay;
    v.z += kick * az;
  }
  if (mode & MODE_DRIFT) {
    p.x += v.x;
    p.y += v.y;
    p.z += v.z;
    bool inside = inside_box(p.x, p.y, p.z, box);
    if (!inside) atomicMin(escape, iteration + 1);
  }
}

template <int SCHEME, int FDS>
struct Geo {
  static constexpr int SLO = SCHEME == PMGPU_TSC ? -1 : 0;
  static constexpr int SHI = SCHEME == PMGPU_NGP ? 0 : 1;
  static constexpr int CNT = SHI - SLO + 1;
  static constexpr int F = FDS == PMGPU_TWO_POINT ? 1 : 2;
  static constexpr int EX = TILE_X + CNT - 1, EY = TILE_Y + CNT - 1, EZ = TILE_Z + CNT - 1;
  static constexpr int PX = EX + 2 * F, PY = EY + 2 * F, PZ = EZ + 2 * F;
  // The staged potential's x rows start at the 4-aligned cell ox - 4 (the
  // region's first cell ox + SLO - F is 1 to 3 cells later: XOFF) and hold
  // NCH aligned float4 chunks, so a row is NCH 16-byte loads; a chunk wraps
  // as a unit because the grid's x side is a multiple of 4.
  static constexpr int XOFF = SLO - F + 4;
  static constexpr int NCH = (XOFF + PX + 3) / 4;
  static constexpr int SPX = 4 * NCH;
};

template <int SCHEME>
__device__ __forceinline__ void interp_weights(float d, float w[3]) {
  if constexpr (SCHEME == PMGPU_TSC) {
    // TSCAssignmentFunc (pmMethod.cpp:187-198); the 0.125 is applied once
    w[0] = (0.5f - d) * (0.5f - d);
    w[1] = 1.5f - 2.0f * d * d;
    w[2] = (0.5f + d) * (0.5f + d);
  } else if constexpr (SCHEME == PMGPU_CIC) {
    w[0] = 1.0f - d;
    w[1] = d;
  } else {
    w[0] = 1.0f;
  }
}

// At most 128 VGPRs: four 256-thread workgroups per CU, as many as the
// LDS block of the largest stencil allows.
template <int SCHEME, int FDS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_interp(
    const float* __restrict__ phi, int nx, int ny, int nz, int ntx, int nty, int ntiles,
    const int32_t* __restrict__ cls_start, const int32_t* __restrict__ chunk_start,
    const int32_t* __restrict__ owner, float4* __restrict__ pm, float4* __restrict__ vel, float4* __restrict__ acc, int mode,
    float kick, float3 box, ExtField ext, int int_chunk, int centred, int32_t* __restrict__ escape,
    int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  using Gm = Geo<SCHEME, FDS>;
  // The potential and the field share one LDS block: each thread forms its
  // share of E in registers, and E overwrites the potential after a barrier,
  // which halves the block and doubles the workgroups a CU holds.
  constexpr int NP = Gm::SPX * Gm::PY * Gm::PZ, NE = Gm::EX * Gm::EY * Gm::EZ;
  constexpr int KE = (NE + 255) / 256;  // field points per thread
  __shared__ __align__(16) float lds[NP > 3 * NE ? NP : 3 * NE];
  float* const sphi = lds;
  float* const sE0 = lds;
  float* const sE1 = lds + NE;
  float* const sE2 = lds + 2 * NE;
  __shared__ int s_tile, s_chunk;
  const int total = chunk_start[ntiles];
  const XcdRange wr = xcd_range(total);
  for (int w = wr.first; w < wr.end; w += wr.stride) {
    if (threadIdx.x == 0) {
      const int t = owner[w];
      s_tile = t;
      s_chunk = w - chunk_start[t];
    }
    __syncthreads();
    const int t = s_tile;
    const int ox = (t % ntx) * TILE_X, oy = ((t / ntx) % nty) * TILE_Y,
              oz = (t / (ntx * nty)) * TILE_Z;
    const int py0 = oy + Gm::SLO - Gm::F, pz0 = oz + Gm::SLO - Gm::F;
    // stage the potential in float4 chunks of x rows (NCH per row)
    constexpr int NF = Gm::NCH * Gm::PY * Gm::PZ, UNR = 4;
    for (int f0 = threadIdx.x; f0 < NF; f0 += 256 * UNR) {
      float4 v[UNR];
#pragma unroll
      for (int q = 0; q < UNR; ++q) {
        const int f = f0 + 256 * q;
        const int cx = f % Gm::NCH, row = f / Gm::NCH;
        const int iy = row % Gm::PY, iz = row / Gm::PY;
        const uint32_t gx = wrapc(ox - 4 + 4 * cx, nx), gy = wrapc(py0 + iy, ny), gz = wrapc(pz0 + iz, nz);
        v[q] = f < NF ? *reinterpret_cast<const float4*>(phi + gx + (uint32_t)nx * (gy + (uint32_t)ny * gz))
                      : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int q = 0; q < UNR; ++q) {
        const int f = f0 + 256 * q;
        if (f < NF) *reinterpret_cast<float4*>(sphi + 4 * f) = v[q];
      }
    }
    __syncthreads();
    float rEx[KE], rEy[KE], rEz[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int e = (int)threadIdx.x + 256 * k;
      if (e < NE) {
        const int ex = e % Gm::EX, ey = (e / Gm::EX) % Gm::EY, ez = e / (Gm::EX * Gm::EY);
        const int c = (ex + Gm::F + Gm::XOFF) + Gm::SPX * ((ey + Gm::F) + Gm::PY * (ez + Gm::F));
        const int sx = 1, sy = Gm::SPX, sz = Gm::SPX * Gm::PY;
        if constexpr (FDS == PMGPU_TWO_POINT) {
          rEx[k] = -0.5f * (sphi[c + sx] - sphi[c - sx]);
          rEy[k] = -0.5f * (sphi[c + sy] - sphi[c - sy]);
          rEz[k] = -0.5f * (sphi[c + sz] - sphi[c - sz]);
        } else {
          rEx[k] = (-1.0f / 12) * (-sphi[c + 2 * sx] + 8 * sphi[c + sx] - 8 * sphi[c - sx] + sphi[c - 2 * sx]);
          rEy[k] = (-1.0f / 12) * (-sphi[c + 2 * sy] + 8 * sphi[c + sy] - 8 * sphi[c - sy] + sphi[c - 2 * sy]);
          rEz[k] = (-1.0f / 12) * (-sphi[c + 2 * sz] + 8 * sphi[c + sz] - 8 * sphi[c - sz] + sphi[c - 2 * sz]);
        }
      }
    }
    __syncthreads();  // every read of the potential is done: E may overwrite it
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int e = (int)threadIdx.x + 256 * k;
      if (e < NE) {
        sE0[e] = rEx[k];
        sE1[e] = rEy[k];
        sE2[e] = rEz[k];
      }
    }
    __syncthreads();
    const int pbeg = cls_start[t * NCLASS] + s_chunk * int_chunk;
    const int pend = min(pbeg + int_chunk, cls_start[t * NCLASS + NCLASS]);
    for (int i = pbeg + (int)threadIdx.x; i < pend; i += 256) {
      float4 p = pm[i];
      const int ax = anchor_of(p.x, SCHEME, centred), ay = anchor_of(p.y, SCHEME, centred), az = anchor_of(p.z, SCHEME, centred);
      float wx[3], wy[3], wz[3];
      interp_weights<SCHEME>(p.x - (float)ax, wx);
      interp_weights<SCHEME>(p.y - (float)ay, wy);
      interp_weights<SCHEME>(p.z - (float)az, wz);
      const int lx = wrapc(ax, nx) - ox, ly = wrapc(ay, ny) - oy, lz = wrapc(az, nz) - oz;
      const float3 g = ext_field(ext, p.x, p.y, p.z);
      float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll 1
      for (int a = 0; a < Gm::CNT; ++a) {
        // the TSC's 1/8 once per x weight, not per tap (a power of two: the
        // same products)
        const float t1 = SCHEME == PMGPU_TSC ? wx[a] * 0.125f : wx[a];
#pragma unroll
        for (int b = 0; b < Gm::CNT; ++b) {
          const float t2 = t1 * wy[b];
#pragma unroll
          for (int c = 0; c < Gm::CNT; ++c) {
            const float t3 = t2 * wz[c];
            const int e = (lx + a) + Gm::EX * ((ly + b) + Gm::EY * (lz + c));
            sx += t3 * sE0[e];
            sy += t3 * sE1[e];
            sz += t3 * sE2[e];
          }
        }
      }
      sx += g.x;
      sy += g.y;
      sz += g.z;
      if (mode & MODE_ACC) acc[i] = make_float4(sx, sy, sz, 0.f);
      if (mode & (MODE_KICK | MODE_DRIFT)) {
        float4 v = vel[i];
        leapfrog_tail(p, v, sx, sy, sz, mode, kick, box, escape, iteration);
        vel[i] = v;
        pm[i] = p;
      }
    }
    __syncthreads();
  }
}

// E on every cell, for Grid::getField (include/grid.h:26).
__global__ void k_field(const float* __restrict__ phi, int nx, int ny, int nz, int fds,
                        float* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t M = (int64_t)nx * ny * nz;
  if (i >= M) return;
  const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
  auto P = [&](int a, int b, int c) {
    return phi[(int64_t)wrapc(a, nx) + (int64_t)nx * (wrapc(b, ny) + (int64_t)ny * wrapc(c, nz))];
  };
  float Ex, Ey, Ez;
  if (fds == PMGPU_TWO_POINT) {
    Ex = -0.5f * (P(x + 1, y, z) - P(x - 1, y, z));
    Ey = -0.5f * (P(x, y + 1, z) - P(x, y - 1, z));
    Ez = -0.5f * (P(x, y, z + 1) - P(x, y, z - 1));
  } else {
    Ex = (-1.0f / 12) * (-P(x + 2, y, z) + 8 * P(x + 1, y, z) - 8 * P(x - 1, y, z) + P(x - 2, y, z));
    Ey = (-1.0f / 12) * (-P(x, y + 2, z) + 8 * P(x, y + 1, z) - 8 * P(x, y - 1, z) + P(x, y - 2, z));
    Ez = (-1.0f / 12) * (-P(x, y, z + 2) + 8 * P(x, y, z + 1) - 8 * P(x, y, z - 1) + P(x, y, z - 2));
  }
  out[3 * i] = Ex;
  out[3 * i + 1] = Ey;
  out[3 * i + 2] = Ez;
}

// E = -grad(phi) on every cell as float4 (x, y, z, 0), for the gather
// variant of the interpolation.
template <int FDS>
__global__ void k_field4(const float* __restrict__ phi, int nx, int ny, int nz, float4* __restrict__ out,
                         const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t M = (int64_t)nx * ny * nz;
  if (i >= M) return;
  const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  auto X = [&](int d) { return phi[i - x + wrapc(x + d, nx)]; };
  auto Y = [&](int d) { return phi[i + (wrapc(y + d, ny) - y) * sy]; };
  auto Z = [&](int d) { return phi[i + (wrapc(z + d, nz) - z) * sz]; };
  float4 e;
  if constexpr (FDS == PMGPU_TWO_POINT) {
    e.x = -0.5f * (X(1) - X(-1));
    e.y = -0.5f * (Y(1) - Y(-1));
    e.z = -0.5f * (Z(1) - Z(-1));
  } else {
    e.x = (-1.0f / 12) * (-X(2) + 8 * X(1) - 8 * X(-1) + X(-2));
    e.y = (-1.0f / 12) * (-Y(2) + 8 * Y(1) - 8 * Y(-1) + Y(-2));
    e.z = (-1.0f / 12) * (-Z(2) + 8 * Z(1) - 8 * Z(-1) + Z(-2));
  }
  e.w = 0.0f;
  out[i] = e;
}

// Interpolation as a gather from the E grid, one particle per thread (the
// particles are tile-sorted, so neighbouring threads read neighbouring
// cells through the caches).  Same weights, operation order and leapfrog
// tail as k_interp; no per-tile staging, so sparse tiles cost nothing extra.
template <int SCHEME>
__global__ __launch_bounds__(256) void k_interp_gather(
    const float4* __restrict__ E, int nx, int ny, int nz, int64_t n, float4* __restrict__ pm,
    float4* __restrict__ vel, float4* __restrict__ acc, int mode, float kick, float3 box, ExtField ext,
    int centred, int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int SLO = SCHEME == PMGPU_TSC ? -1 : 0;
  constexpr int CNT = SCHEME == PMGPU_TSC ? 3 : (SCHEME == PMGPU_CIC ? 2 : 1);
  float4 p = pm[i];
  const int ax = anchor_of(p.x, SCHEME, centred), ay = anchor_of(p.y, SCHEME, centred), az = anchor_of(p.z, SCHEME, centred);
  float wx[3], wy[3], wz[3];
  interp_weights<SCHEME>(p.x - (float)ax, wx);
  interp_weights<SCHEME>(p.y - (float)ay, wy);
  interp_weights<SCHEME>(p.z - (float)az, wz);
  int gx[3], gy[3], gz[3];
#pragma unroll
  for (int a = 0; a < CNT; ++a) {
    gx[a] = wrapc(ax + SLO + a, nx);
    gy[a] = wrapc(ay + SLO + a, ny);
    gz[a] = wrapc(az + SLO + a, nz);
  }
  // 32-bit cell offsets (the grid has at most 2^30 cells)
  uint32_t rowoff[3][3];
#pragma unroll
  for (int b = 0; b < CNT; ++b)
#pragma unroll
    for (int c = 0; c < CNT; ++c) rowoff[b][c] = (uint32_t)nx * ((uint32_t)gy[b] + (uint32_t)ny * (uint32_t)gz[c]);
  float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
  for (int a = 0; a < CNT; ++a) {
    // the TSC's 1/8 once per x weight, not per tap (a power of two: the same
    // products)
    const float t1 = SCHEME == PMGPU_TSC ? wx[a] * 0.125f : wx[a];
#pragma unroll
    for (int b = 0; b < CNT; ++b) {
      const float t2 = t1 * wy[b];
#pragma unroll
      for (int c = 0; c < CNT; ++c) {
        const float t3 = t2 * wz[c];
        const float4 e = E[(uint32_t)gx[a] + rowoff[b][c]];
        sx += t3 * e.x;
        sy += t3 * e.y;
        sz += t3 * e.z;
      }
    }
  }
  const float3 g = ext_field(ext, p.x, p.y, p.z);
  sx += g.x;
  sy += g.y;
  sz += g.z;
  if (mode & MODE_ACC) acc[i] = make_float4(sx, sy, sz, 0.f);
  if (mode & (MODE_KICK | MODE_DRIFT)) {
    float4 v = vel[i];
    leapfrog_tail(p, v, sx, sy, sz, mode, kick, box, escape, iteration);
    vel[i] = v;
    pm[i] = p;
  }
}

// Interpolation straight from the potential, one particle per thread, with
// no field grid: the field interpolated with the assignment weights w and
// the difference stencil s (E = -sum_o s_o phi(x + o)) is, per component,
// a separable sum over the potential with the derivative axis weighted by
// the convolution D = w * s, e.g. a_x = -sum D_x(X) w_y(Y) w_z(Z) phi(X, Y, Z).
// The same quantity as k_interp up to rounding (the order of the sums
// differs).
template <int SCHEME, int FDS>
__global__ __launch_bounds__(256) void k_interp_phi(
    const float* __restrict__ phi, int nx, int ny, int nz, int64_t n, float4* __restrict__ pm,
    float4* __restrict__ vel, float4* __restrict__ acc, int mode, float kick, float3 box, ExtField ext,
    int centred, int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int SLO = SCHEME == PMGPU_TSC ? -1 : 0;
  constexpr int CNT = SCHEME == PMGPU_TSC ? 3 : (SCHEME == PMGPU_CIC ? 2 : 1);
  constexpr int F = FDS == PMGPU_TWO_POINT ? 1 : 2;
  constexpr int EXT = CNT + 2 * F;
  // difference stencil s_o, o = -F .. F (E = -sum_o s_o phi(x + o))
  constexpr float S2[3] = {-0.5f, 0.0f, 0.5f};
  constexpr float S4[5] = {1.0f / 12, -8.0f / 12, 0.0f, 8.0f / 12, -1.0f / 12};
  float4 p = pm[i];
  const int ax = anchor_of(p.x, SCHEME, centred), ay = anchor_of(p.y, SCHEME, centred), az = anchor_of(p.z, SCHEME, centred);
  float wx[3], wy[3], wz[3];
  interp_weights<SCHEME>(p.x - (float)ax, wx);
  interp_weights<SCHEME>(p.y - (float)ay, wy);
  interp_weights<SCHEME>(p.z - (float)az, wz);
  float dx[EXT], dy[EXT], dz[EXT];
#pragma unroll
  for (int e = 0; e < EXT; ++e) dx[e] = dy[e] = dz[e] = 0.0f;
#pragma unroll
  for (int a = 0; a < CNT; ++a)
#pragma unroll
    for (int o = -F; o <= F; ++o) {
      const float so = F == 1 ? S2[o + 1] : S4[o + 2];
      if (so == 0.0f) continue;
      dx[a + F + o] = fmaf(wx[a], so, dx[a + F + o]);
      dy[a + F + o] = fmaf(wy[a], so, dy[a + F + o]);
      dz[a + F + o] = fmaf(wz[a], so, dz[a + F + o]);
    }
  int gx[EXT], gy[EXT], gz[EXT];
#pragma unroll
  for (int e = 0; e < EXT; ++e) {
    gx[e] = wrapc(ax + SLO - F + e, nx);
    gy[e] = wrapc(ay + SLO - F + e, ny) * nx;
    gz[e] = wrapc(az + SLO - F + e, nz) * ny;
  }
  auto P = [&](int a, int b, int c) { return phi[(int64_t)gx[a] + (int64_t)gy[b] + (int64_t)nx * gz[c]]; };
  float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
  for (int b = 0; b < CNT; ++b)
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
      const float wyz = wy[b] * wz[c];
      float line = 0.0f;
#pragma unroll
      for (int e = 0; e < EXT; ++e) line = fmaf(dx[e], P(e, F + b, F + c), line);
      sx = fmaf(wyz, line, sx);
    }
#pragma unroll
  for (int a = 0; a < CNT; ++a)
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
      const float wxz = wx[a] * wz[c];
      float line = 0.0f;
#pragma unroll
      for (int e = 0; e < EXT; ++e) line = fmaf(dy[e], P(F + a, e, F + c), line);
      sy = fmaf(wxz, line, sy);
    }
#pragma unroll
  for (int a = 0; a < CNT; ++a)
#pragma unroll
    for (int b = 0; b < CNT; ++b) {
      const float wxy = wx[a] * wy[b];
      float line = 0.0f;
#pragma unroll
      for (int e = 0; e < EXT; ++e) line = fmaf(dz[e], P(F + a, F + b, e), line);
      sz = fmaf(wxy, line, sz);
    }
  const float k = SCHEME == PMGPU_TSC ? -0.125f : -1.0f;
  sx *= k;
  sy *= k;
  sz *= k;
  const float3 g = ext_field(ext, p.x, p.y, p.z);
  sx += g.x;
  sy += g.y;
  sz += g.z;
  if (mode & MODE_ACC) acc[i] = make_float4(sx, sy, sz, 0.f);
  if (mode & (MODE_KICK | MODE_DRIFT)) {
    float4 v = vel[i];
    leapfrog_tail(p, v, sx, sy, sz, mode, kick, box, escape, iteration);
    vel[i] = v;
    pm[i] = p;
  }
}

// Leapfrog pieces on the sorted state (source/leapfrog.cpp:5-24).
__global__ void k_leapfrog(float4* __restrict__ pm, float4* __restrict__ vel,
                           const float4* __restrict__ acc, float4* __restrict__ vint, int64_t n,
                           int op, float dt, float3 box, int32_t* __restrict__ escape,
                           int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 v = vel[i];
  if (op == PMGPU_DRIFT) {
    float4 p = pm[i];
    p.x += dt * v.x;
    p.y += dt * v.y;
    p.z += dt * v.z;
    pm[i] = p;
    bool inside = inside_box(p.x, p.y, p.z, box);
    // iteration < 0: a standalone updatePositions, which never stops a run
    if (!inside && iteration >= 0) atomicMin(escape, iteration);
    return;
  }
  float4 a = acc[i];
  if (op == PMGPU_INTEGER_STEP) {
    vint[i] = make_float4(v.x + 0.5f * dt * a.x, v.y + 0.5f * dt * a.y, v.z + 0.5f * dt * a.z, v.w);
    return;
  }
  const float f = op == PMGPU_HALF_KICK ? 0.5f * dt : dt;
  v.x += f * a.x;
  v.y += f * a.y;
  v.z += f * a.z;
  vel[i] = v;
}

__global__ void k_escape(const float4* __restrict__ pm, int64_t n, float3 box,
                         int32_t* __restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = pm[i];
  bool inside = inside_box(p.x, p.y, p.z, box);
  if (!inside) atomicOr(flag, 1);
}

template <int S, int F>
void launch_interp(pmgpu_ctx* c, int grid, int mode, float kick, float3 box, ExtField e,
                   int iteration) {
  hipLaunchKernelGGL((k_interp<S, F>), dim3(grid), dim3(256), 0, c->stream, c->phi, c->nx, c->ny,
                     c->nz, c->ntx, c->nty, c->ntiles, c->cls_start, c->int_chunk_start, c->int_owner,
                     c->pm[c->cur], c->vel[c->cur], c->acc, mode, kick, box, e, c->int_chunk, c->tsc_centred,
                     c->escape, iteration);
}

}  // namespace

static ExtField ext_of(pmgpu_ctx* c) {
  ExtField e{};
  e.kind = c->prm.ext_kind;
  e.cx = c->prm.ext_params[0];
  e.cy = c->prm.ext_params[1];
  e.cz = c->prm.ext_params[2];
  e.R = c->prm.ext_params[3];
  e.M = c->prm.ext_params[4];
  e.G = c->prm.ext_params[5];
  e.H = c->prm.H;
  e.DT = c->prm.DT;
  return e;
}

static float3 box_of(pmgpu_ctx* c) {
  return make_float3(c->prm.box[0], c->prm.box[1], c->prm.box[2]);
}

// The gather variant: E on every cell, then one particle per thread.
static void interpolate_gather(pmgpu_ctx* c, int mode, float kick, int iteration) {
  if (!c->efield) HIP_CHECK(hipMalloc((void**)&c->efield, sizeof(float4) * (size_t)c->M));
  const unsigned gm = (unsigned)((c->M + 255) / 256), gp = (unsigned)((c->n + 255) / 256);
  if (c->prm.finite_diff == PMGPU_TWO_POINT)
    hipLaunchKernelGGL(k_field4<PMGPU_TWO_POINT>, dim3(gm), dim3(256), 0, c->stream, c->phi, c->nx, c->ny, c->nz,
                       c->efield, c->escape, iteration);
  else
    hipLaunchKernelGGL(k_field4<PMGPU_FOUR_POINT>, dim3(gm), dim3(256), 0, c->stream, c->phi, c->nx, c->ny, c->nz,
                       c->efield, c->escape, iteration);
  HIP_CHECK(hipGetLastError());
  const ExtField e = ext_of(c);
  const float3 box = box_of(c);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(gp), dim3(256), 0, c->stream, (const float4*)c->efield, c->nx, c->ny, c->nz, c->n,
                       c->pm[c->cur], c->vel[c->cur], c->acc, mode, kick, box, e, c->tsc_centred, c->escape,
                       iteration);
  };
  if (c->prm.interpolation == PMGPU_NGP) go(k_interp_gather<PMGPU_NGP>);
  else if (c->prm.interpolation == PMGPU_CIC) go(k_interp_gather<PMGPU_CIC>);
  else go(k_interp_gather<PMGPU_TSC>);
  HIP_CHECK(hipGetLastError());
}

// The potential-gather variant: no field grid.
static void interpolate_phi(pmgpu_ctx* c, int mode, float kick, int iteration) {
  const unsigned gp = (unsigned)((c->n + 255) / 256);
  const ExtField e = ext_of(c);
  const float3 box = box_of(c);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(gp), dim3(256), 0, c->stream, (const float*)c->phi, c->nx, c->ny, c->nz, c->n,
                       c->pm[c->cur], c->vel[c->cur], c->acc, mode, kick, box, e, c->tsc_centred, c->escape,
                       iteration);
  };
  const int s = c->prm.interpolation, f = c->prm.finite_diff;
  if (s == PMGPU_NGP && f == PMGPU_TWO_POINT) go(k_interp_phi<PMGPU_NGP, PMGPU_TWO_POINT>);
  else if (s == PMGPU_NGP) go(k_interp_phi<PMGPU_NGP, PMGPU_FOUR_POINT>);
  else if (s == PMGPU_CIC && f == PMGPU_TWO_POINT) go(k_interp_phi<PMGPU_CIC, PMGPU_TWO_POINT>);
  else if (s == PMGPU_CIC) go(k_interp_phi<PMGPU_CIC, PMGPU_FOUR_POINT>);
  else if (f == PMGPU_TWO_POINT) go(k_interp_phi<PMGPU_TSC, PMGPU_TWO_POINT>);
  else go(k_interp_phi<PMGPU_TSC, PMGPU_FOUR_POINT>);
  HIP_CHECK(hipGetLastError());
}

void interpolate(pmgpu_ctx* c, int mode, float kick, int iteration) {
  int tok = stage_begin(c, "interp");
  if (c->interp_kernel == 3) {
    interpolate_phi(c, mode, kick, iteration);
    stage_end(c, tok);
    return;
  }
  // auto: the gather variant when tiles are sparse (the tiled variant's
  // per-tile staging of the potential and field then dominates)
  const bool gather = c->general || c->interp_kernel == 2 ||
                      (c->interp_kernel == 0 && c->ntiles < INTERP_GATHER_TILES_PER_CU * c->num_cu);
  if (gather) {
    interpolate_gather(c, mode, kick, iteration);
    stage_end(c, tok);
    return;
  }
  int64_t bound = (int64_t)c->ntiles + c->n / c->int_chunk + 1;
  int grid = (int)std::min<int64_t>(bound, (int64_t)c->num_cu * c->blocks_per_cu);
  const ExtField e = ext_of(c);
  const float3 box = box_of(c);
  const int s = c->prm.interpolation, f = c->prm.finite_diff;
  if (s == PMGPU_NGP && f == PMGPU_TWO_POINT) launch_interp<PMGPU_NGP, PMGPU_TWO_POINT>(c, grid, mode, kick, box, e, iteration);
  else if (s == PMGPU_NGP) launch_interp<PMGPU_NGP, PMGPU_FOUR_POINT>(c, grid, mode, kick, box, e, iteration);
  else if (s == PMGPU_CIC && f == PMGPU_TWO_POINT) launch_interp<PMGPU_CIC, PMGPU_TWO_POINT>(c, grid, mode, kick, box, e, iteration);
  else if (s == PMGPU_CIC) launch_interp<PMGPU_CIC, PMGPU_FOUR_POINT>(c, grid, mode, kick, box, e, iteration);
  else if (f == PMGPU_TWO_POINT) launch_interp<PMGPU_TSC, PMGPU_TWO_POINT>(c, grid, mode, kick, box, e, iteration);
  else launch_interp<PMGPU_TSC, PMGPU_FOUR_POINT>(c, grid, mode, kick, box, e, iteration);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
}

void field_grid(pmgpu_ctx* c) {
  hipLaunchKernelGGL(k_field, dim3((unsigned)((c->M + 255) / 256)), dim3(256), 0, c->stream, c->phi,
                     c->nx, c->ny, c->nz, c->prm.finite_diff, c->field);
  HIP_CHECK(hipGetLastError());
}

void leapfrog_op(pmgpu_ctx* c, int op, float dt, int iteration) {
  int tok = stage_begin(c, "leapfrog");
  hipLaunchKernelGGL(k_leapfrog, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream,
                     c->pm[c->cur], c->vel[c->cur], c->acc, c->vint, c->n, op, dt, box_of(c),
                     c->escape, iteration);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
}

void escape_check(pmgpu_ctx* c, int32_t* flag_dev) {
  hipLaunchKernelGGL(k_escape, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream,
                     c->pm[c->cur], c->n, box_of(c), flag_dev);
  HIP_CHECK(hipGetLastError());
}

__global__ void k_add_accelerations(float4* __restrict__ acc, const float4* __restrict__ vel,
                                    const float* __restrict__ add, int64_t n) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t id = __float_as_int(vel[s].w);
  float4 a = acc[s];
  a.x += add[3 * id];
  a.y += add[3 * id + 1];
  a.z += add[3 * id + 2];
  acc[s] = a;
}

void add_accelerations(pmgpu_ctx* c, const float* add) {
  hipLaunchKernelGGL(k_add_accelerations, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream, c->acc,
                     c->vel[c->cur], add, c->n);
  HIP_CHECK(hipGetLastError());
}

}  // namespace pmgpu
