// The force step on grids whose sides are not powers of two in [32, 1024]
// (the reference takes any grid through FFTW, e.g. the 2Ng x 2Ng x Ng grids
// of its accuracy demo, source/demos.cpp:381-408).  These grids are small
// (at most a few million cells), so this path favours generality:
//  - deposit: one thread per particle, each contribution added to a 64-bit
//    fixed-point grid with a global integer atomic (associative, so the
//    density is bitwise reproducible, as on the tiled path), then converted;
//  - transforms: a direct DFT along each axis, one line per workgroup staged
//    in LDS, twiddles from a table indexed by (j k) mod n -- the same
//    half-spectrum layout, normalisation and G multiply as the FFT path
//    (source/grid.cpp:50-56, source/pmMethod.cpp:340-350);
//  - the interpolation is the gather variant (no tiles).
#include "pmgpu_internal.h"

namespace pmgpu {

namespace {

constexpr int DFT_THREADS = 256;

__device__ __forceinline__ int wrapg(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}

__device__ __forceinline__ int anchor_g(float x, int scheme, int centred) {
  float f = (scheme == PMGPU_NGP || centred) ? roundf(x) : floorf(x);
  f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
  return (int)f;
}

// Per-axis weights of the deposit (PMMethod::spreadMass, pmMethod.cpp:200-277;
// TSCAssignmentFunc, :187-198 with the 1/8 applied to the mass).
__device__ __forceinline__ void weights_g(int scheme, float d, float w[3]) {
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

// The deposit's fixed-point conversion (round to nearest even of x 2^shift
// in FP32 and integer operations; see the tiled deposit).
__device__ __forceinline__ unsigned long long fixed_g(float x, float scale) {
  const float t = x * scale;
  const int hi = (int)(t * 0x1p-31f);
  const float rem = fmaf((float)hi, -0x1p31f, t);
  const int lo = (int)rintf(rem);
  return (unsigned long long)(((long long)hi << 31) + (long long)lo);
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n, const int32_t* __restrict__ escape,
                           int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

template <int SCHEME>
__global__ __launch_bounds__(256) void k_deposit_global(const float4* __restrict__ pm, int64_t n, int nx, int ny,
                                                        int nz, int centred, int shift,
                                                        unsigned long long* __restrict__ fx,
                                                        const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int LO = SCHEME == PMGPU_TSC ? -1 : 0;
  constexpr int CNT = SCHEME == PMGPU_TSC ? 3 : (SCHEME == PMGPU_CIC ? 2 : 1);
  const float scale = ldexpf(1.0f, shift);
  const float4 p = pm[i];
  const int ax = anchor_g(p.x, SCHEME, centred), ay = anchor_g(p.y, SCHEME, centred),
            az = anchor_g(p.z, SCHEME, centred);
  float wx[3], wy[3], wz[3];
  weights_g(SCHEME, p.x - (float)ax, wx);
  weights_g(SCHEME, p.y - (float)ay, wy);
  weights_g(SCHEME, p.z - (float)az, wz);
  const float m = SCHEME == PMGPU_TSC ? p.w / 8.0f : p.w;
#pragma unroll
  for (int a = 0; a < CNT; ++a) {
    const int gx = wrapg(ax + LO + a, nx);
    const float t1 = m * wx[a];
#pragma unroll
    for (int b = 0; b < CNT; ++b) {
      const int gy = wrapg(ay + LO + b, ny);
      const float t2 = t1 * wy[b];
#pragma unroll
      for (int c = 0; c < CNT; ++c) {
        const int gz = wrapg(az + LO + c, nz);
        const float t3 = t2 * wz[c];
        atomicAdd(&fx[(int64_t)gx + (int64_t)nx * ((int64_t)gy + (int64_t)ny * gz)], fixed_g(t3, scale));
      }
    }
  }
}

__global__ void k_fixed_to_rho(const unsigned long long* __restrict__ fx, int64_t m, int shift,
                               float* __restrict__ rho, const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) rho[i] = ldexpf((float)(long long)fx[i], -shift);
}

// x pass, real -> half spectrum: X[k] = sum_j x[j] exp(-2 pi i j k / nx),
// k = 0 .. nx/2, one row per workgroup; tw[m] = exp(-2 pi i m / nx).
__global__ __launch_bounds__(DFT_THREADS) void k_dft_x_r2c(const float* __restrict__ rho, float2* __restrict__ spec,
                                                           int nx, int sp, const float2* __restrict__ tw,
                                                           const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  __shared__ float row[1024];
  __shared__ float2 w[1024];
  const int64_t r = blockIdx.x;
  for (int j = threadIdx.x; j < nx; j += DFT_THREADS) {
    row[j] = rho[r * nx + j];
    w[j] = tw[j];
  }
  __syncthreads();
  for (int k = threadIdx.x; k <= nx / 2; k += DFT_THREADS) {
    float sr = 0.f, si = 0.f;
    int idx = 0;
    for (int j = 0; j < nx; ++j) {
      const float2 t = w[idx];
      sr = fmaf(row[j], t.x, sr);
      si = fmaf(row[j], t.y, si);
      idx += k;
      if (idx >= nx) idx -= nx;
    }
    spec[r * sp + k] = make_float2(sr, si);
  }
}

// x pass, half spectrum -> real (unnormalised, the inverse of the above for
// a Hermitian spectrum): x[j] = Re X[0] + 2 sum_{0<k<nx/2} Re(X[k] e^{+2 pi i jk/nx})
// + (nx even) Re X[nx/2] (-1)^j.
__global__ __launch_bounds__(DFT_THREADS) void k_dft_x_c2r(const float2* __restrict__ spec, float* __restrict__ phi,
                                                           int nx, int sp, const float2* __restrict__ tw,
                                                           const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  __shared__ float2 X[513];
  __shared__ float2 w[1024];
  const int64_t r = blockIdx.x;
  const int nh = nx / 2;
  for (int k = threadIdx.x; k <= nh; k += DFT_THREADS) X[k] = spec[r * sp + k];
  for (int j = threadIdx.x; j < nx; j += DFT_THREADS) w[j] = tw[j];
  __syncthreads();
  const int kmax = (nx - 1) / 2;  // the terms with a conjugate partner
  for (int j = threadIdx.x; j < nx; j += DFT_THREADS) {
    float s = 0.f;
    int idx = j;
    for (int k = 1; k <= kmax; ++k) {
      // e^{+2 pi i jk/nx} = conj(tw[(jk) mod nx])
      const float2 t = w[idx];
      s = fmaf(X[k].x, t.x, s);
      s = fmaf(X[k].y, t.y, s);
      idx += j;
      if (idx >= nx) idx -= nx;
    }
    float v = X[0].x + 2.0f * s;
    if ((nx & 1) == 0) v += (j & 1) ? -X[nh].x : X[nh].x;
    phi[r * nx + j] = v;
  }
}

// A direct complex DFT of length n along one axis of the half spectrum, in
// place, one line per workgroup: line `l` = (outer, kx) with kx < nxh fastest,
// element j at base(l) + j * stride.  MODE 0: forward.  1: inverse.  2:
// forward, multiply by g (same layout), inverse (the z pass of the force).
template <int MODE>
__global__ __launch_bounds__(DFT_THREADS) void k_dft_lines(float2* __restrict__ spec, int n, int nxh,
                                                           int64_t outer_stride, int64_t stride,
                                                           const float* __restrict__ g,
                                                           const float2* __restrict__ tw,
                                                           const int32_t* __restrict__ escape, int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  __shared__ float2 a[1024];
  __shared__ float2 b[1024];
  __shared__ float2 w[1024];
  const int64_t l = blockIdx.x;
  const int64_t base = (l / nxh) * outer_stride + (l % nxh);
  for (int j = threadIdx.x; j < n; j += DFT_THREADS) {
    a[j] = spec[base + j * stride];
    w[j] = tw[j];
  }
  __syncthreads();
  auto pass = [&](const float2* in, float2* out, bool inverse) {
    for (int k = threadIdx.x; k < n; k += DFT_THREADS) {
      float sr = 0.f, si = 0.f;
      int idx = 0;
      for (int j = 0; j < n; ++j) {
        float2 t = w[idx];
        if (inverse) t.y = -t.y;
        sr = fmaf(in[j].x, t.x, fmaf(-in[j].y, t.y, sr));
        si = fmaf(in[j].x, t.y, fmaf(in[j].y, t.x, si));
        idx += k;
        if (idx >= n) idx -= n;
      }
      out[k] = make_float2(sr, si);
    }
    __syncthreads();
  };
  if constexpr (MODE == 0) {
    pass(a, b, false);
  } else if constexpr (MODE == 1) {
    pass(a, b, true);
  } else {
    pass(a, b, false);
    for (int j = threadIdx.x; j < n; j += DFT_THREADS) {
      const float s = g[base + j * stride];
      b[j] = make_float2(b[j].x * s, b[j].y * s);
    }
    __syncthreads();
    pass(b, a, true);
    for (int j = threadIdx.x; j < n; j += DFT_THREADS) b[j] = a[j];
    __syncthreads();
  }
  for (int j = threadIdx.x; j < n; j += DFT_THREADS) spec[base + j * stride] = b[j];
}

unsigned blocks_g(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

bool general_grid(int nx, int ny, int nz) {
  auto fast = [](int v) { return v >= 32 && v <= 1024 && (v & (v - 1)) == 0; };
  return !(fast(nx) && fast(ny) && fast(nz));
}

void deposit_general(pmgpu_ctx* c) {
  int tok = stage_begin(c, "deposit");
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_g(c->M)), dim3(256), 0, c->stream, c->fx_grid, c->M, c->escape,
                     c->iter);
  HIP_CHECK(hipGetLastError());
  if (c->n > 0) {
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(blocks_g(c->n)), dim3(256), 0, c->stream, c->pm[c->cur], c->n, c->nx, c->ny,
                         c->nz, c->tsc_centred, c->fx_shift, c->fx_grid, c->escape, c->iter);
    };
    if (c->prm.interpolation == PMGPU_TSC) go(k_deposit_global<PMGPU_TSC>);
    else if (c->prm.interpolation == PMGPU_CIC) go(k_deposit_global<PMGPU_CIC>);
    else go(k_deposit_global<PMGPU_NGP>);
    HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_fixed_to_rho, dim3(blocks_g(c->M)), dim3(256), 0, c->stream, c->fx_grid, c->M, c->fx_shift,
                     c->rho, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
}

void fft_general(pmgpu_ctx* c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nxh = c->nxh, sp = c->sp;
  int tok = stage_begin(c, "fft_x_r2c");
  hipLaunchKernelGGL(k_dft_x_r2c, dim3((unsigned)(ny * nz)), dim3(DFT_THREADS), 0, c->stream, c->rho, c->spec, nx, sp,
                     c->tw_gx, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
  // y lines: (z, kx), element y at z*ny*sp + y*sp + kx
  tok = stage_begin(c, "fft_y_fwd");
  hipLaunchKernelGGL(k_dft_lines<0>, dim3((unsigned)(nz * nxh)), dim3(DFT_THREADS), 0, c->stream, c->spec, ny, nxh,
                     (int64_t)ny * sp, (int64_t)sp, (const float*)nullptr, c->tw_y, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
  // z lines: (y, kx), element z at z*ny*sp + y*sp + kx; forward, x G, inverse
  tok = stage_begin(c, "fft_z_green");
  hipLaunchKernelGGL(k_dft_lines<2>, dim3((unsigned)(ny * nxh)), dim3(DFT_THREADS), 0, c->stream, c->spec, nz, nxh,
                     (int64_t)sp, (int64_t)ny * sp, c->greens, c->tw_z, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
  tok = stage_begin(c, "fft_y_inv");
  hipLaunchKernelGGL(k_dft_lines<1>, dim3((unsigned)(nz * nxh)), dim3(DFT_THREADS), 0, c->stream, c->spec, ny, nxh,
                     (int64_t)ny * sp, (int64_t)sp, (const float*)nullptr, c->tw_y, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
  tok = stage_begin(c, "fft_x_c2r");
  hipLaunchKernelGGL(k_dft_x_c2r, dim3((unsigned)(ny * nz)), dim3(DFT_THREADS), 0, c->stream, c->spec, c->phi, nx, sp,
                     c->tw_gx, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
}

}  // namespace pmgpu
