// 3-D real-to-complex FFT -> x G(k) -> complex-to-real FFT, replacing
// Grid::fftDensity, PMMethod::findFourierPotential and Grid::invFftPotential
// (source/grid.cpp:50-56, source/pmMethod.cpp:340-350, FFTWAdapter.cpp:27-43).
//
// Five passes over HBM, each a batch of 1-D transforms staged in LDS:
//   x  (R2C): rows of Nx reals -> Nx/2+1 complex, via an Nx/2 complex FFT
//             and the even/odd split;
//   y        : forward along y;
//   z        : forward along z, multiply by the real influence function G,
//              inverse along z -- one kernel, the pencil never leaves LDS;
//   y        : inverse;
//   x  (C2R): Nx/2+1 complex -> Nx reals.
// The spectrum is stored [z][y][kx] (kx fastest) with row pitch SP: Nx/2+1
// rounded up to a multiple of 16, so that every row, and with it every
// column tile of the y and z passes, starts on a 128-byte boundary.  Every
// 1-D transform is a Stockham autosort FFT of radix-8 / radix-4 stages with
// L/8 threads per transform; the inverse is unnormalised and 1/(NxNyNz) is
// folded into G (the reference's adapters divide after the inverse,
// FFTWAdapter.cpp:39-40).
#include "pmgpu_internal.h"

namespace pmgpu {

namespace {

__device__ __forceinline__ float2 operator+(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 operator-(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 scale2(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2* a) {
  float2 t = a[0];
  a[0] = t + a[1];
  a[1] = t - a[1];
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = rot90<INV>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* a) {
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  const float h = 0.70710678118654752f;
  // W8^1 = h(1 -/+ i), W8^2 = -/+ i, W8^3 = h(-1 -/+ i)
  o1 = INV ? make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y))
           : make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
  o2 = rot90<INV>(o2);
  o3 = INV ? make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y))
           : make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
  a[0] = e0 + o0;
  a[4] = e0 - o0;
  a[1] = e1 + o1;
  a[5] = e1 - o1;
  a[2] = e2 + o2;
  a[6] = e2 - o2;
  a[3] = e3 + o3;
  a[7] = e3 - o3;
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* a) {
  if constexpr (R == 8) dft8<INV>(a);
  else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
  else dft2<INV>(a);
}

// One Stockham stage of radix R over a transform of length L stored in LDS
// (element n of column col at buf[col*cs + n*ns]).  The L/8 threads of a
// column each do 8/R butterflies.  tw[m] = exp(-2 pi i m / L).
template <int L, int R, int NS, bool INV>
__device__ __forceinline__ void stockham_stage(float2* buf, int cs, int ns, int col, int t,
                                               bool active, const float2* __restrict__ tw) {
  constexpr int T = L / 8, NB = 8 / R, LR = L / R;
  float2 v[8];
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + b * T;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b * R + r] = buf[col * cs + (j + r * LR) * ns];
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + b * T;
      const int k = j & (NS - 1);
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[(r * k) * (L / (NS * R))];
          if (INV) w.y = -w.y;
          v[b * R + r] = cmul(v[b * R + r], w);
        }
      }
      dft<R, INV>(&v[b * R]);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[col * cs + (base + r * NS) * ns] = v[b * R + r];
    }
  }
  __syncthreads();
}

template <int L, bool INV, int NS, int... Rs>
struct Stages;
template <int L, bool INV, int NS>
struct Stages<L, INV, NS> {
  static __device__ __forceinline__ void run(float2*, int, int, int, int, bool, const float2*) {}
};
template <int L, bool INV, int NS, int R, int... Rs>
struct Stages<L, INV, NS, R, Rs...> {
  static __device__ __forceinline__ void run(float2* buf, int cs, int ns, int col, int t,
                                             bool active, const float2* tw) {
    stockham_stage<L, R, NS, INV>(buf, cs, ns, col, t, active, tw);
    Stages<L, INV, NS * R, Rs...>::run(buf, cs, ns, col, t, active, tw);
  }
};

// Radix plans: L/8 threads per transform, 8 points per thread per stage.
template <int L, bool INV>
__device__ __forceinline__ void fft_lds(float2* buf, int cs, int ns, int col, int t, bool active,
                                        const float2* tw) {
  if constexpr (L == 16) Stages<L, INV, 1, 4, 4>::run(buf, cs, ns, col, t, active, tw);
  else if constexpr (L == 32) Stages<L, INV, 1, 8, 4>::run(buf, cs, ns, col, t, active, tw);
  else if constexpr (L == 64) Stages<L, INV, 1, 8, 8>::run(buf, cs, ns, col, t, active, tw);
  else if constexpr (L == 128) Stages<L, INV, 1, 8, 4, 4>::run(buf, cs, ns, col, t, active, tw);
  else if constexpr (L == 256) Stages<L, INV, 1, 8, 8, 4>::run(buf, cs, ns, col, t, active, tw);
  else if constexpr (L == 512) Stages<L, INV, 1, 8, 8, 8>::run(buf, cs, ns, col, t, active, tw);
  else Stages<L, INV, 1, 8, 8, 4, 4>::run(buf, cs, ns, col, t, active, tw);
}

// ---------------------------------------------------------------------------
// x pass, real -> complex.  L = Nx/2 complex points per row, ROWS rows per
// workgroup of 256 threads.
template <int L>
__global__ __launch_bounds__(256) void k_fft_x_r2c(const float* __restrict__ rho,
                                                   float2* __restrict__ spec, int sp,
                                                   const float2* __restrict__ tw,
                                                   const float2* __restrict__ tw_r2c,
                                                   const int32_t* __restrict__ escape,
                                                   int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  constexpr int T = L / 8, ROWS = 256 / T, P = L + 2;
  __shared__ float2 buf[ROWS * P];
  const int row0 = blockIdx.x * ROWS;
  const float2* src = reinterpret_cast<const float2*>(rho) + (int64_t)row0 * L;
  for (int e = threadIdx.x; e < ROWS * L; e += 256) buf[(e / L) * P + (e % L)] = src[e];
  __syncthreads();
  const int col = threadIdx.x / T, t = threadIdx.x % T;
  fft_lds<L, false>(buf, P, 1, col, t, true, tw);
  float2* dst = spec + (int64_t)(row0 + col) * sp;
  for (int k = t; k <= L; k += T) {
    float2 zk = buf[col * P + (k & (L - 1))];
    float2 zc = conjf2(buf[col * P + ((L - k) & (L - 1))]);
    float2 e = scale2(zk + zc, 0.5f);
    float2 o = scale2(rot90<false>(zk - zc), 0.5f);  // (zk - zc) / (2i)
    dst[k] = e + cmul(tw_r2c[k], o);
  }
}

// x pass, complex -> real (unnormalised).
template <int L>
__global__ __launch_bounds__(256) void k_fft_x_c2r(const float2* __restrict__ spec, int sp,
                                                   float* __restrict__ phi,
                                                   const float2* __restrict__ tw,
                                                   const float2* __restrict__ tw_r2c,
                                                   const int32_t* __restrict__ escape,
                                                   int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  constexpr int T = L / 8, ROWS = 256 / T, P = L + 2;
  __shared__ float2 buf[ROWS * P];
  const int row0 = blockIdx.x * ROWS;
  const float2* src = spec + (int64_t)row0 * sp;
  constexpr int PER = (ROWS * (L + 1) + 255) / 256;
  float2 v[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * 256;
    if (e < ROWS * (L + 1)) v[q] = src[(e / (L + 1)) * sp + e % (L + 1)];
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * 256;
    if (e < ROWS * (L + 1)) buf[(e / (L + 1)) * P + (e % (L + 1))] = v[q];
  }
  __syncthreads();
  const int col = threadIdx.x / T, t = threadIdx.x % T;
  float2 z[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int k = t + i * T;
    float2 xk = buf[col * P + k];
    float2 xc = conjf2(buf[col * P + (L - k)]);
    // Z'[k] = (Xk + Xc) + i W^-k (Xk - Xc), W = exp(-2 pi i / Nx)
    z[i] = (xk + xc) + rot90<true>(cmul(conjf2(tw_r2c[k]), xk - xc));
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 8; ++i) buf[col * P + t + i * T] = z[i];
  __syncthreads();
  fft_lds<L, true>(buf, P, 1, col, t, true, tw);
  float2* dst = reinterpret_cast<float2*>(phi) + (int64_t)row0 * L;
  for (int e = threadIdx.x; e < ROWS * L; e += 256) dst[e] = buf[(e / L) * P + (e % L)];
}

// Column tiles for the y and z passes: TW columns of kx per workgroup (rows
// of TW * 8 bytes: 256 B for TW = 32, 128 B for 16), the last tile also
// takes the Nyquist column kx = Nx/2, transformed after the others by the
// first L/8 threads.  TW = 32 up to L = 256 when Nx/2 >= 32, 16 up to 512,
// 8 for 1024: the LDS tile stays within 68 KB.
template <int L, int TW_>
struct ColTile {
  static constexpr int T = L / 8;
  static constexpr int TW = TW_;
  static constexpr int NC = TW + 1;
  static constexpr int THREADS = TW * T;
  static_assert(THREADS <= 1024, "one workgroup per column tile");
};

template <int L, int TW, bool INV>
__device__ __forceinline__ void col_fft(float2* buf, int ncols, const float2* tw) {
  using CT = ColTile<L, TW>;
  fft_lds<L, INV>(buf, 1, CT::NC, threadIdx.x % CT::TW, threadIdx.x / CT::TW, true, tw);
  if (ncols > CT::TW) fft_lds<L, INV>(buf, 1, CT::NC, CT::TW, threadIdx.x, threadIdx.x < CT::T, tw);
}

// y pass (forward or inverse), in place on the spectrum.
template <int L, int TW, bool INV>
__global__ __launch_bounds__(TW * (L / 8)) void k_fft_y(float2* __restrict__ spec, int sp,
                                                               int ntile,
                                                               const float2* __restrict__ tw,
                                                               const int32_t* __restrict__ escape,
                                                               int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  using CT = ColTile<L, TW>;
  __shared__ float2 buf[L * CT::NC];
  // A (plane, column tile) grid: consecutive workgroups take consecutive
  // planes of one column tile, so the ones running at once read rows spread
  // over many planes (and HBM channels), not the rows of one plane (512^3:
  // 0.25 ms per pass against 0.29).  Each pass also starts where the pass
  // before it ended, so its first reads can find that pass's last writes in
  // the caches: the forward pass (after the x pass, which ends at the last
  // planes) from the last plane down, the inverse pass (after the z pass,
  // which ends at the last column tiles) from the last column tile down
  // (512^3: 0.235 ms against 0.259, the spectrum twice the Infinity Cache).
  const int tile = INV ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;
  const int z = INV ? (int)blockIdx.x : (int)(gridDim.x - 1 - blockIdx.x);
  const int c0 = tile * CT::TW;
  const int ncols = tile == ntile - 1 ? CT::TW + 1 : CT::TW;
  float2* base = spec + (int64_t)z * L * sp + c0;
  // all of a thread's loads are issued before the first LDS store
  constexpr int PER = (L * CT::NC + CT::THREADS - 1) / CT::THREADS;
  float2 v[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * CT::THREADS;
    if (e < L * ncols) {
      const int n = e / ncols, col = e % ncols;
      v[q] = base[(int64_t)n * sp + col];
    }
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * CT::THREADS;
    if (e < L * ncols) buf[(e / ncols) * CT::NC + e % ncols] = v[q];
  }
  __syncthreads();
  col_fft<L, TW, INV>(buf, ncols, tw);
  for (int e = threadIdx.x; e < L * ncols; e += CT::THREADS) {
    const int n = e / ncols, col2 = e % ncols;
    base[(int64_t)n * sp + col2] = buf[n * CT::NC + col2];
  }
}

// Waves per SIMD the z pass is compiled for: eight (two 1024-thread
// workgroups per CU) where its registers fit in 64 without spilling (on
// 512^3: 0.44 ms per launch against 0.57 at one workgroup per CU).
template <int L, int TW>
constexpr int z_waves() {
  return (L == 512 && TW == 16) || (L == 256 && TW == 32) || (L == 64 && TW == 32) || (L == 32 && TW == 16) ? 8 : 1;
}

// z pass.  MODE Z_GREEN: forward along z, multiply by G, inverse along z
// (the force step: the pencil never leaves LDS).  Z_FWD: forward only.
// Z_MUL_INV: multiply src by G, inverse along z, write dst (src is kept,
// so a cached density spectrum can be reused with several G).
constexpr int Z_GREEN = 0, Z_FWD = 1, Z_MUL_INV = 2;

//
// A one-dimensional grid of ntile * ny blocks, y fastest: the workgroups
// running at once take neighbouring y of one column tile (512^3: 0.40 ms
// against 0.44 with the column tiles of one y side by side).  Blocks are
// dealt to the 8 XCDs round-robin (block id mod 8); with pair set (G in the
// spectrum's layout), the two tiles (2q, 2q + 1) at one y get ids 8 apart,
// so they run on one XCD, one after the other, and the 128-byte lines of G
// that their rows of TW = 16 floats share are fetched into that XCD's L2
// once (the spectrum's rows are whole lines).
//
// G is read in the spectrum's layout (gnxh = 0) or in the tile order of
// greens_z (gnxh = Nx/2 + 1): the tile at (y, c0) holds its L x ncols values
// contiguously from (y gnxh + c0) L, element e = n ncols + col, so a
// thread's G loads are consecutive with its spectrum loads' e.
template <int L, int TW, int MODE>
__global__ __launch_bounds__(TW * (L / 8)) __attribute__((amdgpu_waves_per_eu(z_waves<L, TW>()))) void k_fft_z(
    const float2* __restrict__ src, float2* __restrict__ dst, const float* __restrict__ greens, int sp,
    int ny, int ntile, int pair, int gnxh, const float2* __restrict__ tw, const int32_t* __restrict__ escape,
    int iteration) {
  iteration = gate_iteration(escape, iteration);
  if (escape[0] <= iteration) return;
  using CT = ColTile<L, TW>;
  __shared__ float2 buf[L * CT::NC];
  int tile, y;
  if (pair) {
    const int b = blockIdx.x, r = b >> 3, half = ntile >> 1;
    const int p = (r >> 1) * 8 + (b & 7);
    y = p / half;
    tile = 2 * (p - y * half) + (r & 1);
  } else {
    y = blockIdx.x % ny;
    tile = blockIdx.x / ny;
  }
  const int c0 = tile * CT::TW;
  const int ncols = tile == ntile - 1 ? CT::TW + 1 : CT::TW;
  const int64_t zstride = (int64_t)ny * sp;
  const int64_t off = (int64_t)y * sp + c0;
  // all of a thread's loads (and its G values) are issued before the first
  // LDS store
  constexpr int PER = (L * CT::NC + CT::THREADS - 1) / CT::THREADS;
  float2 v[PER];
  float g[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * CT::THREADS;
    if (e < L * ncols) {
      const int n = e / ncols, col = e % ncols;
      v[q] = src[off + n * zstride + col];
      if constexpr (MODE != Z_FWD)
        g[q] = gnxh ? greens[((int64_t)y * gnxh + c0) * L + e] : greens[off + n * zstride + col];
    }
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * CT::THREADS;
    if (e < L * ncols) {
      float2 w = v[q];
      if constexpr (MODE == Z_MUL_INV) w = scale2(w, g[q]);
      buf[(e / ncols) * CT::NC + e % ncols] = w;
    }
  }
  __syncthreads();
  if constexpr (MODE != Z_MUL_INV) col_fft<L, TW, false>(buf, ncols, tw);
  if constexpr (MODE == Z_GREEN) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = threadIdx.x + q * CT::THREADS;
      if (e < L * ncols) {
        const int i = (e / ncols) * CT::NC + e % ncols;
        buf[i] = scale2(buf[i], g[q]);
      }
    }
    __syncthreads();
  }
  if constexpr (MODE != Z_FWD) {
    // the inverse reads the forward transform's twiddles (conjugated): an
    // opaque copy of the pointer makes it load them again instead of keeping
    // the forward transform's values live across the multiply
    const float2* tw_inv = tw;
    asm volatile("" : "+s"(tw_inv));
    col_fft<L, TW, true>(buf, ncols, tw_inv);
  }
  for (int e = threadIdx.x; e < L * ncols; e += CT::THREADS) {
    const int n = e / ncols, col2 = e % ncols;
    dst[off + n * zstride + col2] = buf[n * CT::NC + col2];
  }
}

#define PMGPU_LAUNCH_X(KERNEL, L, grid, ...)                                               \
  switch (L) {                                                                             \
    case 16: hipLaunchKernelGGL(KERNEL<16>, grid, dim3(256), 0, __VA_ARGS__); break;       \
    case 32: hipLaunchKernelGGL(KERNEL<32>, grid, dim3(256), 0, __VA_ARGS__); break;       \
    case 64: hipLaunchKernelGGL(KERNEL<64>, grid, dim3(256), 0, __VA_ARGS__); break;       \
    case 128: hipLaunchKernelGGL(KERNEL<128>, grid, dim3(256), 0, __VA_ARGS__); break;     \
    case 256: hipLaunchKernelGGL(KERNEL<256>, grid, dim3(256), 0, __VA_ARGS__); break;     \
    case 512: hipLaunchKernelGGL(KERNEL<512>, grid, dim3(256), 0, __VA_ARGS__); break;     \
    default: fail(PMGPU_ERR_INVALID_ARGUMENT, "unsupported x length");                    \
  }

int rows_per_block(int L) {
  return 256 / (L / 8);
}

// the tile width for length L with Nx/2 columns (Nx >= 32, so 16 divides Nx/2)
constexpr int tile_width(int L) { return L <= 512 ? 16 : 8; }

template <int L, int TW, bool INV>
void launch_y_tw(pmgpu_ctx* c, float2* spec) {
  using CT = ColTile<L, TW>;
  const int ntile = (c->nx / 2) / CT::TW;
  hipLaunchKernelGGL((k_fft_y<L, TW, INV>), dim3(c->nz, ntile), dim3(CT::THREADS), 0, c->stream, spec,
                     c->sp, ntile, c->tw_y, c->escape, c->iter);
}

template <int L, bool INV>
void launch_y(pmgpu_ctx* c, float2* spec) {
  if constexpr (L <= 256) {
    if (c->nx / 2 >= 32) return launch_y_tw<L, 32, INV>(c, spec);
  }
  launch_y_tw<L, tile_width(L), INV>(c, spec);
}

template <int L, int TW, int MODE>
void launch_z_tw(pmgpu_ctx* c, const float2* src, float2* dst, const float* g, int gnxh) {
  using CT = ColTile<L, TW>;
  const int ntile = (c->nx / 2) / CT::TW;
  // the pairing (for G in the spectrum's layout) is a bijection of the block
  // ids when 8 divides the pairs
  const int pair = !gnxh && ntile % 2 == 0 && (ntile / 2 * c->ny) % 8 == 0;
  hipLaunchKernelGGL((k_fft_z<L, TW, MODE>), dim3(ntile * c->ny), dim3(CT::THREADS), 0, c->stream, src, dst,
                     g, c->sp, c->ny, ntile, pair, gnxh, c->tw_z, c->escape, c->iter);
}

// the z pass's column-tile width for nz = L
template <int L>
int z_tile_width(const pmgpu_ctx* c) {
  if constexpr (L <= 256) {
    if (c->nx / 2 >= 32) return 32;
  }
  return tile_width(L);
}

template <int L, int MODE>
void launch_z(pmgpu_ctx* c, const float2* src, float2* dst, const float* g, int gnxh) {
  if constexpr (L <= 256) {
    if (z_tile_width<L>(c) == 32) return launch_z_tw<L, 32, MODE>(c, src, dst, g, gnxh);
  }
  launch_z_tw<L, tile_width(L), MODE>(c, src, dst, g, gnxh);
}

// greens (spectrum layout, pitch sp) -> out in the z pass's tile order
__global__ void k_greens_block(const float* __restrict__ greens, int sp, int nx, int ny, int nz, int tw,
                               float* __restrict__ out) {
  const int nxh = nx / 2 + 1, ntile = (nx / 2) / tw;
  const int64_t n = (int64_t)nxh * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int kx = (int)(i % nxh), ky = (int)((i / nxh) % ny), kz = (int)(i / ((int64_t)nxh * ny));
  const int tile = min(kx / tw, ntile - 1), c0 = tile * tw, col = kx - c0;
  const int ncols = tile == ntile - 1 ? tw + 1 : tw;
  out[((int64_t)ky * nxh + c0) * nz + (int64_t)kz * ncols + col] = greens[((int64_t)kz * ny + ky) * sp + kx];
}

template <bool INV>
void y_pass(pmgpu_ctx* c, float2* spec) {
  switch (c->ny) {
    case 32: launch_y<32, INV>(c, spec); break;
    case 64: launch_y<64, INV>(c, spec); break;
    case 128: launch_y<128, INV>(c, spec); break;
    case 256: launch_y<256, INV>(c, spec); break;
    case 512: launch_y<512, INV>(c, spec); break;
    case 1024: launch_y<1024, INV>(c, spec); break;
    default: fail(PMGPU_ERR_INVALID_ARGUMENT, "unsupported ny");
  }
  HIP_CHECK(hipGetLastError());
}

template <int MODE>
void z_pass(pmgpu_ctx* c, const float2* src, float2* dst, const float* g, int gnxh = 0) {
  switch (c->nz) {
    case 32: launch_z<32, MODE>(c, src, dst, g, gnxh); break;
    case 64: launch_z<64, MODE>(c, src, dst, g, gnxh); break;
    case 128: launch_z<128, MODE>(c, src, dst, g, gnxh); break;
    case 256: launch_z<256, MODE>(c, src, dst, g, gnxh); break;
    case 512: launch_z<512, MODE>(c, src, dst, g, gnxh); break;
    case 1024: launch_z<1024, MODE>(c, src, dst, g, gnxh); break;
    default: fail(PMGPU_ERR_INVALID_ARGUMENT, "unsupported nz");
  }
  HIP_CHECK(hipGetLastError());
}

int z_width(const pmgpu_ctx* c) {
  switch (c->nz) {
    case 32: return z_tile_width<32>(c);
    case 64: return z_tile_width<64>(c);
    case 128: return z_tile_width<128>(c);
    case 256: return z_tile_width<256>(c);
    case 512: return z_tile_width<512>(c);
    case 1024: return z_tile_width<1024>(c);
    default: fail(PMGPU_ERR_INVALID_ARGUMENT, "unsupported nz");
  }
  return 0;
}

void x_r2c(pmgpu_ctx* c, const float* in, float2* spec) {
  const int L = c->nx / 2;
  const dim3 gx(c->ny * c->nz / rows_per_block(L));
  PMGPU_LAUNCH_X(k_fft_x_r2c, L, gx, c->stream, in, spec, c->sp, c->tw_x, c->tw_r2c, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
}

void x_c2r(pmgpu_ctx* c, const float2* spec, float* out) {
  const int L = c->nx / 2;
  const dim3 gx(c->ny * c->nz / rows_per_block(L));
  PMGPU_LAUNCH_X(k_fft_x_c2r, L, gx, c->stream, spec, c->sp, out, c->tw_x, c->tw_r2c, c->escape, c->iter);
  HIP_CHECK(hipGetLastError());
}

}  // namespace

void fft_forward_green_inverse(pmgpu_ctx* c) {
  if (c->general) return fft_general(c);
  int tok = stage_begin(c, "fft_x_r2c");
  x_r2c(c, c->rho, c->spec);
  stage_end(c, tok);
  tok = stage_begin(c, "fft_y_fwd");
  y_pass<false>(c, c->spec);
  stage_end(c, tok);
  tok = stage_begin(c, "fft_z_green");
  if (c->greens_z_valid) z_pass<Z_GREEN>(c, c->spec, c->spec, c->greens_z, c->nx / 2 + 1);
  else z_pass<Z_GREEN>(c, c->spec, c->spec, c->greens);
  stage_end(c, tok);
  tok = stage_begin(c, "fft_y_inv");
  y_pass<true>(c, c->spec);
  stage_end(c, tok);
  tok = stage_begin(c, "fft_x_c2r");
  x_c2r(c, c->spec, c->phi);
  stage_end(c, tok);
}

void greens_block_for_z(pmgpu_ctx* c) {
  const int64_t n = (int64_t)(c->nx / 2 + 1) * c->ny * c->nz;
  hipLaunchKernelGGL(k_greens_block, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->greens, c->sp,
                     c->nx, c->ny, c->nz, z_width(c), c->greens_z);
  HIP_CHECK(hipGetLastError());
  c->greens_z_valid = true;
}

void fft_r2c_forward(pmgpu_ctx* c, const float* in, float2* spec) {
  x_r2c(c, in, spec);
  y_pass<false>(c, spec);
  z_pass<Z_FWD>(c, spec, spec, nullptr);
}

void fft_mul_c2r_inverse(pmgpu_ctx* c, const float2* src, const float* g, float2* work, float* out) {
  z_pass<Z_MUL_INV>(c, src, work, g);
  y_pass<true>(c, work);
  x_c2r(c, work, out);
}

}  // namespace pmgpu
