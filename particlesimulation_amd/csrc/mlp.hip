// Learned influence-function correction: a 4 -> 64 -> 64 -> 64 -> 1 MLP with
// GELU activations that maps a mode k of the half spectrum to a factor
// (1 + y(k)) multiplying the analytic G(k).  There is no counterpart in the
// reference; it attaches where PMMethod::initGreensFunction fills G
// (source/pmMethod.cpp:164-185).
//
// Inputs per mode: |kx|/(Nx/2), |ky|/(Ny/2), |kz|/(Nz/2) with ky, kz folded to
// [0, N/2], and the assignment order minus 2 (NGP -1, CIC 0, TSC +1).  The
// folded magnitudes make y(k) = y(-k) (a real potential) and give it the
// lattice's reflection symmetry.
//
// The two 64x64 layers run on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation; the 4-wide input and the scalar output layer run on VALU.
// One wavefront owns 16 modes at a time.  The layers are written transposed,
// Z^T = W H^T, so that the accumulator layout of one layer (rows = features
// 4*(lane/16)+r, column = mode lane%16) is exactly the B-operand layout the
// next layer needs, and activations never leave registers in the forward
// pass.  The backward pass computes parameter gradients as MFMA products
// dZ * [H^T | 1] whose reduction index is the mode; those operands are
// transposed through LDS.  Each wave keeps its gradient in accumulators and
// writes it once; a second kernel adds the per-wave gradients in wave order,
// so the gradient is bitwise reproducible.
#include "pmgpu_internal.h"

#include <cmath>

namespace pmgpu {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int HID = 64;
constexpr int OFF_W1 = 0, OFF_B1 = 256, OFF_W2 = 320, OFF_B2 = 4416, OFF_W3 = 4480, OFF_B3 = 8576,
              OFF_W4 = 8640, OFF_B4 = 8704, NPARAM = 8705;
constexpr int MLP_THREADS = 256;  // 4 waves
constexpr int MLP_MAX_BLOCKS = 512;

// Weights rearranged for the kernels (device memory): the four 64x64 A
// operands (W2, W3 for the forward pass, W2^T, W3^T for the backward pass)
// as 16 tiles x 64 lanes x 4 bf16 in lane order, and the FP32 rest.
struct Packed {
  bf16x4 a[4][16][64];
  float w1[256], b1[64], b2[64], b3[64], w4[64], b4;
};
constexpr int A_W2 = 0, A_W3 = 1, A_W2T = 2, A_W3T = 3;

__device__ __forceinline__ short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__global__ void k_mlp_pack(const float* __restrict__ p, Packed* __restrict__ out) {
  for (int e = threadIdx.x; e < 4 * 16 * 64 * 4; e += blockDim.x) {
    const int item = e & 3, lane = (e >> 2) & 63, tile = (e >> 8) & 15, which = e >> 12;
    const int t = tile >> 2, s = tile & 3;
    const int i16 = lane & 15, g4 = 4 * (lane >> 4) + item;
    float v;
    if (which == A_W2 || which == A_W3) {
      // A[i = out feature 16t+i16][k = in feature 16s+g4]
      const float* W = p + (which == A_W2 ? OFF_W2 : OFF_W3);
      v = W[(16 * t + i16) * HID + 16 * s + g4];
    } else {
      // (W^T)[i = in feature 16t+i16][k = out feature 16s+g4]
      const float* W = p + (which == A_W2T ? OFF_W2 : OFF_W3);
      v = W[(16 * s + g4) * HID + 16 * t + i16];
    }
    out->a[which][tile][lane][item] = f2bf(v);
  }
  for (int e = threadIdx.x; e < 256; e += blockDim.x) out->w1[e] = p[OFF_W1 + e];
  for (int e = threadIdx.x; e < 64; e += blockDim.x) {
    out->b1[e] = p[OFF_B1 + e];
    out->b2[e] = p[OFF_B2 + e];
    out->b3[e] = p[OFF_B3 + e];
    out->w4[e] = p[OFF_W4 + e];
  }
  if (threadIdx.x == 0) out->b4 = p[OFF_B4];
}

struct Shared {
  bf16x4 a[4][16][64];
  float w1[256], b1[64], b2[64], b3[64], w4[64], b4;
};

__device__ void load_weights(Shared& sh, const Packed* __restrict__ pk, int nmat) {
  const bf16x4* src = &pk->a[0][0][0];
  bf16x4* dst = &sh.a[0][0][0];
  for (int e = threadIdx.x; e < nmat * 16 * 64; e += blockDim.x) dst[e] = src[e];
  for (int e = threadIdx.x; e < 256; e += blockDim.x) sh.w1[e] = pk->w1[e];
  for (int e = threadIdx.x; e < 64; e += blockDim.x) {
    sh.b1[e] = pk->b1[e];
    sh.b2[e] = pk->b2[e];
    sh.b3[e] = pk->b3[e];
    sh.w4[e] = pk->w4[e];
  }
  if (threadIdx.x == 0) sh.b4 = pk->b4;
  __syncthreads();
}

// Input features of one mode.
struct Inputs {
  const float* list;  // n x 4, or null: modes of the half spectrum
  int nx, ny, nz, pitch;  // pitch: entries per row of the half spectrum (>= nx/2 + 1)
  float order;
};

__device__ __forceinline__ void mode_inputs(const Inputs& in, int64_t m, int64_t n, float x[4]) {
  if (m >= n) {
    x[0] = x[1] = x[2] = x[3] = 0.0f;
    return;
  }
  if (in.list) {
    x[0] = in.list[4 * m];
    x[1] = in.list[4 * m + 1];
    x[2] = in.list[4 * m + 2];
    x[3] = in.list[4 * m + 3];
    return;
  }
  const int kx = (int)(m % in.pitch);  // row padding (kx > nx/2) carries G = 0
  const int y = (int)((m / in.pitch) % in.ny);
  const int z = (int)(m / ((int64_t)in.pitch * in.ny));
  x[0] = (float)kx / (float)(in.nx / 2);
  x[1] = (float)min(y, in.ny - y) / (float)(in.ny / 2);
  x[2] = (float)min(z, in.nz - z) / (float)(in.nz / 2);
  x[3] = in.order;
}

// Forward pass of one 16-mode tile.  Lane l computes mode l%16; zK[t][r] is
// the pre-activation of feature 16t + 4(l/16) + r of layer K.
struct Acts {
  f32x4 z1[4], z2[4], z3[4];
  bf16x4 h1[4], h2[4], h3[4];
  float y;
};

__device__ __forceinline__ void forward_tile(const Shared& sh, const float x[4], Acts& a) {
  const int lane = threadIdx.x & 63, g = 4 * (lane >> 4);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int f = 16 * s + g + r;
      float z = sh.b1[f];
      z = fmaf(sh.w1[4 * f], x[0], z);
      z = fmaf(sh.w1[4 * f + 1], x[1], z);
      z = fmaf(sh.w1[4 * f + 2], x[2], z);
      z = fmaf(sh.w1[4 * f + 3], x[3], z);
      a.z1[s][r] = z;
      a.h1[s][r] = f2bf(gelu(z));
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = sh.b2[16 * t + g + r];
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = mfma(sh.a[A_W2][4 * t + s][lane], a.h1[s], acc);
    a.z2[t] = acc;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) a.h2[t][r] = f2bf(gelu(a.z2[t][r]));
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = sh.b3[16 * t + g + r];
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = mfma(sh.a[A_W3][4 * t + s][lane], a.h2[s], acc);
    a.z3[t] = acc;
  }
  float part = 0.0f;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float h = gelu(a.z3[t][r]);
      a.h3[t][r] = f2bf(h);
      part = fmaf(sh.w4[16 * t + g + r], h, part);
    }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  a.y = part + sh.b4;
}

// out[m] = y(m), or greens[m] * (1 + y(m)) when greens is given.
__global__ __launch_bounds__(MLP_THREADS) void k_mlp_forward(const Packed* __restrict__ pk, Inputs in,
                                                             int64_t n, const float* greens, float* out) {
  __shared__ Shared sh;
  load_weights(sh, pk, 2);
  const int lane = threadIdx.x & 63;
  const int64_t ntiles = (n + 15) / 16;
  const int64_t wave = (int64_t)blockIdx.x * (MLP_THREADS / 64) + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * (MLP_THREADS / 64);
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    const int64_t m = tile * 16 + (lane & 15);
    float x[4];
    mode_inputs(in, m, n, x);
    Acts a;
    forward_tile(sh, x, a);
    if (lane < 16 && m < n) out[m] = greens ? greens[m] * (1.0f + a.y) : a.y;
  }
}

// Per-wave LDS for the transposes of the backward pass: five 64 x 16 bf16
// matrices [feature][mode] and the inputs [mode][4].
struct WaveScratch {
  short dz3[64 * 16], h2[64 * 16], dz2[64 * 16], h1[64 * 16], dz1[64 * 16];
  float x[16 * 4];
};

// store a D-layout matrix (rows 16t + 4(l/16) + r, column = mode l%16)
__device__ __forceinline__ void store_t(short* dst, const bf16x4 v[4]) {
  const int lane = threadIdx.x & 63, g = 4 * (lane >> 4), m = lane & 15;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) dst[(16 * t + g + r) * 16 + m] = v[t][r];
}

// A/B operand whose reduction index is the mode: element (feature 16t + l%16,
// modes 4(l/16) .. +3)
__device__ __forceinline__ bf16x4 load_t(const short* src, int t) {
  const int lane = threadIdx.x & 63;
  const short* p = src + (16 * t + (lane & 15)) * 16 + 4 * (lane >> 4);
  bf16x4 v;
  v[0] = p[0];
  v[1] = p[1];
  v[2] = p[2];
  v[3] = p[3];
  return v;
}

struct Grads {
  f32x4 w3[4][4], b3[4], w2[4][4], b2[4], w1b1[4];
  float w4[4][4], b4;
};

// Parameter gradients for sum_m up[m] * y(m): one partial gradient per wave
// (NPARAM floats at partial + wave * NPARAM).  The tile loop is uniform over
// the block so the workgroup barriers around the LDS transposes are legal.
__global__ __launch_bounds__(MLP_THREADS) void k_mlp_backward(const Packed* __restrict__ pk, Inputs in,
                                                              int64_t n, const float* __restrict__ up,
                                                              float* __restrict__ partial) {
  __shared__ Shared sh;
  __shared__ WaveScratch ws[MLP_THREADS / 64];
  load_weights(sh, pk, 4);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = 4 * (lane >> 4), m16 = lane & 15;
  WaveScratch& S = ws[w];
  const int64_t ntiles = (n + 15) / 16;
  const int64_t iters = (ntiles + (int64_t)gridDim.x * 4 - 1) / ((int64_t)gridDim.x * 4);
  Grads G;
  const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    G.b3[i] = G.b2[i] = G.w1b1[i] = zero4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      G.w3[i][j] = G.w2[i][j] = zero4;
      G.w4[i][j] = 0.0f;
    }
  }
  G.b4 = 0.0f;
  // B operand of the bias columns: column 0 is all ones over the modes
  bf16x4 ones;
#pragma unroll
  for (int r = 0; r < 4; ++r) ones[r] = m16 == 0 ? (short)0x3f80 : (short)0;

  for (int64_t it = 0; it < iters; ++it) {
    const int64_t tile = ((int64_t)blockIdx.x + it * gridDim.x) * 4 + w;
    const int64_t m = tile * 16 + m16;
    float x[4];
    mode_inputs(in, m, n, x);
    Acts a;
    forward_tile(sh, x, a);
    const float d = (m < n) ? up[m] : 0.0f;
    if (lane < 16) G.b4 += d;

    // layer 4 and 3
    bf16x4 dz3b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int f = 16 * t + g + r;
        // h3 is needed in FP32 for dW4; recompute from z3 (bf16 h3 is the MFMA copy)
        G.w4[t][r] = fmaf(d, gelu(a.z3[t][r]), G.w4[t][r]);
        dz3b[t][r] = f2bf(sh.w4[f] * d * gelu_d(a.z3[t][r]));
      }
    store_t(S.dz3, dz3b);
    store_t(S.h2, a.h2);
    if (lane < 16) {
#pragma unroll
      for (int i = 0; i < 4; ++i) S.x[4 * m16 + i] = x[i];
    }
    // dh2 = W3^T dz3 (B operand = dz3 in accumulator layout)
    bf16x4 dz2b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = zero4;
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mfma(sh.a[A_W3T][4 * t + s][lane], dz3b[s], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) dz2b[t][r] = f2bf(acc[r] * gelu_d(a.z2[t][r]));
    }
    store_t(S.dz2, dz2b);
    store_t(S.h1, a.h1);
    bf16x4 dz1b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = zero4;
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mfma(sh.a[A_W2T][4 * t + s][lane], dz2b[s], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) dz1b[t][r] = f2bf(acc[r] * gelu_d(a.z1[t][r]));
    }
    store_t(S.dz1, dz1b);
    __syncthreads();

    // parameter gradients: dW = dZ [H^T | 1] with the mode as reduction index
    bf16x4 hB[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) hB[j] = load_t(S.h2, j);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bf16x4 dA = load_t(S.dz3, i);
#pragma unroll
      for (int j = 0; j < 4; ++j) G.w3[i][j] = mfma(dA, hB[j], G.w3[i][j]);
      G.b3[i] = mfma(dA, ones, G.b3[i]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) hB[j] = load_t(S.h1, j);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bf16x4 dA = load_t(S.dz2, i);
#pragma unroll
      for (int j = 0; j < 4; ++j) G.w2[i][j] = mfma(dA, hB[j], G.w2[i][j]);
      G.b2[i] = mfma(dA, ones, G.b2[i]);
    }
    // layer 1: B columns 0..3 = inputs, 4 = ones
    bf16x4 xB;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mm = 4 * (lane >> 4) + r;
      xB[r] = m16 < 4 ? f2bf(S.x[4 * mm + m16]) : (m16 == 4 ? (short)0x3f80 : (short)0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) G.w1b1[i] = mfma(load_t(S.dz1, i), xB, G.w1b1[i]);
    __syncthreads();
  }

  // write this wave's gradient
  float* P = partial + ((int64_t)blockIdx.x * 4 + w) * NPARAM;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * i + g + r;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        P[OFF_W3 + row * HID + 16 * j + m16] = G.w3[i][j][r];
        P[OFF_W2 + row * HID + 16 * j + m16] = G.w2[i][j][r];
      }
      if (m16 == 0) {
        P[OFF_B3 + row] = G.b3[i][r];
        P[OFF_B2 + row] = G.b2[i][r];
      }
      if (m16 < 4) P[OFF_W1 + row * 4 + m16] = G.w1b1[i][r];
      if (m16 == 4) P[OFF_B1 + row] = G.w1b1[i][r];
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = G.w4[t][r];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      if (m16 == 0) P[OFF_W4 + 16 * t + g + r] = v;
    }
  float b4 = G.b4;
  for (int o = 1; o < 64; o <<= 1) b4 += __shfl_xor(b4, o, 64);
  if (lane == 0) P[OFF_B4] = b4;
}

// grad[p] (+)= scale * sum over waves of partial[wave][p], in wave order.
__global__ void k_mlp_reduce(const float* __restrict__ partial, int nwaves, float scale, int accumulate,
                             float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float s = 0.0f;
  for (int w = 0; w < nwaves; ++w) s += partial[(int64_t)w * NPARAM + p];
  grad[p] = accumulate ? grad[p] + scale * s : scale * s;
}

int mlp_blocks(int64_t n) {
  const int64_t tiles = (n + 15) / 16;
  const int64_t b = (tiles + 3) / 4;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, MLP_MAX_BLOCKS));
}

}  // namespace

// Device-side MLP state: packed weights and the partial-gradient scratch.
struct MlpDevice {
  Packed* packed = nullptr;
  float* partial = nullptr;
  int partial_waves = 0;
};

MlpDevice* mlp_device_create() {
  MlpDevice* d = new MlpDevice();
  HIP_CHECK(hipMalloc((void**)&d->packed, sizeof(Packed)));
  return d;
}

void mlp_device_destroy(MlpDevice* d) {
  if (!d) return;
  (void)hipFree(d->packed);
  (void)hipFree(d->partial);
  delete d;
}

void mlp_pack(MlpDevice* d, const float* params_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(256), 0, s, params_dev, d->packed);
  HIP_CHECK(hipGetLastError());
}

static Inputs grid_inputs(int nx, int ny, int nz, int pitch, int scheme) {
  Inputs in{};
  in.list = nullptr;
  in.nx = nx;
  in.ny = ny;
  in.nz = nz;
  in.pitch = pitch;
  in.order = (float)(scheme - PMGPU_CIC);  // NGP -1, CIC 0, TSC +1
  return in;
}

void mlp_forward_grid(MlpDevice* d, int nx, int ny, int nz, int pitch, int scheme, const float* greens,
                      float* out, hipStream_t s) {
  const int64_t K = (int64_t)pitch * ny * nz;
  hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_blocks(K)), dim3(MLP_THREADS), 0, s, d->packed,
                     grid_inputs(nx, ny, nz, pitch, scheme), K, greens, out);
  HIP_CHECK(hipGetLastError());
}

static void backward(MlpDevice* d, Inputs in, int64_t n, const float* up, float scale, bool accumulate,
                     float* grad, hipStream_t s) {
  const int nb = mlp_blocks(n);
  if (d->partial_waves < nb * 4) {
    HIP_CHECK(hipStreamSynchronize(s));
    (void)hipFree(d->partial);
    d->partial = nullptr;
    HIP_CHECK(hipMalloc((void**)&d->partial, sizeof(float) * (size_t)nb * 4 * NPARAM));
    d->partial_waves = nb * 4;
  }
  hipLaunchKernelGGL(k_mlp_backward, dim3(nb), dim3(MLP_THREADS), 0, s, d->packed, in, n, up, d->partial);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0, s, d->partial, nb * 4, scale,
                     accumulate ? 1 : 0, grad);
  HIP_CHECK(hipGetLastError());
}

void mlp_backward_grid(MlpDevice* d, int nx, int ny, int nz, int pitch, int scheme, const float* up, float scale,
                       bool accumulate, float* grad, hipStream_t s) {
  const int64_t K = (int64_t)pitch * ny * nz;
  backward(d, grid_inputs(nx, ny, nz, pitch, scheme), K, up, scale, accumulate, grad, s);
}

// The factor grid for pmgpu_init_greens: greens[k] *= (1 + y(k)).
void mlp_apply_to_greens(pmgpu_ctx* c) {
  float* p = nullptr;
  MlpDevice* d = mlp_device_create();
  try {
    HIP_CHECK(hipMalloc((void**)&p, sizeof(float) * NPARAM));
    HIP_CHECK(hipMemcpyAsync(p, c->mlp_weights.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice,
                             c->stream));
    mlp_pack(d, p, c->stream);
    mlp_forward_grid(d, c->nx, c->ny, c->nz, c->sp, c->prm.interpolation, c->greens, c->greens, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  } catch (...) {
    (void)hipFree(p);
    mlp_device_destroy(d);
    throw;
  }
  (void)hipFree(p);
  mlp_device_destroy(d);
}

}  // namespace pmgpu

using namespace pmgpu;

extern "C" {

int64_t pmgpu_mlp_num_params(void) { return NPARAM; }

int pmgpu_set_greens_correction(pmgpu_ctx* c, const float* weights) {
  if (!c) return PMGPU_ERR_INVALID_ARGUMENT;
  if (!weights) {
    c->mlp_weights.clear();
    return PMGPU_OK;
  }
  c->mlp_weights.assign(weights, weights + NPARAM);
  return PMGPU_OK;
}

// Standalone evaluation on host arrays (inputs n x 4): out[m] = y(m), and the
// gradient of sum_m up[m] * y(m) with respect to the parameters.
int pmgpu_mlp_forward(const float* params, int64_t n, const float* inputs, float* out) {
  if (!params || !inputs || !out || n < 0) return PMGPU_ERR_INVALID_ARGUMENT;
  if (n == 0) return PMGPU_OK;
  float *dp = nullptr, *din = nullptr, *dout = nullptr;
  MlpDevice* d = nullptr;
  int rc = PMGPU_OK;
  try {
    d = mlp_device_create();
    HIP_CHECK(hipMalloc((void**)&dp, sizeof(float) * NPARAM));
    HIP_CHECK(hipMalloc((void**)&din, sizeof(float) * 4 * n));
    HIP_CHECK(hipMalloc((void**)&dout, sizeof(float) * n));
    HIP_CHECK(hipMemcpy(dp, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(din, inputs, sizeof(float) * 4 * n, hipMemcpyHostToDevice));
    mlp_pack(d, dp, nullptr);
    Inputs in{};
    in.list = din;
    hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_blocks(n)), dim3(MLP_THREADS), 0, nullptr, d->packed, in, n,
                       (const float*)nullptr, dout);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(out, dout, sizeof(float) * n, hipMemcpyDeviceToHost));
  } catch (const Error& e) {
    rc = e.code;
  }
  (void)hipFree(dp);
  (void)hipFree(din);
  (void)hipFree(dout);
  mlp_device_destroy(d);
  return rc;
}

int pmgpu_mlp_backward(const float* params, int64_t n, const float* inputs, const float* up, float* grad) {
  if (!params || !inputs || !up || !grad || n < 0) return PMGPU_ERR_INVALID_ARGUMENT;
  float *dp = nullptr, *din = nullptr, *dup = nullptr, *dg = nullptr;
  MlpDevice* d = nullptr;
  int rc = PMGPU_OK;
  try {
    d = mlp_device_create();
    HIP_CHECK(hipMalloc((void**)&dp, sizeof(float) * NPARAM));
    HIP_CHECK(hipMalloc((void**)&din, sizeof(float) * 4 * std::max<int64_t>(n, 1)));
    HIP_CHECK(hipMalloc((void**)&dup, sizeof(float) * std::max<int64_t>(n, 1)));
    HIP_CHECK(hipMalloc((void**)&dg, sizeof(float) * NPARAM));
    HIP_CHECK(hipMemcpy(dp, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    if (n > 0) {
      HIP_CHECK(hipMemcpy(din, inputs, sizeof(float) * 4 * n, hipMemcpyHostToDevice));
      HIP_CHECK(hipMemcpy(dup, up, sizeof(float) * n, hipMemcpyHostToDevice));
    }
    mlp_pack(d, dp, nullptr);
    Inputs in{};
    in.list = din;
    backward(d, in, n, dup, 1.0f, false, dg, nullptr);
    HIP_CHECK(hipMemcpy(grad, dg, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  } catch (const Error& e) {
    rc = e.code;
  }
  (void)hipFree(dp);
  (void)hipFree(din);
  (void)hipFree(dup);
  (void)hipFree(dg);
  mlp_device_destroy(d);
  return rc;
}

}  // extern "C"
