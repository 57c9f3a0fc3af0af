// Training of the learned influence-function correction on the GPU, and the
// direct-sum force used to make its targets.  Neither has a counterpart in
// the reference; the loss is the reference's own accuracy measure, the
// squared error of the PM force against the direct sum
// (script/pm_accuracy.py:44-48), normalised: e_rms^2 = sum|a - t|^2 / sum|t|^2.
//
// A snapshot is a particle set with some of its particles chosen as targets
// and their direct-sum accelerations.  Its density spectrum rho^ is computed
// once (deposit + forward R2C) and kept in HBM.  With g(k) = G_sym(k)/M (1 +
// y(k)) the PM acceleration at a target is a_i = sum_x W_i(x) E(x),
// E = -D phi, phi = C2R(rho^ g).  The gradient of the loss with respect to
// the factor c(k) = 1 + y(k) follows from the adjoint of that chain:
//   r_i = dL/da_i = 2 w (a_i - t_i)              (w = 1 / sum|t|^2)
//   R_c(x) = sum_i W_i(x) r_ic                    (transpose of interpolation)
//   lambda(x) = sum_c (D_c R_c)(x)                (transpose of -grad, D antisymmetric)
//   dL/dc(k) = mult(kx) G_sym(k)/M Re(rho^(k) conj(lambda^(k)))
// with lambda^ the forward R2C of lambda and mult 2 for 0 < kx < Nx/2 (the
// mode and its mirror), else 1.  The MLP backward pass then turns dL/dc into
// parameter gradients, and Adam updates the parameters on the device.
//
// Each target's stencil (cells and weights, the deposit's floor-anchored
// weights) is built once on the host in two orders: per target for the
// forward interpolation and per cell (CSR) for the transpose, so both are
// gathers and the gradient is deterministic.
#include "pmgpu_internal.h"

#include <algorithm>
#include <cstring>
#include <atomic>
#include <cmath>
#include <array>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>

namespace pmgpu {

namespace {

constexpr float kGc = 0.079577471545947668f;  // 1/(4 pi)

struct Snap {
  float2* rho_hat = nullptr;   // K
  int64_t nt = 0;
  int32_t* tcell = nullptr;    // nt * S
  float* tw = nullptr;         // nt * S
  float4* tacc = nullptr;      // nt target accelerations
  float4* resid = nullptr;     // nt dL/da
  float* sq = nullptr;         // nt |a - t|^2
  int64_t ncell = 0;
  int32_t* cell_id = nullptr;     // ncell
  int32_t* cell_start = nullptr;  // ncell + 1
  int32_t* ent_target = nullptr;  // nt * S
  float* ent_w = nullptr;         // nt * S
  double tnorm = 0.0;             // sum |t|^2
};

template <typename T>
void dmalloc(T** p, int64_t n) {
  HIP_CHECK(hipMalloc((void**)p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

__device__ __forceinline__ int wrapc(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}

// E = -D phi at one cell, 2- or 4-point (PMMethod::getFieldInCell,
// source/pmMethod.cpp:355-365), periodic.
template <int FDS>
__device__ __forceinline__ float3 field_at(const float* __restrict__ phi, int x, int y, int z, int nx, int ny,
                                           int nz) {
  auto P = [&](int a, int b, int c) {
    return phi[(int64_t)wrapc(a, nx) + (int64_t)nx * (wrapc(b, ny) + (int64_t)ny * wrapc(c, nz))];
  };
  float3 e;
  if constexpr (FDS == PMGPU_TWO_POINT) {
    e.x = -0.5f * (P(x + 1, y, z) - P(x - 1, y, z));
    e.y = -0.5f * (P(x, y + 1, z) - P(x, y - 1, z));
    e.z = -0.5f * (P(x, y, z + 1) - P(x, y, z - 1));
  } else {
    e.x = (-1.0f / 12) * (-P(x + 2, y, z) + 8 * P(x + 1, y, z) - 8 * P(x - 1, y, z) + P(x - 2, y, z));
    e.y = (-1.0f / 12) * (-P(x, y + 2, z) + 8 * P(x, y + 1, z) - 8 * P(x, y - 1, z) + P(x, y - 2, z));
    e.z = (-1.0f / 12) * (-P(x, y, z + 2) + 8 * P(x, y, z + 1) - 8 * P(x, y, z - 1) + P(x, y, z - 2));
  }
  return e;
}

template <int FDS>
__global__ void k_target_forward(const float* __restrict__ phi, int nx, int ny, int nz,
                                 const int32_t* __restrict__ tcell, const float* __restrict__ tw, int S,
                                 const float4* __restrict__ tacc, int64_t nt, float two_w,
                                 float4* __restrict__ resid, float* __restrict__ sq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (int e = 0; e < S; ++e) {
    const int cell = tcell[i * S + e];
    const float w = tw[i * S + e];
    const int x = cell % nx, y = (cell / nx) % ny, z = cell / (nx * ny);
    const float3 E = field_at<FDS>(phi, x, y, z, nx, ny, nz);
    ax = fmaf(w, E.x, ax);
    ay = fmaf(w, E.y, ay);
    az = fmaf(w, E.z, az);
  }
  const float4 t = tacc[i];
  const float dx = ax - t.x, dy = ay - t.y, dz = az - t.z;
  resid[i] = make_float4(two_w * dx, two_w * dy, two_w * dz, 0.f);
  sq[i] = dx * dx + dy * dy + dz * dz;
}

// R_c at the touched cells (gather over the CSR).
__global__ void k_adj_gather(const int32_t* __restrict__ cell_id, const int32_t* __restrict__ cell_start,
                             const int32_t* __restrict__ ent_target, const float* __restrict__ ent_w,
                             int64_t ncell, const float4* __restrict__ resid, int64_t M,
                             float* __restrict__ R) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncell) return;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int e = cell_start[k]; e < cell_start[k + 1]; ++e) {
    const float w = ent_w[e];
    const float4 r = resid[ent_target[e]];
    sx = fmaf(w, r.x, sx);
    sy = fmaf(w, r.y, sy);
    sz = fmaf(w, r.z, sz);
  }
  const int64_t c = cell_id[k];
  R[c] = sx;
  R[M + c] = sy;
  R[2 * M + c] = sz;
}

__global__ void k_adj_clear(const int32_t* __restrict__ cell_id, int64_t ncell, int64_t M,
                            float* __restrict__ R) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncell) return;
  const int64_t c = cell_id[k];
  R[c] = 0.f;
  R[M + c] = 0.f;
  R[2 * M + c] = 0.f;
}

// lambda = sum_c D_c R_c with the same stencil as E = -D phi.
template <int FDS>
__global__ void k_adj_div(const float* __restrict__ R, int nx, int ny, int nz, float* __restrict__ lam) {
  const int64_t M = (int64_t)nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
  auto Q = [&](int c, int a, int b, int d) {
    return R[(int64_t)c * M + (int64_t)wrapc(a, nx) + (int64_t)nx * (wrapc(b, ny) + (int64_t)ny * wrapc(d, nz))];
  };
  float v;
  if constexpr (FDS == PMGPU_TWO_POINT) {
    v = 0.5f * (Q(0, x + 1, y, z) - Q(0, x - 1, y, z)) + 0.5f * (Q(1, x, y + 1, z) - Q(1, x, y - 1, z)) +
        0.5f * (Q(2, x, y, z + 1) - Q(2, x, y, z - 1));
  } else {
    v = (1.0f / 12) * (-Q(0, x + 2, y, z) + 8 * Q(0, x + 1, y, z) - 8 * Q(0, x - 1, y, z) + Q(0, x - 2, y, z)) +
        (1.0f / 12) * (-Q(1, x, y + 2, z) + 8 * Q(1, x, y + 1, z) - 8 * Q(1, x, y - 1, z) + Q(1, x, y - 2, z)) +
        (1.0f / 12) * (-Q(2, x, y, z + 2) + 8 * Q(2, x, y, z + 1) - 8 * Q(2, x, y, z - 1) + Q(2, x, y, z - 2));
  }
  lam[i] = v;
}

__global__ void k_grad_acc(const float2* __restrict__ rho_hat, const float2* __restrict__ lam_hat,
                           const float* __restrict__ greens, int64_t K, int sp, int nx,
                           float* __restrict__ gradc) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const int kx = (int)(k % sp);
  if (kx > nx / 2) return;  // row padding: never written by the transforms
  const float mult = (kx == 0 || kx == nx / 2) ? 1.0f : 2.0f;
  const float2 a = rho_hat[k], b = lam_hat[k];
  gradc[k] += mult * greens[k] * (a.x * b.x + a.y * b.y);
}

__global__ void k_fill(float* __restrict__ p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// out[slot] = w * sum sq, one block, fixed order.
__global__ void k_loss_sum(const float* __restrict__ sq, int64_t n, double w, double* __restrict__ out,
                           int slot) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += (double)sq[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[slot] = w * sh[0];
}

__global__ void k_axpy_add(float* __restrict__ a, const float* __restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Direct sum a_i = -G' sum_j m_j d_ij / (|d_ij|^2 + eps^2)^(3/2), d_ij = x_i - x_j,
// skipping coincident particles, FP32 with compensated (Neumaier) sums.
// PERIODIC: d_ij is the nearest image in the box and the Ewald correction
// (the periodic force minus the nearest-image Newtonian force, for a unit
// mass on a uniform neutralising background) is added by trilinear
// interpolation in a table over the positive octant [0, L/2]^3; the
// correction is odd in the matching component and even in the others.
constexpr int DS_THREADS = 256;
constexpr int EW_N = 64;  // table intervals per half box side

__device__ __forceinline__ void neumaier(float& s, float& c, float v) {
  const float t = s + v;
  c += fabsf(s) >= fabsf(v) ? (s - t) + v : (v - t) + s;
  s = t;
}

struct Box {
  float L[3], invL[3], tab_scale[3];  // tab_scale = EW_N / (L/2)
};

__device__ __forceinline__ float nearest(float d, float L, float invL) { return d - L * rintf(d * invL); }

__device__ __forceinline__ float3 ewald_corr(const float4* __restrict__ tab, const Box& b, float dx, float dy,
                                             float dz) {
  const float ux = fminf(fabsf(dx) * b.tab_scale[0], (float)EW_N - 1e-3f);
  const float uy = fminf(fabsf(dy) * b.tab_scale[1], (float)EW_N - 1e-3f);
  const float uz = fminf(fabsf(dz) * b.tab_scale[2], (float)EW_N - 1e-3f);
  const int ix = (int)ux, iy = (int)uy, iz = (int)uz;
  const float fx = ux - ix, fy = uy - iy, fz = uz - iz;
  constexpr int S1 = EW_N + 1, S2 = (EW_N + 1) * (EW_N + 1);
  const float4* t = tab + ix + S1 * iy + S2 * iz;
  auto lerp4 = [](float4 a, float4 c, float f) {
    return make_float4(fmaf(f, c.x - a.x, a.x), fmaf(f, c.y - a.y, a.y), fmaf(f, c.z - a.z, a.z), 0.f);
  };
  const float4 c00 = lerp4(t[0], t[1], fx), c10 = lerp4(t[S1], t[S1 + 1], fx);
  const float4 c01 = lerp4(t[S2], t[S2 + 1], fx), c11 = lerp4(t[S2 + S1], t[S2 + S1 + 1], fx);
  const float4 c0 = lerp4(c00, c10, fy), c1 = lerp4(c01, c11, fy);
  const float4 c = lerp4(c0, c1, fz);
  // odd in its own coordinate: c(-d) = -c(d) componentwise (the tabulated value may be negative)
  return make_float3(dx < 0.f ? -c.x : c.x, dy < 0.f ? -c.y : c.y, dz < 0.f ? -c.z : c.z);
}

template <bool PERIODIC>
__global__ __launch_bounds__(DS_THREADS) void k_direct_sum(const float4* __restrict__ src, int64_t n,
                                                           const float4* __restrict__ tgt, int64_t nt,
                                                           float eps2, Box box, const float4* __restrict__ tab,
                                                           float4* __restrict__ out) {
  __shared__ float4 sh[DS_THREADS];
  const int64_t i = (int64_t)blockIdx.x * DS_THREADS + threadIdx.x;
  const float4 p = i < nt ? tgt[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
  for (int64_t base = 0; base < n; base += DS_THREADS) {
    const int64_t j = base + threadIdx.x;
    sh[threadIdx.x] = j < n ? src[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    const int cnt = (int)std::min<int64_t>(DS_THREADS, n - base);
    float bx = 0.f, by = 0.f, bz = 0.f;  // block partials, then compensated into the total
    for (int k = 0; k < cnt; ++k) {
      const float4 q = sh[k];
      float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      if constexpr (PERIODIC) {
        dx = nearest(dx, box.L[0], box.invL[0]);
        dy = nearest(dy, box.L[1], box.invL[1]);
        dz = nearest(dz, box.L[2], box.invL[2]);
      }
      const float r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > 0.f) {
        const float s = r2 + eps2;
        const float inv = 1.0f / (s * sqrtf(s));
        const float f = q.w * inv;
        bx = fmaf(f, dx, bx);
        by = fmaf(f, dy, by);
        bz = fmaf(f, dz, bz);
        if constexpr (PERIODIC) {
          // the correction is an acceleration (G = 1 units); the Newtonian part
          // above enters with a minus sign below
          const float3 c = ewald_corr(tab, box, dx, dy, dz);
          bx = fmaf(-q.w, c.x, bx);
          by = fmaf(-q.w, c.y, by);
          bz = fmaf(-q.w, c.z, bz);
        }
      }
    }
    neumaier(sx, cx, bx);
    neumaier(sy, cy, by);
    neumaier(sz, cz, bz);
    __syncthreads();
  }
  if (i < nt) out[i] = make_float4(-kGc * (sx + cx), -kGc * (sy + cy), -kGc * (sz + cz), 0.f);
}

// Periodic acceleration of a unit mass at the origin on a neutralising
// background, G = 1 (Ewald summation):
//   a(x) = -sum_n (x-nL)/r^3 [erfc(alpha r) + 2 alpha r/sqrt(pi) exp(-alpha^2 r^2)]
//          - (4 pi / V) sum_{k != 0} k/k^2 exp(-k^2 / (4 alpha^2)) sin(k.x)
// with k = 2 pi (hx/Lx, hy/Ly, hz/Lz); alpha = 2 / min(L); the sums are cut
// where the terms fall below 1e-12 of the leading one.
void ewald_accel(const double L[3], const double x[3], double out[3]) {
  const double Lmin = std::min(L[0], std::min(L[1], L[2]));
  const double alpha = 2.0 / Lmin, V = L[0] * L[1] * L[2];
  double a[3] = {0.0, 0.0, 0.0};
  int R[3], Hm[3];
  for (int c = 0; c < 3; ++c) {
    R[c] = (int)std::ceil(5.0 / (alpha * L[c]) + 0.5);
    Hm[c] = (int)std::ceil(std::sqrt(28.0) * 2.0 * alpha * L[c] / (2.0 * M_PI));
  }
  for (int i = -R[0]; i <= R[0]; ++i)
    for (int j = -R[1]; j <= R[1]; ++j)
      for (int k = -R[2]; k <= R[2]; ++k) {
        const double d[3] = {x[0] - i * L[0], x[1] - j * L[1], x[2] - k * L[2]};
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (r2 == 0.0) continue;
        const double r = std::sqrt(r2);
        const double f = (std::erfc(alpha * r) + 2.0 * alpha * r / std::sqrt(M_PI) * std::exp(-alpha * alpha * r2)) /
                         (r2 * r);
        for (int c = 0; c < 3; ++c) a[c] -= f * d[c];
      }
  for (int i = -Hm[0]; i <= Hm[0]; ++i)
    for (int j = -Hm[1]; j <= Hm[1]; ++j)
      for (int k = -Hm[2]; k <= Hm[2]; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        const double kv[3] = {2.0 * M_PI * i / L[0], 2.0 * M_PI * j / L[1], 2.0 * M_PI * k / L[2]};
        const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        const double e = k2 / (4.0 * alpha * alpha);
        if (e > 28.0) continue;
        const double f = 4.0 * M_PI / V * std::exp(-e) / k2 * std::sin(kv[0] * x[0] + kv[1] * x[1] + kv[2] * x[2]);
        for (int c = 0; c < 3; ++c) a[c] -= f * kv[c];
      }
  for (int c = 0; c < 3; ++c) out[c] = a[c];
}

// Table of the correction a_ewald(x) - (-x/|x|^3) on the octant grid; 0 at x = 0.
constexpr int64_t EW_ENTRIES = (int64_t)(EW_N + 1) * (EW_N + 1) * (EW_N + 1);

void ewald_table_fill(const float Lf[3], float4* tab) {
  const double L[3] = {Lf[0], Lf[1], Lf[2]};
  constexpr int S = EW_N + 1;
  auto work = [&](int z0, int z1) {
    for (int iz = z0; iz < z1; ++iz)
      for (int iy = 0; iy < S; ++iy)
        for (int ix = 0; ix < S; ++ix) {
          const double x[3] = {0.5 * L[0] * ix / EW_N, 0.5 * L[1] * iy / EW_N, 0.5 * L[2] * iz / EW_N};
          const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
          float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
          if (r2 > 0.0) {
            double a[3];
            ewald_accel(L, x, a);
            const double r3 = r2 * std::sqrt(r2);
            v = make_float4((float)(a[0] + x[0] / r3), (float)(a[1] + x[1] / r3), (float)(a[2] + x[2] / r3), 0.f);
          }
          tab[ix + (size_t)S * (iy + (size_t)S * iz)] = v;
        }
  };
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int t = 0; t < nth; ++t) th.emplace_back(work, S * t / nth, S * (t + 1) / nth);
  for (auto& t : th) t.join();
}

std::vector<float4> ewald_table_compute(const float Lf[3]) {
  std::vector<float4> tab((size_t)EW_ENTRIES);
  ewald_table_fill(Lf, tab.data());
  return tab;
}

// Tables are cached per box (a training set reuses one box for every snapshot).
const std::vector<float4>& ewald_table(const float L[3]) {
  static std::mutex mu;
  static std::map<std::array<float, 3>, std::vector<float4>> cache;
  std::lock_guard<std::mutex> lock(mu);
  const std::array<float, 3> key = {L[0], L[1], L[2]};
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, ewald_table_compute(L)).first;
  return it->second;
}

// Staging and table memory of pmgpu_direct_sum_noalloc, in one file-scope
// device array (each device holds its own copy until the process ends):
//   [0, 16 MiB)        sources as float4 (x, y, z, m), DS_MAX_SOURCES
//   [16, 28 MiB)       source positions as given, 3 floats each
//   [28, 32 MiB)       source masses
//   then               targets as float4, results as float4 (DS_MAX_TARGETS
//                      each), target positions and results as 3 floats each,
//                      and the Ewald table (EW_ENTRIES float4).
constexpr int64_t DS_MAX_SOURCES = PMGPU_DS_NOALLOC_MAX_SOURCES;
constexpr int64_t DS_MAX_TARGETS = PMGPU_DS_NOALLOC_MAX_TARGETS;
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t SCR_SRC4 = 0;
constexpr size_t SCR_POS = SCR_SRC4 + align256(sizeof(float4) * DS_MAX_SOURCES);
constexpr size_t SCR_MASS = SCR_POS + align256(sizeof(float) * 3 * DS_MAX_SOURCES);
constexpr size_t SCR_TGT4 = SCR_MASS + align256(sizeof(float) * DS_MAX_SOURCES);
constexpr size_t SCR_RES4 = SCR_TGT4 + align256(sizeof(float4) * DS_MAX_TARGETS);
constexpr size_t SCR_TPOS = SCR_RES4 + align256(sizeof(float4) * DS_MAX_TARGETS);
constexpr size_t SCR_OUT = SCR_TPOS + align256(sizeof(float) * 3 * DS_MAX_TARGETS);
constexpr size_t SCR_TAB = SCR_OUT + align256(sizeof(float) * 3 * DS_MAX_TARGETS);
constexpr size_t SCR_BYTES = SCR_TAB + align256(sizeof(float4) * EW_ENTRIES);

}  // namespace

}  // namespace pmgpu

static __device__ __align__(256) unsigned char scratch[pmgpu::SCR_BYTES];

namespace pmgpu {
namespace {

__global__ void k_pack_sources(const float* __restrict__ pos, const float* __restrict__ mass, int64_t n,
                               float4* __restrict__ src) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) src[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass ? mass[i] : 0.0f);
}

__global__ void k_unpack_results(const float4* __restrict__ res, int64_t nt, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const float4 r = res[i];
  out[3 * i] = r.x;
  out[3 * i + 1] = r.y;
  out[3 * i + 2] = r.z;
}

// Device address of `scratch` on each device (hipGetSymbolAddress with the
// device current, once per device), and the box whose Ewald table the
// device's copy holds.
constexpr int SCR_MAX_DEVICES = 64;
std::mutex g_scr_mu;
unsigned char* g_scr_ptr[SCR_MAX_DEVICES] = {};
float g_scr_box[SCR_MAX_DEVICES][3] = {};
// host copy of the last table computed for the no-alloc path, and its box
float4 g_tab_host[EW_ENTRIES];
float g_tab_host_box[3] = {0.f, 0.f, 0.f};

unsigned char* scratch_ptr(int dev) {
  if (dev < 0 || dev >= SCR_MAX_DEVICES) fail(PMGPU_ERR_UNSUPPORTED, "device index beyond the scratch table");
  if (!g_scr_ptr[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scr_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_scr_ptr[dev];
}

}  // namespace

}  // namespace pmgpu

using namespace pmgpu;

struct pmgpu_trainer {
  pmgpu_ctx* ctx = nullptr;
  int scheme = 0, centred = 0, fds = 0, S = 1;
  std::vector<Snap> snaps;
  float* geff = nullptr;   // K
  float* gradc = nullptr;  // K
  float* R = nullptr;      // 3M, zero outside a snapshot's adjoint pass
  float* lam = nullptr;    // M
  double* loss_parts = nullptr;
  int loss_cap = 0;
  MlpDevice* mlp = nullptr;
  float* params = nullptr;  // NPARAM (device)
  float* grad = nullptr;    // NPARAM (device)
  std::string err;
};

namespace pmgpu {
namespace {

constexpr int NPARAM = 8705;

// Host copy of the deposit's per-axis weights (csrc/deposit.hip), floor
// anchor for CIC/TSC, round for NGP and the centred TSC.
void axis_stencil(int scheme, int centred, float x, int* first, float w[3], int* cnt) {
  if (scheme == PMGPU_NGP) {
    *first = (int)std::round(x);
    w[0] = 1.0f;
    *cnt = 1;
    return;
  }
  const float f = centred ? std::round(x) : std::floor(x);
  const float d = x - f;
  if (scheme == PMGPU_CIC) {
    *first = (int)f;
    w[0] = 1.0f - d;
    w[1] = d;
    *cnt = 2;
  } else {
    *first = (int)f - 1;
    w[0] = (0.5f - d) * (0.5f - d);
    w[1] = 1.5f - 2.0f * d * d;
    w[2] = (0.5f + d) * (0.5f + d);
    *cnt = 3;
  }
}

void ensure_loss(pmgpu_trainer* t) {
  if ((int)t->snaps.size() > t->loss_cap) {
    HIP_CHECK(hipStreamSynchronize(t->ctx->stream));
    dfree(t->loss_parts);
    t->loss_cap = (int)t->snaps.size() * 2 + 8;
    dmalloc(&t->loss_parts, t->loss_cap);
  }
}

unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Loss and dL/dc over every snapshot of t for the parameters in t->params
// (already packed into t->mlp).  w is the loss normaliser 1/sum|t|^2.
void loss_and_gradc(pmgpu_trainer* t, double w) {
  pmgpu_ctx* c = t->ctx;
  hipStream_t s = c->stream;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  mlp_forward_grid(t->mlp, nx, ny, nz, c->sp, t->scheme, c->greens, t->geff, s);
  hipLaunchKernelGGL(k_fill, dim3(blocks(c->K)), dim3(256), 0, s, t->gradc, c->K, 0.0f);
  ensure_loss(t);
  for (size_t k = 0; k < t->snaps.size(); ++k) {
    Snap& sn = t->snaps[k];
    fft_mul_c2r_inverse(c, sn.rho_hat, t->geff, c->spec, c->phi);
    if (sn.nt > 0) {
      if (t->fds == PMGPU_TWO_POINT)
        hipLaunchKernelGGL(k_target_forward<PMGPU_TWO_POINT>, dim3(blocks(sn.nt)), dim3(256), 0, s, c->phi, nx,
                           ny, nz, sn.tcell, sn.tw, t->S, sn.tacc, sn.nt, (float)(2.0 * w), sn.resid, sn.sq);
      else
        hipLaunchKernelGGL(k_target_forward<PMGPU_FOUR_POINT>, dim3(blocks(sn.nt)), dim3(256), 0, s, c->phi,
                           nx, ny, nz, sn.tcell, sn.tw, t->S, sn.tacc, sn.nt, (float)(2.0 * w), sn.resid, sn.sq);
      HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_adj_gather, dim3(blocks(sn.ncell)), dim3(256), 0, s, sn.cell_id, sn.cell_start,
                         sn.ent_target, sn.ent_w, sn.ncell, sn.resid, c->M, t->R);
    }
    if (t->fds == PMGPU_TWO_POINT)
      hipLaunchKernelGGL(k_adj_div<PMGPU_TWO_POINT>, dim3(blocks(c->M)), dim3(256), 0, s, t->R, nx, ny, nz,
                         t->lam);
    else
      hipLaunchKernelGGL(k_adj_div<PMGPU_FOUR_POINT>, dim3(blocks(c->M)), dim3(256), 0, s, t->R, nx, ny, nz,
                         t->lam);
    HIP_CHECK(hipGetLastError());
    if (sn.nt > 0)
      hipLaunchKernelGGL(k_adj_clear, dim3(blocks(sn.ncell)), dim3(256), 0, s, sn.cell_id, sn.ncell, c->M, t->R);
    fft_r2c_forward(c, t->lam, c->spec);
    hipLaunchKernelGGL(k_grad_acc, dim3(blocks(c->K)), dim3(256), 0, s, sn.rho_hat, c->spec, c->greens, c->K,
                       c->sp, nx, t->gradc);
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, s, sn.sq, sn.nt, w, t->loss_parts, (int)k);
    HIP_CHECK(hipGetLastError());
  }
}

double read_loss(pmgpu_trainer* t) {
  std::vector<double> parts(t->snaps.size());
  if (!parts.empty())
    HIP_CHECK(hipMemcpyAsync(parts.data(), t->loss_parts, sizeof(double) * parts.size(), hipMemcpyDeviceToHost,
                             t->ctx->stream));
  HIP_CHECK(hipStreamSynchronize(t->ctx->stream));
  double L = 0.0;
  for (double v : parts) L += v;
  return L;
}

void upload_params(pmgpu_trainer* t, const float* params) {
  if (params) {
    HIP_CHECK(hipMemcpyAsync(t->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, t->ctx->stream));
  } else {
    hipLaunchKernelGGL(k_fill, dim3(blocks(NPARAM)), dim3(256), 0, t->ctx->stream, t->params, (int64_t)NPARAM,
                       0.0f);
  }
  mlp_pack(t->mlp, t->params, t->ctx->stream);
}

double total_norm(pmgpu_trainer* const* trs, int ntr) {
  double s = 0.0;
  for (int i = 0; i < ntr; ++i)
    for (const Snap& sn : trs[i]->snaps) s += sn.tnorm;
  return s;
}

}  // namespace
}  // namespace pmgpu

#define TRAINER_BEGIN(t) \
  if (!(t)) return PMGPU_ERR_INVALID_ARGUMENT; \
  try {
#define TRAINER_END(t)             \
  }                                \
  catch (const Error& e) {         \
    (t)->err = e.msg;              \
    return e.code;                 \
  }                                \
  catch (const std::exception& e) { \
    (t)->err = e.what();           \
    return PMGPU_ERR_DEVICE;       \
  }                                \
  return PMGPU_OK;

extern "C" {

int pmgpu_trainer_create(const pmgpu_pm_params* params, pmgpu_trainer** out) {
  if (!out || !params) return PMGPU_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  pmgpu_ctx* c = nullptr;
  int rc = pmgpu_create(params, &c);
  if (rc != PMGPU_OK) return rc;
  if (c->general) {
    pmgpu_destroy(c);
    return PMGPU_ERR_UNSUPPORTED;  // the trainer's transforms use the power-of-two FFT passes
  }
  pmgpu_trainer* t = new pmgpu_trainer();
  t->ctx = c;
  try {
    t->scheme = c->prm.interpolation;
    t->centred = c->tsc_centred;
    t->fds = params->finite_diff;
    t->S = t->scheme == PMGPU_NGP ? 1 : (t->scheme == PMGPU_CIC ? 8 : 27);
    greens_compute(c, nullptr);  // analytic G_sym / M
    dmalloc(&t->geff, c->K);
    dmalloc(&t->gradc, c->K);
    dmalloc(&t->R, 3 * c->M);
    dmalloc(&t->lam, c->M);
    dmalloc(&t->params, NPARAM);
    dmalloc(&t->grad, NPARAM);
    hipLaunchKernelGGL(k_fill, dim3(blocks(3 * c->M)), dim3(256), 0, c->stream, t->R, 3 * c->M, 0.0f);
    HIP_CHECK(hipGetLastError());
    t->mlp = mlp_device_create();
    HIP_CHECK(hipStreamSynchronize(c->stream));
  } catch (const Error& e) {
    pmgpu_trainer_destroy(t);
    return e.code;
  }
  *out = t;
  return PMGPU_OK;
}

void pmgpu_trainer_destroy(pmgpu_trainer* t) {
  if (!t) return;
  if (t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
  for (Snap& s : t->snaps) {
    dfree(s.rho_hat);
    dfree(s.tcell);
    dfree(s.tw);
    dfree(s.tacc);
    dfree(s.resid);
    dfree(s.sq);
    dfree(s.cell_id);
    dfree(s.cell_start);
    dfree(s.ent_target);
    dfree(s.ent_w);
  }
  dfree(t->geff);
  dfree(t->gradc);
  dfree(t->R);
  dfree(t->lam);
  dfree(t->loss_parts);
  dfree(t->params);
  dfree(t->grad);
  mlp_device_destroy(t->mlp);
  if (t->ctx) pmgpu_destroy(t->ctx);
  delete t;
}

const char* pmgpu_trainer_last_error(const pmgpu_trainer* t) { return t ? t->err.c_str() : "null trainer"; }

int pmgpu_trainer_add_snapshot(pmgpu_trainer* t, int64_t n, const float* pos, const float* mass, int64_t nt,
                               const int64_t* target_index, const float* target_acc) {
  TRAINER_BEGIN(t)
  if (n <= 0 || !pos || !mass) fail(PMGPU_ERR_INVALID_ARGUMENT, "empty snapshot");
  if (nt < 0 || (nt > 0 && (!target_index || !target_acc))) fail(PMGPU_ERR_INVALID_ARGUMENT, "bad targets");
  pmgpu_ctx* c = t->ctx;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int64_t i = 0; i < nt; ++i)
    if (target_index[i] < 0 || target_index[i] >= n) fail(PMGPU_ERR_INVALID_ARGUMENT, "target index out of range");
  // density spectrum through the force step's own deposit and FFT
  int rc = pmgpu_set_particles(c, n, pos, nullptr, mass);
  if (rc != PMGPU_OK) fail(rc, c->err);
  tile_sort(c);
  deposit(c);
  Snap sn;
  dmalloc(&sn.rho_hat, c->K);
  fft_r2c_forward(c, c->rho, sn.rho_hat);
  // target stencils
  const int S = t->S;
  std::vector<int32_t> tcell((size_t)nt * S);
  std::vector<float> tw((size_t)nt * S);
  std::vector<float4> tacc(nt);
  for (int64_t i = 0; i < nt; ++i) {
    const int64_t j = target_index[i];
    int f[3], cnt[3];
    float w[3][3];
    for (int a = 0; a < 3; ++a) axis_stencil(t->scheme, t->centred, pos[3 * j + a], &f[a], w[a], &cnt[a]);
    int e = 0;
    for (int a = 0; a < cnt[0]; ++a)
      for (int b = 0; b < cnt[1]; ++b)
        for (int d = 0; d < cnt[2]; ++d) {
          float wt = w[0][a] * w[1][b];
          wt = wt * w[2][d];
          if (t->scheme == PMGPU_TSC) wt *= 0.125f;
          const int gx = ((f[0] + a) % nx + nx) % nx, gy = ((f[1] + b) % ny + ny) % ny,
                    gz = ((f[2] + d) % nz + nz) % nz;
          tcell[(size_t)i * S + e] = gx + nx * (gy + ny * gz);
          tw[(size_t)i * S + e] = wt;
          ++e;
        }
    const float* ta = target_acc + 3 * i;
    tacc[i] = make_float4(ta[0], ta[1], ta[2], 0.f);
    sn.tnorm += (double)ta[0] * ta[0] + (double)ta[1] * ta[1] + (double)ta[2] * ta[2];
  }
  // transpose: entries sorted by cell, then target
  std::vector<int64_t> order((size_t)nt * S);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return tcell[a] < tcell[b]; });
  std::vector<int32_t> cell_id, cell_start, ent_target(order.size());
  std::vector<float> ent_w(order.size());
  for (size_t k = 0; k < order.size(); ++k) {
    const int64_t e = order[k];
    if (k == 0 || tcell[e] != tcell[order[k - 1]]) {
      cell_id.push_back(tcell[e]);
      cell_start.push_back((int32_t)k);
    }
    ent_target[k] = (int32_t)(e / S);
    ent_w[k] = tw[e];
  }
  cell_start.push_back((int32_t)order.size());
  sn.nt = nt;
  sn.ncell = (int64_t)cell_id.size();
  dmalloc(&sn.tcell, tcell.size());
  dmalloc(&sn.tw, tw.size());
  dmalloc(&sn.tacc, nt);
  dmalloc(&sn.resid, nt);
  dmalloc(&sn.sq, nt);
  dmalloc(&sn.cell_id, sn.ncell);
  dmalloc(&sn.cell_start, sn.ncell + 1);
  dmalloc(&sn.ent_target, ent_target.size());
  dmalloc(&sn.ent_w, ent_w.size());
  hipStream_t s = c->stream;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  up(sn.tcell, tcell.data(), tcell.size() * 4);
  up(sn.tw, tw.data(), tw.size() * 4);
  up(sn.tacc, tacc.data(), tacc.size() * sizeof(float4));
  up(sn.cell_id, cell_id.data(), cell_id.size() * 4);
  up(sn.cell_start, cell_start.data(), cell_start.size() * 4);
  up(sn.ent_target, ent_target.data(), ent_target.size() * 4);
  up(sn.ent_w, ent_w.data(), ent_w.size() * 4);
  HIP_CHECK(hipStreamSynchronize(s));  // host vectors die here
  t->snaps.push_back(sn);
  TRAINER_END(t)
}

int pmgpu_trainer_loss_grad(pmgpu_trainer* t, const float* params, double* loss, float* grad, float* grad_c) {
  TRAINER_BEGIN(t)
  if (t->snaps.empty()) fail(PMGPU_ERR_NOT_READY, "no snapshots");
  const double norm = total_norm(&t, 1);
  if (!(norm > 0.0)) fail(PMGPU_ERR_INVALID_ARGUMENT, "target accelerations are all zero");
  upload_params(t, params);
  loss_and_gradc(t, 1.0 / norm);
  pmgpu_ctx* c = t->ctx;
  if (grad) {
    mlp_backward_grid(t->mlp, c->nx, c->ny, c->nz, c->sp, t->scheme, t->gradc, 1.0f, false, t->grad, c->stream);
    HIP_CHECK(hipMemcpyAsync(grad, t->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c->stream));
  }
  if (grad_c) {  // [z][y][kx <= nx/2], without the row padding
    std::vector<float> g((size_t)c->K);
    HIP_CHECK(hipMemcpyAsync(g.data(), t->gradc, sizeof(float) * c->K, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int64_t r = 0; r < (int64_t)c->ny * c->nz; ++r)
      std::memcpy(grad_c + r * c->nxh, g.data() + r * c->sp, sizeof(float) * c->nxh);
  }
  const double L = read_loss(t);
  if (loss) *loss = L;
  TRAINER_END(t)
}

int pmgpu_train(pmgpu_trainer** trs, int32_t ntr, float* params, int32_t steps, float lr, float* loss_history) {
  if (!trs || ntr <= 0 || !params || steps < 0 || !(lr > 0.0f)) return PMGPU_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < ntr; ++i)
    if (!trs[i] || trs[i]->snaps.empty()) return PMGPU_ERR_INVALID_ARGUMENT;
  pmgpu_trainer* t0 = trs[0];
  float *m = nullptr, *v = nullptr;
  TRAINER_BEGIN(t0)
  const double norm = total_norm(trs, ntr);
  if (!(norm > 0.0)) fail(PMGPU_ERR_INVALID_ARGUMENT, "target accelerations are all zero");
  hipStream_t s0 = t0->ctx->stream;
  dmalloc(&m, NPARAM);
  dmalloc(&v, NPARAM);
  hipLaunchKernelGGL(k_fill, dim3(blocks(NPARAM)), dim3(256), 0, s0, m, (int64_t)NPARAM, 0.0f);
  hipLaunchKernelGGL(k_fill, dim3(blocks(NPARAM)), dim3(256), 0, s0, v, (int64_t)NPARAM, 0.0f);
  upload_params(t0, params);
  HIP_CHECK(hipStreamSynchronize(s0));
  const float b1 = 0.9f, b2 = 0.999f;
  for (int step = 0; step < steps; ++step) {
    double L = 0.0;
    for (int i = 0; i < ntr; ++i) {
      pmgpu_trainer* t = trs[i];
      pmgpu_ctx* c = t->ctx;
      if (i > 0) {
        HIP_CHECK(hipMemcpyAsync(t->params, t0->params, sizeof(float) * NPARAM, hipMemcpyDeviceToDevice, c->stream));
        mlp_pack(t->mlp, t->params, c->stream);
      }
      loss_and_gradc(t, 1.0 / norm);
      mlp_backward_grid(t->mlp, c->nx, c->ny, c->nz, c->sp, t->scheme, t->gradc, 1.0f, false, t->grad, c->stream);
      L += read_loss(t);  // synchronises this trainer's stream
      if (i > 0) {
        hipLaunchKernelGGL(k_axpy_add, dim3(blocks(NPARAM)), dim3(256), 0, s0, t0->grad, t->grad, NPARAM);
        HIP_CHECK(hipGetLastError());
      }
    }
    if (loss_history) loss_history[step] = (float)L;
    const float c1 = 1.0f - std::pow(b1, (float)(step + 1)), c2 = 1.0f - std::pow(b2, (float)(step + 1));
    hipLaunchKernelGGL(k_adam, dim3(blocks(NPARAM)), dim3(256), 0, s0, t0->params, t0->grad, m, v, NPARAM, lr, b1,
                       b2, 1e-8f, c1, c2);
    HIP_CHECK(hipGetLastError());
    mlp_pack(t0->mlp, t0->params, s0);
  }
  HIP_CHECK(hipMemcpyAsync(params, t0->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, s0));
  HIP_CHECK(hipStreamSynchronize(s0));
  dfree(m);
  dfree(v);
  return PMGPU_OK;
  }
  catch (const Error& e) {
    dfree(m);
    dfree(v);
    t0->err = e.msg;
    return e.code;
  }
}

int pmgpu_direct_sum(int64_t n, const float* pos, const float* mass, int64_t nt, const float* tpos,
                     float softening, const float* box, float* out) {
  if (n < 0 || nt < 0 || (n > 0 && (!pos || !mass)) || (nt > 0 && (!tpos || !out))) return PMGPU_ERR_INVALID_ARGUMENT;
  const bool periodic = box != nullptr;
  if (periodic && !(box[0] > 0.f && box[1] > 0.f && box[2] > 0.f)) return PMGPU_ERR_INVALID_ARGUMENT;
  if (nt == 0) return PMGPU_OK;
  float4 *src = nullptr, *tgt = nullptr, *res = nullptr, *tab = nullptr;
  int rc = PMGPU_OK;
  try {
    std::vector<float4> hs(std::max<int64_t>(n, 1)), ht(nt), hr(nt);
    for (int64_t i = 0; i < n; ++i) hs[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass[i]);
    for (int64_t i = 0; i < nt; ++i) ht[i] = make_float4(tpos[3 * i], tpos[3 * i + 1], tpos[3 * i + 2], 0.f);
    dmalloc(&src, n);
    dmalloc(&tgt, nt);
    dmalloc(&res, nt);
    if (n > 0) HIP_CHECK(hipMemcpy(src, hs.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(tgt, ht.data(), sizeof(float4) * nt, hipMemcpyHostToDevice));
    Box b{};
    if (periodic) {
      const std::vector<float4>& t = ewald_table(box);
      dmalloc(&tab, (int64_t)t.size());
      HIP_CHECK(hipMemcpy(tab, t.data(), sizeof(float4) * t.size(), hipMemcpyHostToDevice));
      for (int c = 0; c < 3; ++c) {
        b.L[c] = box[c];
        b.invL[c] = 1.0f / box[c];
        b.tab_scale[c] = (float)EW_N / (0.5f * box[c]);
      }
      hipLaunchKernelGGL(k_direct_sum<true>, dim3(blocks(nt, DS_THREADS)), dim3(DS_THREADS), 0, nullptr, src, n,
                         tgt, nt, softening * softening, b, tab, res);
    } else {
      hipLaunchKernelGGL(k_direct_sum<false>, dim3(blocks(nt, DS_THREADS)), dim3(DS_THREADS), 0, nullptr, src, n,
                         tgt, nt, softening * softening, b, (const float4*)nullptr, res);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(hr.data(), res, sizeof(float4) * nt, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nt; ++i) {
      out[3 * i] = hr[i].x;
      out[3 * i + 1] = hr[i].y;
      out[3 * i + 2] = hr[i].z;
    }
  } catch (const Error& e) {
    rc = e.code;
  }
  dfree(src);
  dfree(tgt);
  dfree(res);
  dfree(tab);
  return rc;
}

// pmgpu_direct_sum without device allocation: the staging copies of the
// inputs and outputs and the Ewald table live in the file-scope `scratch`
// array (at most DS_MAX_SOURCES sources and DS_MAX_TARGETS targets); the
// kernels are pmgpu_direct_sum's.  Runs on the null stream of the current
// device and returns when `out` is written.  Calls that overlap on one
// device share its scratch and race.
int pmgpu_direct_sum_noalloc(int64_t n, const float* pos, const float* mass, int64_t nt, const float* tpos,
                             float softening, const float* box, float* out) {
  if (n < 0 || nt < 0 || (n > 0 && (!pos || !mass)) || (nt > 0 && (!tpos || !out))) return PMGPU_ERR_INVALID_ARGUMENT;
  if (n > DS_MAX_SOURCES || nt > DS_MAX_TARGETS) return PMGPU_ERR_UNSUPPORTED;
  const bool periodic = box != nullptr;
  if (periodic && !(box[0] > 0.f && box[1] > 0.f && box[2] > 0.f)) return PMGPU_ERR_INVALID_ARGUMENT;
  if (nt == 0) return PMGPU_OK;
  try {
    std::lock_guard<std::mutex> lock(g_scr_mu);
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    unsigned char* base = scratch_ptr(dev);
    float4* src = reinterpret_cast<float4*>(base + SCR_SRC4);
    float* spos = reinterpret_cast<float*>(base + SCR_POS);
    float* smass = reinterpret_cast<float*>(base + SCR_MASS);
    float4* tgt = reinterpret_cast<float4*>(base + SCR_TGT4);
    float4* res = reinterpret_cast<float4*>(base + SCR_RES4);
    float* stpos = reinterpret_cast<float*>(base + SCR_TPOS);
    float* sout = reinterpret_cast<float*>(base + SCR_OUT);
    float4* tab = reinterpret_cast<float4*>(base + SCR_TAB);
    if (n > 0) {
      HIP_CHECK(hipMemcpy(spos, pos, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice));
      HIP_CHECK(hipMemcpy(smass, mass, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_pack_sources, dim3(blocks(n)), dim3(256), 0, nullptr, spos, smass, n, src);
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipMemcpy(stpos, tpos, sizeof(float) * 3 * (size_t)nt, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pack_sources, dim3(blocks(nt)), dim3(256), 0, nullptr, stpos, (const float*)nullptr, nt,
                       tgt);
    HIP_CHECK(hipGetLastError());
    Box b{};
    if (periodic) {
      if (std::memcmp(g_scr_box[dev], box, sizeof(float) * 3) != 0) {
        if (std::memcmp(g_tab_host_box, box, sizeof(float) * 3) != 0) {
          ewald_table_fill(box, g_tab_host);
          std::memcpy(g_tab_host_box, box, sizeof(float) * 3);
        }
        HIP_CHECK(hipMemcpy(tab, g_tab_host, sizeof(float4) * EW_ENTRIES, hipMemcpyHostToDevice));
        std::memcpy(g_scr_box[dev], box, sizeof(float) * 3);
      }
      for (int c = 0; c < 3; ++c) {
        b.L[c] = box[c];
        b.invL[c] = 1.0f / box[c];
        b.tab_scale[c] = (float)EW_N / (0.5f * box[c]);
      }
      hipLaunchKernelGGL(k_direct_sum<true>, dim3(blocks(nt, DS_THREADS)), dim3(DS_THREADS), 0, nullptr, src, n,
                         tgt, nt, softening * softening, b, tab, res);
    } else {
      hipLaunchKernelGGL(k_direct_sum<false>, dim3(blocks(nt, DS_THREADS)), dim3(DS_THREADS), 0, nullptr, src, n,
                         tgt, nt, softening * softening, b, (const float4*)nullptr, res);
    }
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_unpack_results, dim3(blocks(nt)), dim3(256), 0, nullptr, res, nt, sout);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(out, sout, sizeof(float) * 3 * (size_t)nt, hipMemcpyDeviceToHost));
  } catch (const Error& e) {
    return e.code;
  }
  return PMGPU_OK;
}

// The same sums on the host CPU in FP64 (an error evaluator, the FP64
// counterpart of the reference's direct PP sum, source/ppMethod.cpp:23-36):
// sources as doubles, one target per task over `threads` std::threads; the
// periodic correction is read from the same Ewald table, interpolated in FP64.
int pmgpu_direct_sum_host(int64_t n, const float* pos, const float* mass, int64_t nt, const float* tpos,
                          float softening, const float* box, int32_t threads, double* out) {
  if (n < 0 || nt < 0 || (n > 0 && (!pos || !mass)) || (nt > 0 && (!tpos || !out))) return PMGPU_ERR_INVALID_ARGUMENT;
  const bool periodic = box != nullptr;
  if (periodic && !(box[0] > 0.f && box[1] > 0.f && box[2] > 0.f)) return PMGPU_ERR_INVALID_ARGUMENT;
  if (nt == 0) return PMGPU_OK;
  try {
    std::vector<double> sx(n), sy(n), sz(n), sm(n);
    for (int64_t j = 0; j < n; ++j) {
      sx[j] = pos[3 * j];
      sy[j] = pos[3 * j + 1];
      sz[j] = pos[3 * j + 2];
      sm[j] = mass[j];
    }
    const std::vector<float4>* tab = periodic ? &ewald_table(box) : nullptr;
    double L[3] = {0, 0, 0}, scale[3] = {0, 0, 0};
    if (periodic)
      for (int c = 0; c < 3; ++c) {
        L[c] = box[c];
        scale[c] = EW_N / (0.5 * box[c]);
      }
    const double eps2 = (double)softening * softening, Gc = 1.0 / (4.0 * M_PI);
    constexpr int S1 = EW_N + 1;
    auto corr = [&](double dx, double dy, double dz, double c[3]) {
      const double u[3] = {std::min(std::fabs(dx) * scale[0], EW_N - 1e-9),
                           std::min(std::fabs(dy) * scale[1], EW_N - 1e-9),
                           std::min(std::fabs(dz) * scale[2], EW_N - 1e-9)};
      const int ix = (int)u[0], iy = (int)u[1], iz = (int)u[2];
      const double fx = u[0] - ix, fy = u[1] - iy, fz = u[2] - iz;
      double acc[3] = {0, 0, 0};
      for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
        const double w = (ox ? fx : 1 - fx) * (oy ? fy : 1 - fy) * (oz ? fz : 1 - fz);
        const float4 v = (*tab)[(ix + ox) + (size_t)S1 * ((iy + oy) + (size_t)S1 * (iz + oz))];
        acc[0] += w * v.x;
        acc[1] += w * v.y;
        acc[2] += w * v.z;
      }
      c[0] = dx < 0 ? -acc[0] : acc[0];
      c[1] = dy < 0 ? -acc[1] : acc[1];
      c[2] = dz < 0 ? -acc[2] : acc[2];
    };
    std::atomic<int64_t> next{0};
    auto work = [&]() {
      for (int64_t t = next++; t < nt; t = next++) {
        const double px = tpos[3 * t], py = tpos[3 * t + 1], pz = tpos[3 * t + 2];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int64_t j = 0; j < n; ++j) {
          double dx = px - sx[j], dy = py - sy[j], dz = pz - sz[j];
          if (periodic) {
            dx -= L[0] * std::nearbyint(dx / L[0]);
            dy -= L[1] * std::nearbyint(dy / L[1]);
            dz -= L[2] * std::nearbyint(dz / L[2]);
          }
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 <= 0.0) continue;
          const double s = r2 + eps2;
          const double f = sm[j] / (s * std::sqrt(s));
          ax += f * dx;
          ay += f * dy;
          az += f * dz;
          if (periodic) {
            double c[3];
            corr(dx, dy, dz, c);
            ax -= sm[j] * c[0];
            ay -= sm[j] * c[1];
            az -= sm[j] * c[2];
          }
        }
        out[3 * t] = -Gc * ax;
        out[3 * t + 1] = -Gc * ay;
        out[3 * t + 2] = -Gc * az;
      }
    };
    const int nth = std::max(1, (int)threads);
    std::vector<std::thread> th;
    for (int i = 0; i < nth; ++i) th.emplace_back(work);
    for (auto& x : th) x.join();
  } catch (const Error& e) {
    return e.code;
  } catch (const std::exception&) {
    return PMGPU_ERR_OUT_OF_MEMORY;
  }
  return PMGPU_OK;
}

}  // extern "C"
