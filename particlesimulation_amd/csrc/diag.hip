// Energy / momentum diagnostics (the reductions of SimInfo,
// source/simInfo.cpp:50-118) as FP64 reductions on the device, so a run with
// diagnostics does not copy the particles to the host every step (the
// reference's CUDA path does, source/PMMethodGPU.cu:121-139).
//
// Two passes with a fixed combination order: each block reduces a contiguous
// slice in FP64 and writes its partial sums, then one block adds the partials
// in block order.  The result is bitwise reproducible.
#include "pmgpu_internal.h"

namespace pmgpu {

namespace {

constexpr int DIAG_THREADS = 256;
constexpr int DIAG_BLOCKS = 512;
constexpr int NQ = PMGPU_DIAG_COUNT;

__device__ double block_sum(double v, double* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < DIAG_THREADS / 64; ++w) t += sh[w];
  __syncthreads();
  return t;
}

// quantities: 0 kinetic 1/2 m v^2, 1-3 momentum m vi, 4-6 angular momentum
// m x cross vi, 7 grid sum rho*phi, 8 total mass, 9-11 mass moment m x
__global__ __launch_bounds__(DIAG_THREADS) void k_diag_partial(
    const float4* __restrict__ pm, const float4* __restrict__ vel, const float4* __restrict__ vint,
    int64_t n, const float* __restrict__ rho, const float* __restrict__ phi, int64_t M,
    double* __restrict__ partial) {
  __shared__ double sh[DIAG_THREADS / 64];
  double q[NQ];
  for (int k = 0; k < NQ; ++k) q[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * DIAG_THREADS;
  for (int64_t i = (int64_t)blockIdx.x * DIAG_THREADS + threadIdx.x; i < n; i += stride) {
    const float4 p = pm[i], v = vel[i];
    const float4 u = vint ? vint[i] : v;
    const double m = p.w;
    q[0] += 0.5 * m * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
    q[1] += m * u.x;
    q[2] += m * u.y;
    q[3] += m * u.z;
    q[4] += m * ((double)p.y * u.z - (double)p.z * u.y);
    q[5] += m * ((double)p.z * u.x - (double)p.x * u.z);
    q[6] += m * ((double)p.x * u.y - (double)p.y * u.x);
    q[8] += m;
    q[9] += m * p.x;
    q[10] += m * p.y;
    q[11] += m * p.z;
  }
  for (int64_t c = (int64_t)blockIdx.x * DIAG_THREADS + threadIdx.x; c < M; c += stride)
    q[7] += (double)rho[c] * (double)phi[c];
  for (int k = 0; k < NQ; ++k) {
    const double t = block_sum(q[k], sh);
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * NQ + k] = t;
  }
}

__global__ __launch_bounds__(DIAG_THREADS) void k_diag_final(const double* __restrict__ partial,
                                                             int nblocks, double* __restrict__ out) {
  if (threadIdx.x < NQ) {
    double t = 0.0;
    for (int b = 0; b < nblocks; ++b) t += partial[(int64_t)b * NQ + threadIdx.x];
    out[threadIdx.x] = t;
  }
}

}  // namespace

}  // namespace pmgpu

using namespace pmgpu;

extern "C" int pmgpu_diagnostics(pmgpu_ctx* c, double* out) {
  if (!c) return PMGPU_ERR_INVALID_ARGUMENT;
  try {
    if (!out) fail(PMGPU_ERR_INVALID_ARGUMENT, "out is NULL");
    if (!c->diag_scratch) {
      HIP_CHECK(hipMalloc((void**)&c->diag_scratch, sizeof(double) * (DIAG_BLOCKS * NQ + NQ)));
    }
    hipLaunchKernelGGL(k_diag_partial, dim3(DIAG_BLOCKS), dim3(DIAG_THREADS), 0, c->stream,
                       c->pm[c->cur], c->vel[c->cur], (const float4*)c->vint, c->n, c->rho, c->phi,
                       c->M, c->diag_scratch);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(DIAG_THREADS), 0, c->stream, c->diag_scratch,
                       DIAG_BLOCKS, c->diag_scratch + DIAG_BLOCKS * NQ);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, c->diag_scratch + DIAG_BLOCKS * NQ, sizeof(double) * NQ,
                             hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  } catch (const Error& e) {
    c->err = e.msg;
    return e.code;
  }
  return PMGPU_OK;
}
