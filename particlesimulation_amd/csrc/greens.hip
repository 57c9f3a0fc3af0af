// Influence function G(k) in FP64 (PMMethod::initGreensFunction,
// source/pmMethod.cpp:164-185, with the four kinds of
// source/greensFunctions.cpp).
//
// The reference evaluates G on the full unfolded grid k_i in [0, N) and
// multiplies a complex-to-complex spectrum; only Re(phi) is used
// (source/grid.cpp:66-68).  Because G is real, Re(IFFT(rho_hat G)) equals
// IFFT(rho_hat G_sym) with G_sym(k) = (G(k) + G(-k mod N)) / 2, which is
// inversion-symmetric, so the half spectrum of a real-to-complex transform
// carries it exactly.  This kernel stores G_sym / (Nx Ny Nz) (the inverse
// FFT's normalisation) on the half spectrum kx in [0, Nx/2].
//
// The optional learned correction multiplies G_sym by (1 + MLP(k)); the MLP
// forward pass is in mlp.hip and writes its factor into `corr`.
#include "pmgpu_internal.h"

namespace pmgpu {

namespace {

constexpr double kPi = 3.14159265358979323846;

__host__ __device__ __forceinline__ double sinc_d(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

// S1Fourier / S2Fourier (greensFunctions.cpp:42-50), with the Taylor series
// for small u where the closed forms cancel catastrophically.
__host__ __device__ double s_fourier(int cs, double k, double a) {
  const double u = k * a / 2.0;
  const double u2 = u * u;
  if (fabs(u) < 0.5) {
    double term = 1.0, sum = 0.0;
    if (cs == PMGPU_S1) {
      double fact = 6.0;
      for (int m = 1; m <= 8; ++m) {
        sum += ((m & 1) ? 1.0 : -1.0) * 2.0 * m * term / fact;
        term *= u2;
        fact *= (2.0 * m + 2.0) * (2.0 * m + 3.0);
      }
      return 3.0 * sum;
    }
    double fact = 24.0;
    for (int m = 2; m <= 9; ++m) {
      sum += ((m & 1) ? -1.0 : 1.0) * (2.0 * m - 2.0) * term / fact;
      term *= u2;
      fact *= (2.0 * m + 1.0) * (2.0 * m + 2.0);
    }
    return 12.0 * sum;
  }
  if (cs == PMGPU_S1) return -3.0 / (u2 * u) * (u * cos(u) - sin(u));
  return 12.0 / (u2 * u2) * (2.0 - 2.0 * cos(u) - u * sin(u));
}

__host__ __device__ double green_optimal(int is, int fds, int cs, double a, int kx, int ky, int kz, int nx,
                                int ny, int nz) {
  if (kx == 0 && ky == 0 && kz == 0) return 0.0;
  const int ki[3] = {kx, ky, kz}, ni[3] = {nx, ny, nz};
  double k[3];
  bool d_zero = true;
  for (int i = 0; i < 3; ++i) {
    k[i] = 2.0 * kPi * (double)ki[i] / (double)ni[i];
    if ((2 * ki[i]) % ni[i] != 0) d_zero = false;
  }
  // every k_i in {0, pi}: D(k) = 0, the mode exerts no force
  if (d_zero) return 0.0;
  // closed-form alias sums (greensFunctions.cpp:102-120)
  double denom = 1.0;
  for (int i = 0; i < 3; ++i) {
    if (is == PMGPU_TSC) {
      double s = sin(k[i] / 2.0), s2 = s * s;
      denom *= 1.0 - s2 + 2.0 / 15.0 * s2 * s2;
    } else if (is == PMGPU_CIC) {
      double c = cos(k[i] / 2.0);
      denom *= (1.0 + 2.0 * c * c) / 3.0;
    }
  }
  // D(k)/i (greensFunctions.cpp:70-89)
  double d[3], dn2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    d[i] = fds == PMGPU_TWO_POINT ? sin(k[i])
                                  : (4.0 / 3.0) * sin(k[i]) + (1.0 - 4.0 / 3.0) * sin(2.0 * k[i]) / 2.0;
    dn2 += d[i] * d[i];
  }
  const int p = is == PMGPU_TSC ? 3 : (is == PMGPU_CIC ? 2 : 1);
  double usq[3][5], kn[3][5];
  for (int i = 0; i < 3; ++i)
    for (int n = 0; n < 5; ++n) {
      double q = k[i] + 2.0 * kPi * (n - 2);
      kn[i][n] = q;
      double s = sinc_d(q / 2.0), s2 = s * s;
      usq[i][n] = p == 3 ? s2 * s2 * s2 : (p == 2 ? s2 * s2 : s2);
    }
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  for (int n1 = 0; n1 < 5; ++n1)
    for (int n2 = 0; n2 < 5; ++n2)
      for (int n3 = 0; n3 < 5; ++n3) {
        const double qx = kn[0][n1], qy = kn[1][n2], qz = kn[2][n3];
        const double q2 = qx * qx + qy * qy + qz * qz;
        const double s = s_fourier(cs, sqrt(q2), a);
        const double f = usq[0][n1] * usq[1][n2] * usq[2][n3] * s * s / q2;
        r0 += f * qx;
        r1 += f * qy;
        r2 += f * qz;
      }
  // numerator D . conj(R): D_i = i d_i, R_i = -i r_i
  const double num = -(d[0] * r0 + d[1] * r1 + d[2] * r2);
  return num / (dn2 * denom * denom);
}

__host__ __device__ double green_at(int gf, int is, int fds, double a, int kx, int ky, int kz, int nx,
                           int ny, int nz) {
  if (kx == 0 && ky == 0 && kz == 0) return 0.0;
  if (gf == PMGPU_DISCRETE_LAPLACIAN) {
    // greensFunctions.cpp:191-200
    double sx = sin(kPi * kx / nx), sy = sin(kPi * ky / ny), sz = sin(kPi * kz / nz);
    return -0.25 / (sx * sx + sy * sy + sz * sz);
  }
  if (gf == PMGPU_POOR_MAN) {
    // greensFunctions.cpp:202-220 (folded k)
    int fx = kx <= nx / 2 ? kx : kx - nx, fy = ky <= ny / 2 ? ky : ky - ny,
        fz = kz <= nz / 2 ? kz : kz - nz;
    double qx = 2.0 * kPi * fx / nx, qy = 2.0 * kPi * fy / ny, qz = 2.0 * kPi * fz / nz;
    return -1.0 / (qx * qx + qy * qy + qz * qz);
  }
  return green_optimal(is, fds, gf == PMGPU_S1_OPTIMAL ? PMGPU_S1 : PMGPU_S2, a, kx, ky, kz, nx, ny,
                       nz);
}

__global__ void k_greens(int gf, int is, int fds, double a, int nx, int ny, int nz, int sp,
                         const float* __restrict__ corr, float* __restrict__ out) {
  const int nxh = nx / 2 + 1;
  const int64_t K = (int64_t)sp * ny * nz;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  const int kx = (int)(i % sp), ky = (int)((i / sp) % ny), kz = (int)(i / ((int64_t)sp * ny));
  if (kx >= nxh) {  // row padding
    out[i] = 0.0f;
    return;
  }
  const double g1 = green_at(gf, is, fds, a, kx, ky, kz, nx, ny, nz);
  const double g2 = green_at(gf, is, fds, a, (nx - kx) % nx, (ny - ky) % ny, (nz - kz) % nz, nx,
                             ny, nz);
  double g = 0.5 * (g1 + g2) / ((double)nx * (double)ny * (double)nz);
  if (corr) g *= 1.0 + (double)corr[i];
  out[i] = (float)g;
}

}  // namespace

void greens_compute(pmgpu_ctx* c, const float* corr_dev) {
  c->greens_z_valid = false;  // the z pass's copy is rebuilt by pmgpu_init_greens
  ++c->graph_gen;            // captured run-loop graphs may read that copy
  int tok = stage_begin(c, "greens_fp64");
  const unsigned g = (unsigned)((c->K + 127) / 128);
  hipLaunchKernelGGL(k_greens, dim3(g), dim3(128), 0, c->stream, c->prm.greens_function,
                     c->prm.interpolation, c->prm.finite_diff, (double)c->prm.particle_diameter,
                     c->nx, c->ny, c->nz, c->sp, corr_dev, c->greens);
  HIP_CHECK(hipGetLastError());
  stage_end(c, tok);
}

}  // namespace pmgpu

extern "C" double pmgpu_greens_value(int32_t greens_function, int32_t interpolation, int32_t finite_diff,
                                     double particle_diameter, int32_t kx, int32_t ky, int32_t kz,
                                     int32_t nx, int32_t ny, int32_t nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || greens_function < 0 || greens_function > 3 || interpolation < 0 ||
      interpolation > PMGPU_TSC_CENTRED || finite_diff < 0 || finite_diff > 1)
    return NAN;
  if (interpolation == PMGPU_TSC_CENTRED) interpolation = PMGPU_TSC;  // same assignment function
  return pmgpu::green_at(greens_function, interpolation, finite_diff, particle_diameter, kx, ky, kz,
                         nx, ny, nz);
}
