/*
 * pm_oracle.c -- CPU restatement of the reference PM / P3M force step.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests check the HIP library in
 * particlesimulation_amd/csrc against this file: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the AleksyBalazinski/ParticleSimulation tree).  All quantities are in the
 * reference's code units (include/unitConversions.h:8-46): positions in grid
 * cells, dt = 1, masses pre-multiplied by 4*pi*G*DT^2/H^3, so that
 * lap(phi) = rho and a = -grad(phi).
 *
 * Precision policy
 *   - deposit, FFT, Green's function, finite differences, interpolation and the
 *     direct sum run in double: the oracle is more accurate than the FP32
 *     reference, so the difference oracle-vs-reference is FP32 rounding only.
 *   - the P3M short-range pass runs in float with the reference's exact
 *     accumulation structure (own accumulator + 13 neighbour slots), because
 *     that pass is also timed as the CPU baseline ("port") of the reference.
 *
 * Pinning (what checks each function independently of its own formula) is
 * listed in tests/test_oracle.py; every function here has at least one check.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_NGP 0
#define ORC_CIC 1
#define ORC_TSC 2
#define ORC_TSC_CENTRED 3 /* textbook TSC (nearest grid point as the centre); not in the reference */

#define ORC_TWO_POINT 0
#define ORC_FOUR_POINT 1

#define ORC_DISCRETE_LAPLACIAN 0
#define ORC_S1_OPTIMAL 1
#define ORC_S2_OPTIMAL 2
#define ORC_POOR_MAN 3

#define ORC_S1 0
#define ORC_S2 1

static const double kPi = 3.14159265358979323846;

static inline int wrapi(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

static inline int64_t flat(int x, int y, int z, int nx, int ny) {
  return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * (int64_t)z);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* Assignment weights                                                        */
/* ------------------------------------------------------------------------ */

/* PMMethod::TSCAssignmentFunc, source/pmMethod.cpp:187-198. */
static inline double tsc_f(double d, int t) {
  if (t == 1) return (0.5 + d) * (0.5 + d);
  if (t == 0) return 1.5 - 2.0 * d * d;
  if (t == -1) return (0.5 - d) * (0.5 - d);
  return 0.0;
}

/*
 * Per-axis stencil of one particle coordinate: anchor cell, first offset and
 * up to three weights.  NGP: round (pmMethod.cpp:207-211).  CIC: (int)
 * truncation, weights 1-d and d (pmMethod.cpp:219-230).  TSC: (int)
 * truncation, offsets -1..1 with f(d,t)/2 so that the product of the three
 * axes equals the reference's (d/8)*f*f*f (pmMethod.cpp:250-266).
 * For x >= 0, (int)x == floor(x); floor() extends the rule to x < 0.
 */
static inline int axis_stencil(int scheme, double x, int* first, double w[3]) {
  if (scheme == ORC_NGP) {
    *first = (int)floor(x + 0.5); /* std::round, halves away from zero for x>=0 */
    w[0] = 1.0;
    return 1;
  }
  int c = (int)floor(x);
  double d = x - c;
  if (scheme == ORC_CIC) {
    *first = c;
    w[0] = 1.0 - d;
    w[1] = d;
    return 2;
  }
  if (scheme == ORC_TSC_CENTRED) {
    /* Hockney & Eastwood's TSC: the same per-axis quadratics about the
       nearest grid point, d = x - round(x) in [-1/2, 1/2), so no weight is
       negative (an extension; the reference anchors at floor(x), Q1) */
    c = (int)floor(x + 0.5);
    d = x - c;
  }
  *first = c - 1;
  w[0] = 0.5 * tsc_f(d, -1);
  w[1] = 0.5 * tsc_f(d, 0);
  w[2] = 0.5 * tsc_f(d, 1);
  return 3;
}

/*
 * PMMethod::spreadMass, source/pmMethod.cpp:200-277 with Grid::assignDensity
 * (source/grid.cpp:28-32).  rho[x + nx*(y + ny*z)].
 * wrap != 0: indices are taken modulo the grid (periodic extension).
 * wrap == 0: the reference's unwrapped indexing; an out-of-range index is
 *            undefined behaviour there and an error (-1) here.
 */
int orc_deposit(int64_t n, const float* pos, const float* mass, int nx, int ny, int nz,
                int scheme, int wrap, double* rho) {
  if (scheme < 0 || scheme > 3) return -2;
  int64_t m = (int64_t)nx * ny * nz;
  memset(rho, 0, sizeof(double) * (size_t)m);
  int bad = 0;
  for (int64_t p = 0; p < n; ++p) {
    int fx, fy, fz;
    double wx[3], wy[3], wz[3];
    int cx = axis_stencil(scheme, pos[3 * p + 0], &fx, wx);
    axis_stencil(scheme, pos[3 * p + 1], &fy, wy);
    axis_stencil(scheme, pos[3 * p + 2], &fz, wz);
    double mp = mass[p];
    for (int a = 0; a < cx; ++a) {
      for (int b = 0; b < cx; ++b) {
        for (int c = 0; c < cx; ++c) {
          int x = fx + a, y = fy + b, z = fz + c;
          if (wrap) {
            x = wrapi(x, nx);
            y = wrapi(y, ny);
            z = wrapi(z, nz);
          } else if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) {
            bad = 1;
            continue;
          }
          rho[flat(x, y, z, nx, ny)] += mp * wx[a] * wy[b] * wz[c];
        }
      }
    }
  }
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* FFT (complex, double; radix 2 for powers of two, a direct DFT otherwise) */
/* ------------------------------------------------------------------------ */

/* Iterative radix-2 DIT on one line; sign = -1 forward, +1 backward
   (FFTW/kissfft sign convention, source/FFTWAdapter.cpp:15-17). */
static void fft_line(double* re, double* im, int n, int sign) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1;
    for (int k = 0; k < half; ++k) {
      double ang = sign * 2.0 * kPi * (double)k / (double)len;
      double wr = cos(ang), wi = sin(ang);
      for (int s = 0; s < n; s += len) {
        int a = s + k, b = s + k + half;
        double tr = re[b] * wr - im[b] * wi;
        double ti = re[b] * wi + im[b] * wr;
        re[b] = re[a] - tr;
        im[b] = im[a] - ti;
        re[a] += tr;
        im[a] += ti;
      }
    }
  }
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* Direct DFT of one line of any length (FFTW's definition, the same sign
   convention): X[k] = sum_j x[j] exp(sign 2 pi i j k / n), with (j k) mod n
   indexing an FP64 twiddle table.  tmp holds 2n doubles. */
static void dft_line(double* re, double* im, int n, int sign, double* tmp) {
  double* tr = tmp;
  double* ti = tmp + n;
  for (int k = 0; k < n; ++k) {
    double sr = 0.0, si = 0.0;
    int64_t idx = 0;
    for (int j = 0; j < n; ++j) {
      double ang = sign * 2.0 * kPi * (double)idx / (double)n;
      double wr = cos(ang), wi = sin(ang);
      sr += re[j] * wr - im[j] * wi;
      si += re[j] * wi + im[j] * wr;
      idx += k;
      if (idx >= n) idx -= n;
    }
    tr[k] = sr;
    ti[k] = si;
  }
  memcpy(re, tr, sizeof(double) * (size_t)n);
  memcpy(im, ti, sizeof(double) * (size_t)n);
}

/*
 * Unnormalised 3-D complex DFT in place on separate re/im arrays laid out
 * x-fastest (the FFTW plan {Nz,Ny,Nx}, source/FFTWAdapter.cpp:15 with dims from
 * source/demos.cpp:126-127).  sign = -1 forward, +1 backward.
 */
int orc_fft3d(int nx, int ny, int nz, double* re, double* im, int sign) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return -2;
  int dims[3] = {nx, ny, nz};
  int64_t strides[3] = {1, nx, (int64_t)nx * ny};
  for (int ax = 0; ax < 3; ++ax) {
    int len = dims[ax];
    int64_t st = strides[ax];
    int o1 = ax == 0 ? ny : nx;
    int o2 = ax == 2 ? ny : nz;
    int64_t s1 = ax == 0 ? nx : 1;
    int64_t s2 = ax == 2 ? nx : (int64_t)nx * ny;
#pragma omp parallel
    {
      double* lr = (double*)malloc(sizeof(double) * (size_t)len);
      double* li = (double*)malloc(sizeof(double) * (size_t)len);
      double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)len);
#pragma omp for collapse(2) schedule(static)
      for (int b = 0; b < o2; ++b) {
        for (int a = 0; a < o1; ++a) {
          int64_t base = a * s1 + b * s2;
          for (int i = 0; i < len; ++i) {
            lr[i] = re[base + i * st];
            li[i] = im[base + i * st];
          }
          if (is_pow2(len)) fft_line(lr, li, len, sign);
          else dft_line(lr, li, len, sign, tmp);
          for (int i = 0; i < len; ++i) {
            re[base + i * st] = lr[i];
            im[base + i * st] = li[i];
          }
        }
      }
      free(lr);
      free(li);
      free(tmp);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Green's functions, source/greensFunctions.cpp (restated in double)        */
/* ------------------------------------------------------------------------ */

static inline double sinc(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

/* S1Fourier(k, a), greensFunctions.cpp:42-45: 3(sin u - u cos u)/u^3, u = ka/2.
   Small u uses the series 3*sum_{m>=1} (-1)^(m+1) 2m u^(2m-2)/(2m+1)!. */
double orc_s1_fourier(double k, double a) {
  double u = k * a / 2.0;
  if (fabs(u) < 0.5) {
    double u2 = u * u, term = 1.0, sum = 0.0, fact = 6.0; /* (2m+1)! for m=1 */
    for (int m = 1; m <= 8; ++m) {
      sum += ((m & 1) ? 1.0 : -1.0) * 2.0 * m * term / fact;
      term *= u2;
      fact *= (2.0 * m + 2.0) * (2.0 * m + 3.0);
    }
    return 3.0 * sum;
  }
  return -3.0 / (u * u * u) * (u * cos(u) - sin(u));
}

/* S2Fourier(k, a), greensFunctions.cpp:47-50: 12(2 - 2cos u - u sin u)/u^4.
   Small u uses 12*sum_{m>=2} (-1)^m (2m-2) u^(2m-4)/(2m)!. */
double orc_s2_fourier(double k, double a) {
  double u = k * a / 2.0;
  if (fabs(u) < 0.5) {
    double u2 = u * u, term = 1.0, sum = 0.0, fact = 24.0; /* (2m)! for m=2 */
    for (int m = 2; m <= 9; ++m) {
      sum += ((m & 1) ? -1.0 : 1.0) * (2.0 * m - 2.0) * term / fact;
      term *= u2;
      fact *= (2.0 * m + 1.0) * (2.0 * m + 2.0);
    }
    return 12.0 * sum;
  }
  double u2 = u * u;
  return 12.0 / (u2 * u2) * (2.0 - 2.0 * cos(u) - u * sin(u));
}

/* Closed-form alias sums, greensFunctions.cpp:102-120. */
static double alias_sum(int is, const double k[3]) {
  double s = 1.0;
  if (is == ORC_TSC) {
    for (int i = 0; i < 3; ++i) {
      double sn = sin(k[i] / 2.0), s2 = sn * sn;
      s *= 1.0 - s2 + 2.0 / 15.0 * s2 * s2;
    }
    return s;
  }
  if (is == ORC_CIC) {
    for (int i = 0; i < 3; ++i) {
      double c = cos(k[i] / 2.0);
      s *= 1.0 + 2.0 * c * c;
    }
    return s / 27.0;
  }
  return 1.0;
}

/* Finite-difference operator D(k) / i: D2Fourier and D4Fourier,
   greensFunctions.cpp:70-89 (alpha = 4/3). */
static inline double fd_d(int fds, double k) {
  if (fds == ORC_TWO_POINT) return sin(k);
  const double alpha = 4.0 / 3.0;
  return alpha * sin(k) + (1.0 - alpha) * sin(2.0 * k) / 2.0;
}

/* D(k)/i for one axis, exported so tests can hold it against the field
   stencils of orc_field (a plane wave through the stencil). */
double orc_fd_d(int fds, double k) { return fd_d(fds, k); }

/* Exact alias sum of U^2 with window [-lim, lim]^3: used by tests to pin the
   closed forms of alias_sum() (they must agree as lim grows). */
double orc_alias_sum_bruteforce(int is, double kx, double ky, double kz, int lim) {
  int p = is == ORC_TSC ? 3 : (is == ORC_CIC ? 2 : 1);
  double k[3] = {kx, ky, kz};
  double tot = 1.0;
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int n = -lim; n <= lim; ++n) {
      double u = sinc((k[i] + 2.0 * kPi * n) / 2.0);
      s += pow(u, 2 * p);
    }
    tot *= s;
  }
  return tot;
}

double orc_alias_sum_closed(int is, double kx, double ky, double kz) {
  double k[3] = {kx, ky, kz};
  return alias_sum(is, k);
}

/*
 * GreenOptimal, greensFunctions.cpp:122-189, for integer mode (kx,ky,kz)
 * taken UNFOLDED in [0,N) exactly as PMMethod::initGreensFunction passes it
 * (source/pmMethod.cpp:164-185, source/grid.cpp:21-26).
 * The numerator D . conj(sum U^2 R) is real: D_i = i d_i and R_i = -i r_i
 * give D_i conj(R_i) = -d_i r_i.
 * Where sum d_i^2 == 0 exactly (every k_i in {0, pi}) the mode carries no
 * force (E = -D phi = 0) and G is set to 0; the FP32 reference divides by
 * sinf(pi_f)^2 ~ 8e-15 there.
 */
double orc_green_optimal(int is, int kx, int ky, int kz, int nx, int ny, int nz, double a,
                         int cs, int fds) {
  if (kx == 0 && ky == 0 && kz == 0) return 0.0;
  int ki[3] = {kx, ky, kz}, ni[3] = {nx, ny, nz};
  double k[3];
  int d_zero = 1;
  for (int i = 0; i < 3; ++i) {
    k[i] = 2.0 * kPi * (double)ki[i] / (double)ni[i];
    if ((2 * ki[i]) % ni[i] != 0) d_zero = 0;
  }
  if (d_zero) return 0.0;
  double denom = alias_sum(is, k);
  double d[3], dnorm2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    d[i] = fd_d(fds, k[i]);
    dnorm2 += d[i] * d[i];
  }
  int p = is == ORC_TSC ? 3 : (is == ORC_CIC ? 2 : 1);
  const int lim = 2;
  double usq_axis[3][5], kn_axis[3][5];
  for (int i = 0; i < 3; ++i) {
    for (int n = -lim; n <= lim; ++n) {
      double kn = k[i] + 2.0 * kPi * n;
      kn_axis[i][n + lim] = kn;
      usq_axis[i][n + lim] = pow(sinc(kn / 2.0), 2 * p);
    }
  }
  double r[3] = {0.0, 0.0, 0.0};
  for (int n1 = 0; n1 < 5; ++n1) {
    for (int n2 = 0; n2 < 5; ++n2) {
      for (int n3 = 0; n3 < 5; ++n3) {
        double usq = usq_axis[0][n1] * usq_axis[1][n2] * usq_axis[2][n3];
        double kx_ = kn_axis[0][n1], ky_ = kn_axis[1][n2], kz_ = kn_axis[2][n3];
        double k2 = kx_ * kx_ + ky_ * ky_ + kz_ * kz_;
        double kl = sqrt(k2);
        double s = cs == ORC_S1 ? orc_s1_fourier(kl, a) : orc_s2_fourier(kl, a);
        double f = usq * s * s / k2;
        r[0] += f * kx_;
        r[1] += f * ky_;
        r[2] += f * kz_;
      }
    }
  }
  double num = -(d[0] * r[0] + d[1] * r[1] + d[2] * r[2]);
  return num / (dnorm2 * denom * denom);
}

/* GreenDiscreteLaplacian, greensFunctions.cpp:191-200. */
double orc_green_discrete_laplacian(int kx, int ky, int kz, int nx, int ny, int nz) {
  if (kx == 0 && ky == 0 && kz == 0) return 0.0;
  double sx = sin(kPi * kx / nx), sy = sin(kPi * ky / ny), sz = sin(kPi * kz / nz);
  return -0.25 / (sx * sx + sy * sy + sz * sz);
}

/* GreenPoorMan, greensFunctions.cpp:202-220 (folded k). */
double orc_green_poor_man(int i, int j, int k, int nx, int ny, int nz) {
  if (i == 0 && j == 0 && k == 0) return 0.0;
  int ki = i <= nx / 2 ? i : i - nx;
  int kj = j <= ny / 2 ? j : j - ny;
  int kk = k <= nz / 2 ? k : k - nz;
  double kx = 2.0 * kPi * ki / nx, ky = 2.0 * kPi * kj / ny, kz = 2.0 * kPi * kk / nz;
  return -1.0 / (kx * kx + ky * ky + kz * kz);
}

static double green_at(int gf, int is, int fds, double a, int kx, int ky, int kz, int nx,
                       int ny, int nz) {
  switch (gf) {
    case ORC_DISCRETE_LAPLACIAN: return orc_green_discrete_laplacian(kx, ky, kz, nx, ny, nz);
    case ORC_POOR_MAN: return orc_green_poor_man(kx, ky, kz, nx, ny, nz);
    case ORC_S1_OPTIMAL: return orc_green_optimal(is, kx, ky, kz, nx, ny, nz, a, ORC_S1, fds);
    case ORC_S2_OPTIMAL: return orc_green_optimal(is, kx, ky, kz, nx, ny, nz, a, ORC_S2, fds);
    default: return NAN;
  }
}

/* PMMethod::initGreensFunction, source/pmMethod.cpp:164-185: G on the full
   grid, G[kx + nx*(ky + ny*kz)], unfolded indices, no symmetrisation. */
int orc_green_grid(int gf, int is, int fds, double a, int nx, int ny, int nz, double* G) {
  if (gf < 0 || gf > 3 || is < 0 || is > 3 || fds < 0 || fds > 1) return -2;
  if (is == ORC_TSC_CENTRED) is = ORC_TSC; /* the same assignment function, only re-centred */
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int kz = 0; kz < nz; ++kz)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx)
        G[flat(kx, ky, kz, nx, ny)] = green_at(gf, is, fds, a, kx, ky, kz, nx, ny, nz);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Field, interpolation, external field                                      */
/* ------------------------------------------------------------------------ */

/* getFieldInCell, source/pmMethod.cpp:352-371; phi read with wrapped indices
   (Grid::getPotential, source/grid.cpp:66-68, include/grid.h:48-51). */
static void field_in_cell(const double* phi, int x, int y, int z, int nx, int ny, int nz,
                          int fds, double E[3]) {
#define PHI(i, j, k) phi[flat(wrapi(i, nx), wrapi(j, ny), wrapi(k, nz), nx, ny)]
  if (fds == ORC_TWO_POINT) {
    E[0] = -0.5 * (PHI(x + 1, y, z) - PHI(x - 1, y, z));
    E[1] = -0.5 * (PHI(x, y + 1, z) - PHI(x, y - 1, z));
    E[2] = -0.5 * (PHI(x, y, z + 1) - PHI(x, y, z - 1));
  } else {
    E[0] = (-1.0 / 12.0) * (-PHI(x + 2, y, z) + 8.0 * PHI(x + 1, y, z) -
                            8.0 * PHI(x - 1, y, z) + PHI(x - 2, y, z));
    E[1] = (-1.0 / 12.0) * (-PHI(x, y + 2, z) + 8.0 * PHI(x, y + 1, z) -
                            8.0 * PHI(x, y - 1, z) + PHI(x, y - 2, z));
    E[2] = (-1.0 / 12.0) * (-PHI(x, y, z + 2) + 8.0 * PHI(x, y, z + 1) -
                            8.0 * PHI(x, y, z - 1) + PHI(x, y, z - 2));
  }
#undef PHI
}

/* PMMethod::findFieldInCells, source/pmMethod.cpp:373-382.  E[3*cell + c]. */
int orc_field(const double* phi, int nx, int ny, int nz, int fds, double* E) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) field_in_cell(phi, x, y, z, nx, ny, nz, fds, E + 3 * flat(x, y, z, nx, ny));
  return 0;
}

/* sphRadDecrField, source/externalFields.cpp:4-16 (original units). */
static void sph_rad_decr_field(const double pos[3], const double* prm, double out[3]) {
  /* prm: center xyz, R, M, G */
  double dx = pos[0] - prm[0], dy = pos[1] - prm[1], dz = pos[2] - prm[2];
  double r = sqrt(dx * dx + dy * dy + dz * dz);
  double R = prm[3], M = prm[4], G = prm[5];
  double g = r > R ? -G * M / (r * r) : -(G * M / (R * R * R)) * r * (4.0 - 3.0 * r / R);
  out[0] = g * dx / r;
  out[1] = g * dy / r;
  out[2] = g * dz / r;
}

/*
 * PMMethod::interpolateField + updateAccelerations, source/pmMethod.cpp:279-338,
 * 384-390: acc = sum_w w * E(cell) + accelerationToCodeUnits(ext(x*H)).
 * ext: 0 none, 1 sphRadDecrField with ext_prm = {cx,cy,cz,R,M,G,H,DT}.
 * wrap as in orc_deposit.
 */
int orc_interpolate(int64_t n, const float* pos, const double* E, int nx, int ny, int nz,
                    int scheme, int wrap, int ext, const double* ext_prm, double* acc) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t p = 0; p < n; ++p) {
    int fx, fy, fz;
    double wx[3], wy[3], wz[3];
    int cnt = axis_stencil(scheme, pos[3 * p + 0], &fx, wx);
    axis_stencil(scheme, pos[3 * p + 1], &fy, wy);
    axis_stencil(scheme, pos[3 * p + 2], &fz, wz);
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < cnt; ++i)
      for (int j = 0; j < cnt; ++j)
        for (int k = 0; k < cnt; ++k) {
          int x = fx + i, y = fy + j, z = fz + k;
          if (wrap) {
            x = wrapi(x, nx);
            y = wrapi(y, ny);
            z = wrapi(z, nz);
          } else if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) {
            bad = 1;
            continue;
          }
          const double* e = E + 3 * flat(x, y, z, nx, ny);
          double w = wx[i] * wy[j] * wz[k];
          a[0] += w * e[0];
          a[1] += w * e[1];
          a[2] += w * e[2];
        }
    if (ext == 1) {
      double H = ext_prm[6], DT = ext_prm[7];
      double po[3] = {pos[3 * p] * H, pos[3 * p + 1] * H, pos[3 * p + 2] * H}, g[3];
      sph_rad_decr_field(po, ext_prm, g);
      for (int c = 0; c < 3; ++c) a[c] += DT * DT * g[c] / H; /* unitConversions.h:22-24 */
    }
    acc[3 * p + 0] = a[0];
    acc[3 * p + 1] = a[1];
    acc[3 * p + 2] = a[2];
  }
  return bad ? -1 : 0;
}

/*
 * PMMethod::pmMethodStep, source/pmMethod.cpp:137-144, following the
 * reference's complex-to-complex path literally: complex density with zero
 * imaginary part (source/grid.cpp:5-19), forward C2C, phi_hat = rho_hat * G
 * with phi_hat(0) = 0 (pmMethod.cpp:340-350), backward C2C / N
 * (FFTWAdapter.cpp:39-40), phi = Re (grid.cpp:66-68), then E and the
 * interpolation.  G is the unsymmetrised full grid from orc_green_grid.
 * Optional outputs rho/phi/E may be NULL.
 */
int orc_pm_accel(int64_t n, const float* pos, const float* mass, int nx, int ny, int nz,
                 int scheme, int fds, const double* G, int wrap, int ext,
                 const double* ext_prm, double* acc, double* rho_out, double* phi_out,
                 double* E_out) {
  int64_t m = (int64_t)nx * ny * nz;
  double* re = (double*)malloc(sizeof(double) * (size_t)m);
  double* im = (double*)calloc((size_t)m, sizeof(double));
  double* E = E_out ? E_out : (double*)malloc(sizeof(double) * 3 * (size_t)m);
  if (!re || !im || !E) return -3;
  int rc = orc_deposit(n, pos, mass, nx, ny, nz, scheme, wrap, re);
  if (rc != 0) goto done;
  if (rho_out) memcpy(rho_out, re, sizeof(double) * (size_t)m);
  rc = orc_fft3d(nx, ny, nz, re, im, -1);
  if (rc != 0) goto done;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    re[i] *= G[i];
    im[i] *= G[i];
  }
  re[0] = 0.0;
  im[0] = 0.0;
  rc = orc_fft3d(nx, ny, nz, re, im, +1);
  if (rc != 0) goto done;
  {
    double inv = 1.0 / (double)m;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) re[i] *= inv;
  }
  if (phi_out) memcpy(phi_out, re, sizeof(double) * (size_t)m);
  orc_field(re, nx, ny, nz, fds, E);
  rc = orc_interpolate(n, pos, E, nx, ny, nz, scheme, wrap, ext, ext_prm, acc);
done:
  free(re);
  free(im);
  if (!E_out) free(E);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* P3M short range, source/p3mMethod.cpp + source/chainingMesh.cpp (float)   */
/* ------------------------------------------------------------------------ */

static const float kPiF = 3.14159265358979323846f;

/* P3MMethod::referenceForceS1 / S2, p3mMethod.cpp:194-218 (float). */
float orc_reference_force(int cs, float r, float a) {
  const float G = 1.0f / (4.0f * kPiF);
  if (cs == ORC_S1) {
    if (r >= a) return G / (r * r);
    float q = r / a;
    return G / (a * a) * (8.0f * r / a - 9.0f * r * r / (a * a) + 2.0f * q * q * q * q);
  }
  float u = 2.0f * r / a;
  float u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u, u6 = u3 * u3;
  if (u <= 1.0f)
    return G / (35.0f * a * a) * (224.0f * u - 224.0f * u3 + 70.0f * u4 + 48.0f * u5 - 21.0f * u6);
  if (u <= 2.0f)
    return G / (35.0f * a * a) *
           (12.0f / u2 - 224.0f + 896.0f * u - 840.0f * u2 + 224.0f * u3 + 70.0f * u4 -
            48.0f * u5 + 7.0f * u6);
  return G / (r * r);
}

/* P3MMethod::initSRForceTable, p3mMethod.cpp:275-294 with the constructor's
   500 entries and deltaSquared = re^2/(cnt-1) (p3mMethod.cpp:42-44). */
int orc_sr_table(float rc, float a, float eps, int cs, int cnt, float* table, float* delta2) {
  float d2 = rc * rc / (float)(cnt - 1);
  const float G = 1.0f / (4.0f * kPiF);
  for (int i = 0; i < cnt; ++i) {
    float r2 = (float)i * d2;
    float r = sqrtf(r2);
    float R = -orc_reference_force(cs, r, a);
    float total = -G / (r * r + eps * eps);
    table[i] = r == 0.0f ? 0.0f : (total - R) / sqrtf(r * r + eps * eps);
  }
  *delta2 = d2;
  return 0;
}

/* referenceForceS1/S2 (p3mMethod.cpp:194-218) in double. */
static double ref_force_d(int cs, double r, double a) {
  const double G = 1.0 / (4.0 * kPi);
  if (cs == ORC_S1) {
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

/*
 * Accurate short-range table (an extension; use_table == 2): the direct
 * formula's pair factor F = (-G/(r^2+eps^2) + R(r))/r of shortRangeForce
 * (p3mMethod.cpp:220-238), tabulated over s = r^2 scale in ORC_LOG_ENTRIES
 * bins cut from the float bits of s: bin i holds the floats whose bits >> 16
 * are (lo >> 16) + i, 128 bins per binade (log-spaced, so the r^-3
 * singularity at small r keeps a fixed relative accuracy), and the last bin
 * (bits hi) is zero and takes every larger s.  Within a bin of binade 2^E,
 * s = s0 + (y - 1) 2^E with y the low 16 bits of s set into the mantissa of
 * 1.0f, and F is taken linear in y: F = T + S y, the chord through the bin's
 * ends lowered by half the sum of its largest deviations above and below F.
 * Bits below lo clamp to the first bin.  With no softening and a <= rc
 * (F = 0 for r >= a) the scale is 1 and hi the bin boundary at or above the
 * bits of a^2; otherwise the scale is 1/rc^2 and hi the bits of 1.0 (the
 * cutoff).  Built in FP64.  table receives 2 * ORC_LOG_ENTRIES floats (T, S).
 */
#define ORC_LOG_ENTRIES 4096

int orc_sr_table_log_entries(void) { return ORC_LOG_ENTRIES; }

static double orc_bits_to_double(uint32_t b) {
  float f;
  memcpy(&f, &b, sizeof f);
  return (double)f;
}

static uint32_t orc_float_bits(float f) {
  uint32_t b;
  memcpy(&b, &f, sizeof b);
  return b;
}

static double orc_pair_factor_d(int cs, double s, double a, double e2) {
  const double G = 1.0 / (4.0 * kPi), r = sqrt(s);
  return (-G / (s + e2) + ref_force_d(cs, r, a)) / r;
}

/* The pair factor F(r^2) of shortRangeForce (p3mMethod.cpp:220-238) in
   double, for n squared separations: a_i += m_j F (x_i - x_j) for
   0 < r^2 < rc^2.  r2[k] <= 0 gives 0. */
int orc_sr_pair_factor(int cs, double a, double eps, const double* r2, int64_t n, double* out) {
  for (int64_t k = 0; k < n; ++k) out[k] = r2[k] > 0.0 ? orc_pair_factor_d(cs, r2[k], a, eps * eps) : 0.0;
  return 0;
}

int orc_sr_table_log(float rc, float a, float eps, int cs, float* table, float* scale, uint32_t* lo_bits) {
  const uint32_t bin = 1u << 16;
  float c;
  uint32_t hi;
  if (eps == 0.0f && a <= rc && a >= 1.0e-6f) {
    c = 1.0f;
    hi = (orc_float_bits(a * a) + bin - 1) / bin * bin;
  } else {
    c = 1.0f / (rc * rc);
    hi = orc_float_bits(1.0f);
  }
  const uint32_t lo = hi - (uint32_t)(ORC_LOG_ENTRIES - 1) * bin;
  const double e2 = (double)eps * eps, cd = (double)c;
  for (int i = 0; i + 1 < ORC_LOG_ENTRIES; ++i) {
    uint32_t b0 = lo + (uint32_t)i * bin;
    double s0 = orc_bits_to_double(b0), s1 = orc_bits_to_double(b0 + bin);
    int E;
    frexp(s0, &E);
    double unit = ldexp(1.0, E - 1);
    double f0 = orc_pair_factor_d(cs, s0 / cd, a, e2), f1 = orc_pair_factor_d(cs, s1 / cd, a, e2);
    double dlo = 0.0, dhi = 0.0;
    for (int t = 1; t < 32; ++t) {
      double u = t / 32.0;
      double e = f0 + u * (f1 - f0) - orc_pair_factor_d(cs, (s0 + u * (s1 - s0)) / cd, a, e2);
      if (e < dlo) dlo = e;
      if (e > dhi) dhi = e;
    }
    double S = (f1 - f0) / (s1 - s0) * unit;
    table[2 * i] = (float)(f0 - S - 0.5 * (dlo + dhi));
    table[2 * i + 1] = (float)S;
  }
  table[2 * (ORC_LOG_ENTRIES - 1)] = 0.0f;
  table[2 * (ORC_LOG_ENTRIES - 1) + 1] = 0.0f;
  *scale = c;
  *lo_bits = lo;
  return 0;
}

static inline float orc_sr_table_log_eval(const float* table, float r2, float scale, uint32_t lo) {
  const uint32_t hi = lo + (uint32_t)(ORC_LOG_ENTRIES - 1) * (1u << 16);
  float s = r2 * scale;
  uint32_t b = orc_float_bits(s);
  if (b < lo) b = lo;
  if (b > hi) b = hi;
  uint32_t k = (b >> 16) - (lo >> 16);
  uint32_t yb = (b & 0xFFFFu) | 0x3F800000u;
  float y;
  memcpy(&y, &yb, sizeof y);
  return table[2 * k] + table[2 * k + 1] * y;
}

/* P3MMethod::shortRangeForce, p3mMethod.cpp:220-238 (direct formula). */
static inline void sr_force_direct(const float rij[3], float mi, float mj, int cs, float a,
                                   float eps, float out[3]) {
  float l = sqrtf(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
  float dir[3] = {rij[0] / l, rij[1] / l, rij[2] / l};
  const float G = 1.0f / (4.0f * kPiF);
  float R = orc_reference_force(cs, l, a);
  float tot = -G * mi * mj / (l * l + eps * eps);
  for (int c = 0; c < 3; ++c) out[c] = tot * dir[c] - (-mi * mj * R * dir[c]);
}

typedef struct {
  int mx, my, mz;
  float hcx, hcy, hcz;
} chain_mesh;

/* ChainingMesh::tripleToFlatIndex, chainingMesh.cpp:79-84 (non-periodic). */
static inline int cm_flat(const chain_mesh* cm, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= cm->mx || y >= cm->my || z >= cm->mz) return -1;
  return x + y * cm->mx + z * cm->mx * cm->my;
}

/* ChainingMesh::getNeighborsAndSelf, chainingMesh.cpp:60-77: 9 cells at y-1,
   3 at (y, z-1), (x-1, y, z), then self. */
static void cm_neighbors(const chain_mesh* cm, int q, int nb[14]) {
  int cx = q % cm->mx, cy = (q / cm->mx) % cm->my, cz = q / (cm->mx * cm->my);
  int i = 0;
  for (int t = -1; t <= 1; ++t)
    for (int s = -1; s <= 1; ++s) nb[i++] = cm_flat(cm, cx + t, cy - 1, cz + s);
  for (int s = -1; s <= 1; ++s) nb[i++] = cm_flat(cm, cx + s, cy, cz - 1);
  nb[12] = cm_flat(cm, cx - 1, cy, cz);
  nb[13] = q;
}

typedef struct {
  float y;
  int id;
} yid;

static int cmp_yid(const void* pa, const void* pb) {
  const yid* a = (const yid*)pa;
  const yid* b = (const yid*)pb;
  if (a->y < b->y) return -1;
  if (a->y > b->y) return 1;
  return a->id < b->id ? -1 : (a->id > b->id);
}

/*
 * P3MMethod::calculateShortRangeForces + correctAccelerations,
 * p3mMethod.cpp:50-57, 168-192, 247-273, 296-322, with the chaining mesh of
 * chainingMesh.cpp:6-58.
 *   box_orig[3], rc_orig, H: ChainingMesh(compBoxSize, cutoffRadius, H) in
 *                           original units (M = int(L/rc), HC = (L/M)/H).
 *   rc, a, eps:             cutoff, particle diameter, softening in code units.
 * Cell lists: enableYSorting builds each list in ascending y with ties in
 * insertion (index) order -- the order fillWithYSorting produces -- using a
 * sort instead of the reference's O(n^2) insertion; without y-sorting the
 * list is in descending index order (head insertion, chainingMesh.cpp:45-58).
 * Every (particle, slot) accumulator has a single writer cell, so the cell
 * loop runs on OpenMP threads exactly as the reference's std::thread pool.
 * use_table: 0 direct formula, 1 the reference's table, 2 the accurate table
 * of orc_sr_table_log (an extension, same chaining and accumulation).
 * acc_out (n x 3) receives (sum of slots + own) / m.  sample_stride > 1 takes
 * as home particles only every sample_stride-th entry of each cell's list
 * (the results are then partial): a bounded sample of the pass whose work per
 * cell, and so its balance over the threads, is the whole pass's divided by
 * the stride.  It exists to time the pass on large inputs.
 * Returns the number of ordered pairs within the cutoff in *pairs_out.
 */
int orc_p3m_sr(int64_t n, const float* pos, const float* mass, const float box_orig[3],
               float rc_orig, float H, float rc, float a, float eps, int cs, int use_table,
               int ysort, int sample_stride, float* acc_out,
               int64_t* pairs_out) {
  chain_mesh cm;
  cm.mx = (int)(box_orig[0] / rc_orig);
  cm.my = (int)(box_orig[1] / rc_orig);
  cm.mz = (int)(box_orig[2] / rc_orig);
  if (cm.mx <= 0 || cm.my <= 0 || cm.mz <= 0) return -2;
  cm.hcx = (box_orig[0] / cm.mx) / H;
  cm.hcy = (box_orig[1] / cm.my) / H;
  cm.hcz = (box_orig[2] / cm.mz) / H;
  int ncell = cm.mx * cm.my * cm.mz;
  int* cell_of = (int*)malloc(sizeof(int) * (size_t)n);
  int* count = (int*)calloc((size_t)ncell + 1, sizeof(int));
  int* start = (int*)malloc(sizeof(int) * ((size_t)ncell + 1));
  int* list = (int*)malloc(sizeof(int) * (size_t)n);
  float* own = (float*)calloc(3 * (size_t)n, sizeof(float));
  float* slots = (float*)calloc(39 * (size_t)n, sizeof(float));
  if (sample_stride < 1) return -2;
  if (!cell_of || !count || !start || !list || !own || !slots) return -3;
  for (int64_t i = 0; i < n; ++i) {
    int cx = (int)(pos[3 * i] / cm.hcx), cy = (int)(pos[3 * i + 1] / cm.hcy),
        cz = (int)(pos[3 * i + 2] / cm.hcz);
    int c = cm_flat(&cm, cx, cy, cz);
    if (c < 0) { /* outside the chaining mesh: undefined behaviour in the reference */
      free(cell_of); free(count); free(start); free(list); free(own); free(slots);
      return -1;
    }
    cell_of[i] = c;
    count[c]++;
  }
  start[0] = 0;
  for (int c = 0; c < ncell; ++c) start[c + 1] = start[c] + count[c];
  {
    int* fillpos = (int*)malloc(sizeof(int) * (size_t)ncell);
    memcpy(fillpos, start, sizeof(int) * (size_t)ncell);
    for (int64_t i = 0; i < n; ++i) list[fillpos[cell_of[i]]++] = (int)i;
    free(fillpos);
  }
  if (ysort) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int c = 0; c < ncell; ++c) {
      int s = start[c], e = start[c + 1];
      if (e - s < 2) continue;
      yid* tmp = (yid*)malloc(sizeof(yid) * (size_t)(e - s));
      for (int k = s; k < e; ++k) {
        tmp[k - s].y = pos[3 * (int64_t)list[k] + 1];
        tmp[k - s].id = list[k];
      }
      qsort(tmp, (size_t)(e - s), sizeof(yid), cmp_yid);
      for (int k = s; k < e; ++k) list[k] = tmp[k - s].id;
      free(tmp);
    }
  } else {
    for (int c = 0; c < ncell; ++c) {
      int s = start[c], e = start[c + 1];
      for (int k = 0; k < (e - s) / 2; ++k) {
        int t = list[s + k];
        list[s + k] = list[e - 1 - k];
        list[e - 1 - k] = t;
      }
    }
  }
  float* table = (float*)malloc(sizeof(float) * 2 * (size_t)orc_sr_table_log_entries());
  float d2 = 0.0f, log_scale = 1.0f;
  uint32_t log_lo = 0;
  if (!table) return -1;
  if (use_table == 2) orc_sr_table_log(rc, a, eps, cs, table, &log_scale, &log_lo);
  else orc_sr_table(rc, a, eps, cs, 500, table, &d2);
  const float rc2 = rc * rc;
  int64_t pairs = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : pairs)
  for (int q = 0; q < ncell; ++q) {
    if (count[q] == 0) continue;
    int nb[14];
    cm_neighbors(&cm, q, nb);
    for (int slot = 0; slot < 14; ++slot) {
      int qn = nb[slot];
      if (qn < 0) continue;
      for (int ii = start[q]; ii < start[q + 1]; ii += sample_stride) {
        int i = list[ii];
        for (int jj = start[qn]; jj < start[qn + 1]; ++jj) {
          int j = list[jj];
          if (ysort && pos[3 * (int64_t)j + 1] - pos[3 * (int64_t)i + 1] > rc) break;
          if (i == j) continue;
          float rij[3] = {pos[3 * (int64_t)i] - pos[3 * (int64_t)j],
                          pos[3 * (int64_t)i + 1] - pos[3 * (int64_t)j + 1],
                          pos[3 * (int64_t)i + 2] - pos[3 * (int64_t)j + 2]};
          float r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
          if (r2 >= rc2) continue;
          pairs += qn == q ? 1 : 2;
          float f[3];
          float mi = mass[i], mj = mass[j];
          if (!use_table) {
            sr_force_direct(rij, mi, mj, cs, a, eps, f);
          } else if (use_table == 2) {
            float F = mi * mj * orc_sr_table_log_eval(table, r2, log_scale, log_lo);
            f[0] = F * rij[0];
            f[1] = F * rij[1];
            f[2] = F * rij[2];
          } else {
            /* shortRangeForceFromTable, p3mMethod.cpp:240-245; the index is
               clamped to the table (t+1 <= cnt-1), where the reference reads
               past its end for r^2 within one rounding of rc^2. */
            float ksi = r2 / d2;
            int t = (int)ksi;
            if (t > 498) t = 498;
            float F = mi * mj * (table[t] + (ksi - t) * (table[t + 1] - table[t]));
            f[0] = F * rij[0];
            f[1] = F * rij[1];
            f[2] = F * rij[2];
          }
          own[3 * (int64_t)i] += f[0];
          own[3 * (int64_t)i + 1] += f[1];
          own[3 * (int64_t)i + 2] += f[2];
          if (qn != q) {
            float* sl = slots + 39 * (int64_t)j + 3 * slot;
            sl[0] += -1.0f * f[0];
            sl[1] += -1.0f * f[1];
            sl[2] += -1.0f * f[2];
          }
        }
      }
    }
  }
  /* correctAccelerations, p3mMethod.cpp:50-57: accumulate 13 slots then own. */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    for (int s = 0; s < 13; ++s) {
      tx += slots[39 * i + 3 * s];
      ty += slots[39 * i + 3 * s + 1];
      tz += slots[39 * i + 3 * s + 2];
    }
    tx += own[3 * i];
    ty += own[3 * i + 1];
    tz += own[3 * i + 2];
    acc_out[3 * i] = tx / mass[i];
    acc_out[3 * i + 1] = ty / mass[i];
    acc_out[3 * i + 2] = tz / mass[i];
  }
  if (pairs_out) *pairs_out = pairs;
  free(cell_of); free(count); free(start); free(list); free(own); free(slots);
  free(table);
  return 0;
}

/* Chaining-mesh dimensions M = int(L/rc) per axis (chainingMesh.cpp:10-12). */
int orc_p3m_mesh_dims(const float box_orig[3], float rc_orig, int out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = (int)(box_orig[i] / rc_orig);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Direct sum (FP64), the accuracy oracle of source/ppMethod.cpp:23-36       */
/* ------------------------------------------------------------------------ */

/*
 * acc[t] for each target index: sum_j -Gc * m_j * (r_t - r_j) * K(r), r = |r_t - r_j|.
 *   kernel 0: K = 1/r^3 (eps ignored)               -- Newtonian, used for errors
 *   kernel 1: K = 1/(r^2+eps^2)^(3/2)               -- ppMethod.cpp:31-33 (Plummer)
 *   kernel 2: K = 1/((r^2+eps^2) r)                  -- the P3M "total" force,
 *                                                      p3mMethod.cpp:236
 * Gc is the coupling (1/(4 pi) in code units).
 */
int orc_direct_sum(int64_t nt, const int64_t* targets, int64_t n, const float* pos,
                   const float* mass, double Gc, double eps, int kernel, double* acc) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < nt; ++t) {
    int64_t i = targets ? targets[t] : t;
    double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 == 0.0) continue;
      double k;
      if (kernel == 0) {
        k = 1.0 / (r2 * sqrt(r2));
      } else if (kernel == 1) {
        double s = r2 + eps * eps;
        k = 1.0 / (s * sqrt(s));
      } else {
        k = 1.0 / ((r2 + eps * eps) * sqrt(r2));
      }
      double f = -Gc * (double)mass[j] * k;
      ax += f * dx;
      ay += f * dy;
      az += f * dz;
    }
    acc[3 * t] = ax;
    acc[3 * t + 1] = ay;
    acc[3 * t + 2] = az;
  }
  return 0;
}

/* ------------------------------------------------------------------------
 * Learned influence-function correction (no reference counterpart): the
 * 4 -> 64 -> 64 -> 64 -> 1 GELU network of particlesimulation_amd/csrc/mlp.hip
 * in FP64.  Parameters packed as W1[64][4], b1, W2[64][64], b2, W3[64][64],
 * b3, w4[64], b4 (8705 values, W[out][in]).  Pinned by closed-form single-path
 * networks (forward) and central finite differences (backward).
 * ------------------------------------------------------------------------ */
#define ORC_HID 64
enum { ORC_W1 = 0, ORC_B1 = 256, ORC_W2 = 320, ORC_B2 = 4416, ORC_W3 = 4480, ORC_B3 = 8576, ORC_W4 = 8640,
       ORC_B4 = 8704 };

static inline double orc_gelu(double z) { return 0.5 * z * (1.0 + erf(z * 0.70710678118654752440)); }
static inline double orc_gelu_d(double z) {
  return 0.5 * (1.0 + erf(z * 0.70710678118654752440)) + z * 0.39894228040143267794 * exp(-0.5 * z * z);
}

static double mlp_one(const double* p, const double* x, double z1[ORC_HID], double z2[ORC_HID],
                      double z3[ORC_HID]) {
  double h1[ORC_HID], h2[ORC_HID];
  for (int f = 0; f < ORC_HID; ++f) {
    double z = p[ORC_B1 + f];
    for (int i = 0; i < 4; ++i) z += p[ORC_W1 + 4 * f + i] * x[i];
    z1[f] = z;
    h1[f] = orc_gelu(z);
  }
  for (int f = 0; f < ORC_HID; ++f) {
    double z = p[ORC_B2 + f];
    for (int k = 0; k < ORC_HID; ++k) z += p[ORC_W2 + ORC_HID * f + k] * h1[k];
    z2[f] = z;
    h2[f] = orc_gelu(z);
  }
  double y = p[ORC_B4];
  for (int f = 0; f < ORC_HID; ++f) {
    double z = p[ORC_B3 + f];
    for (int k = 0; k < ORC_HID; ++k) z += p[ORC_W3 + ORC_HID * f + k] * h2[k];
    z3[f] = z;
    y += p[ORC_W4 + f] * orc_gelu(z);
  }
  return y;
}

void orc_mlp_forward(const double* p, int64_t n, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t m = 0; m < n; ++m) {
    double z1[ORC_HID], z2[ORC_HID], z3[ORC_HID];
    y[m] = mlp_one(p, x + 4 * m, z1, z2, z3);
  }
}

/* grad = d/dp sum_m up[m] y(m) */
void orc_mlp_backward(const double* p, int64_t n, const double* x, const double* up, double* grad) {
  memset(grad, 0, sizeof(double) * 8705);
  for (int64_t m = 0; m < n; ++m) {
    double z1[ORC_HID], z2[ORC_HID], z3[ORC_HID], h1[ORC_HID], h2[ORC_HID];
    double dz3[ORC_HID], dz2[ORC_HID], dz1[ORC_HID];
    const double* xm = x + 4 * m;
    mlp_one(p, xm, z1, z2, z3);
    for (int f = 0; f < ORC_HID; ++f) {
      h1[f] = orc_gelu(z1[f]);
      h2[f] = orc_gelu(z2[f]);
    }
    const double d = up[m];
    grad[ORC_B4] += d;
    for (int f = 0; f < ORC_HID; ++f) {
      grad[ORC_W4 + f] += d * orc_gelu(z3[f]);
      dz3[f] = d * p[ORC_W4 + f] * orc_gelu_d(z3[f]);
      grad[ORC_B3 + f] += dz3[f];
      for (int k = 0; k < ORC_HID; ++k) grad[ORC_W3 + ORC_HID * f + k] += dz3[f] * h2[k];
    }
    for (int k = 0; k < ORC_HID; ++k) {
      double s = 0.0;
      for (int f = 0; f < ORC_HID; ++f) s += p[ORC_W3 + ORC_HID * f + k] * dz3[f];
      dz2[k] = s * orc_gelu_d(z2[k]);
    }
    for (int f = 0; f < ORC_HID; ++f) {
      grad[ORC_B2 + f] += dz2[f];
      for (int k = 0; k < ORC_HID; ++k) grad[ORC_W2 + ORC_HID * f + k] += dz2[f] * h1[k];
    }
    for (int k = 0; k < ORC_HID; ++k) {
      double s = 0.0;
      for (int f = 0; f < ORC_HID; ++f) s += p[ORC_W2 + ORC_HID * f + k] * dz2[f];
      dz1[k] = s * orc_gelu_d(z1[k]);
    }
    for (int f = 0; f < ORC_HID; ++f) {
      grad[ORC_B1 + f] += dz1[f];
      for (int i = 0; i < 4; ++i) grad[ORC_W1 + 4 * f + i] += dz1[f] * xm[i];
    }
  }
}

/* ------------------------------------------------------------------------
 * Periodic direct sum (no reference counterpart; the reference's PM is
 * periodic through its FFT, source/FFTWAdapter.cpp, while its direct sum
 * source/ppMethod.cpp:23-36 is isolated).  Ewald summation of the
 * acceleration of a unit mass at the origin on a neutralising background,
 * G = 1:
 *   a(x) = -sum_n (x-nL)/r^3 [erfc(alpha r) + 2 alpha r/sqrt(pi) e^{-alpha^2 r^2}]
 *          - (4 pi/V) sum_{k != 0} k/k^2 e^{-k^2/(4 alpha^2)} sin(k.x),
 *   k = 2 pi (hx/Lx, hy/Ly, hz/Lz).
 * Pinned by independence of alpha, periodicity, antisymmetry and the
 * near-field limit -x/|x|^3 + (4 pi / 3V) x of a cubic box.
 * ------------------------------------------------------------------------ */
void orc_ewald_accel(const double L[3], double alpha, const double x[3], double out[3]) {
  double V = L[0] * L[1] * L[2];
  double a[3] = {0.0, 0.0, 0.0};
  int R[3], Hm[3];
  for (int c = 0; c < 3; ++c) {
    R[c] = (int)ceil(6.0 / (alpha * L[c]) + 0.5);
    Hm[c] = (int)ceil(sqrt(36.0) * 2.0 * alpha * L[c] / (2.0 * kPi));
  }
  for (int i = -R[0]; i <= R[0]; ++i)
    for (int j = -R[1]; j <= R[1]; ++j)
      for (int k = -R[2]; k <= R[2]; ++k) {
        double d[3] = {x[0] - i * L[0], x[1] - j * L[1], x[2] - k * L[2]};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (r2 == 0.0) continue;
        double r = sqrt(r2);
        double f = (erfc(alpha * r) + 2.0 * alpha * r / sqrt(kPi) * exp(-alpha * alpha * r2)) / (r2 * r);
        for (int c = 0; c < 3; ++c) a[c] -= f * d[c];
      }
  for (int i = -Hm[0]; i <= Hm[0]; ++i)
    for (int j = -Hm[1]; j <= Hm[1]; ++j)
      for (int k = -Hm[2]; k <= Hm[2]; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        double kv[3] = {2.0 * kPi * i / L[0], 2.0 * kPi * j / L[1], 2.0 * kPi * k / L[2]};
        double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double e = k2 / (4.0 * alpha * alpha);
        if (e > 36.0) continue;
        double f = 4.0 * kPi / V * exp(-e) / k2 * sin(kv[0] * x[0] + kv[1] * x[1] + kv[2] * x[2]);
        for (int c = 0; c < 3; ++c) a[c] -= f * kv[c];
      }
  for (int c = 0; c < 3; ++c) out[c] = a[c];
}

/* a_t = Gc sum_j m_j a_ewald(x_t - x_j) over all j != t (exact per pair). */
int orc_direct_sum_periodic(int64_t nt, const int64_t* targets, int64_t n, const float* pos, const float* mass,
                            double Gc, const double L[3], double* acc) {
  double Lmin = L[0] < L[1] ? (L[0] < L[2] ? L[0] : L[2]) : (L[1] < L[2] ? L[1] : L[2]);
  double alpha = 2.0 / Lmin;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t t = 0; t < nt; ++t) {
    int64_t i = targets ? targets[t] : t;
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      double d[3], a[3];
      for (int c = 0; c < 3; ++c) {
        d[c] = (double)pos[3 * i + c] - (double)pos[3 * j + c];
        d[c] -= L[c] * floor(d[c] / L[c] + 0.5); /* nearest image; a is periodic */
      }
      if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
      orc_ewald_accel(L, alpha, d, a);
      for (int c = 0; c < 3; ++c) s[c] += (double)mass[j] * a[c];
    }
    for (int c = 0; c < 3; ++c) acc[3 * t + c] = Gc * s[c];
  }
  return 0;
}
