/*
 * pfc_oracle.c -- CPU restatement of OpenPFC's spectral PFC path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark load it to produce the
 * expected answers.  Nothing in the product (the HIP library, the C++ host
 * layer, the Python bindings) links or calls it.
 *
 * What it restates (reference paths relative to the OpenPFC source tree):
 *   - 3-D R2C/C2R on one rank, r2c along x, x-fastest layout, unnormalised
 *     forward, backward scaled by 1/(nx*ny*nz):
 *       src/openpfc/runtime/cpu/fft.cpp:58-70,114-129
 *       include/openpfc/kernel/fft/detail/fft_heffte_backend.hpp:177-250
 *     The FFT arithmetic itself lives in HeFFTe 2.4.1 -> FFTW3 in the
 *     reference (third-party, not buildable here); this file uses its own
 *     mixed-radix Stockham transform, pinned by brute-force DFTs and by
 *     numpy.fft fixtures.
 *   - k-space helpers: include/openpfc/kernel/fft/kspace.hpp:97-103,126-131,165-167
 *   - ETD coefficients: include/openpfc/kernel/integrator/spectral_exp_coefficients.hpp:68-82
 *   - plain conserved PFC operators and step:
 *       docs/api/examples/10_complete_pfc_model.cpp:121-144 (step)
 *       docs/api/examples/10_complete_pfc_model.cpp:166-213 (operators)
 *   - initial conditions:
 *       include/openpfc/kernel/simulation/initial_conditions/constant.hpp:82-89
 *       include/openpfc/kernel/simulation/initial_conditions/single_seed.hpp:51-83
 *       include/openpfc/kernel/field/operations.hpp:81-105 (x-fastest loop)
 *       include/openpfc/kernel/data/domain.hpp:185-189 (x = origin + i*dx)
 *   - adjoint of the plain PFC step (its transpose; no reference code)
 *   - free energy of the plain PFC model (the functional of that step's
 *     PDE; no reference code)
 *   - FixedBC boundary condition:
 *       include/openpfc/kernel/simulation/boundary_conditions/fixed_bc.hpp
 *   - rotated seeds, SeedGrid, RandomSeeds (and the libstdc++ evaluation of
 *     std::uniform_real_distribution over mt19937_64 they draw from):
 *       include/openpfc/kernel/simulation/initial_conditions/seed.hpp:39-143
 *       include/openpfc/kernel/simulation/initial_conditions/seed_grid.hpp:76-119
 *       include/openpfc/kernel/simulation/initial_conditions/random_seeds.hpp:54-90
 *   - MovingBC boundary condition (one rank):
 *       include/openpfc/kernel/simulation/boundary_conditions/moving_bc.hpp:107-219
 *   - Tungsten mean-field PFC operators and step:
 *       apps/tungsten/include/tungsten/common/tungsten_spectral.hpp:83-104
 *       apps/tungsten/include/tungsten/cpu/tungsten_model.hpp:223-366
 *       apps/tungsten/include/tungsten/common/tungsten_etd_workspace.hpp:99-148
 *
 * The uniform-noise initial condition has no reference counterpart; it is
 * defined here as lo + (hi-lo) * ((mt19937_64() >> 11) * 2^-53) in x-fastest
 * order, so it does not depend on a C++ library's distribution code.
 *
 * Build: see oracle/Makefile.  Compiled with -ffp-contract=off so products
 * and sums round exactly as written (x86-64 baseline code, as the reference
 * is compiled without FMA contraction by default).
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cdbl;

/* wall clock and the transform-time meter of orc_pfc_step */
static double g_fft_seconds = 0.0;
static double orc_wtime(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

/* ------------------------------------------------------------------------ */
/* mt19937_64 (the C++ standard's std::mt19937_64 engine, [rand.predef]).    */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t mt[312];
  int mti;
} orc_mt64;

void orc_mt64_seed(orc_mt64 *g, uint64_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 312; ++i) {
    g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) +
               (uint64_t)i;
  }
  g->mti = 312;
}

uint64_t orc_mt64_next(orc_mt64 *g) {
  static const uint64_t mag01[2] = {0ULL, 0xB5026F5AA96619E9ULL};
  const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
  if (g->mti >= 312) {
    int i;
    for (i = 0; i < 312 - 156; ++i) {
      uint64_t x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
      g->mt[i] = g->mt[i + 156] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    }
    for (; i < 311; ++i) {
      uint64_t x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
      g->mt[i] = g->mt[i + (156 - 312)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    }
    uint64_t x = (g->mt[311] & UM) | (g->mt[0] & LM);
    g->mt[311] = g->mt[155] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    g->mti = 0;
  }
  uint64_t x = g->mt[g->mti++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= (x >> 43);
  return x;
}

/* n-th output (1-based) of a generator seeded with `seed`, for the
 * standard's known-answer check (10000th output of seed 5489). */
uint64_t orc_mt64_nth(uint64_t seed, long n) {
  orc_mt64 g;
  orc_mt64_seed(&g, seed);
  uint64_t v = 0;
  for (long i = 0; i < n; ++i) v = orc_mt64_next(&g);
  return v;
}

/* out[i] += lo + (hi-lo)*u_i, u_i = (r_i >> 11) * 2^-53, i in x-fastest order. */
void orc_add_uniform_noise(double *out, size_t n, uint64_t seed, double lo,
                           double hi) {
  orc_mt64 g;
  orc_mt64_seed(&g, seed);
  const double w = hi - lo;
  for (size_t i = 0; i < n; ++i) {
    double u = (double)(orc_mt64_next(&g) >> 11) * 0x1.0p-53;
    out[i] += lo + w * u;
  }
}

/* ------------------------------------------------------------------------ */
/* 1-D mixed-radix Stockham FFT.                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  int nf;
  int fac[40];
  cdbl *w; /* w[m] = exp(-2*pi*i*m/n), m = 0..n-1 */
} orc_plan1d;

static void plan1d_init(orc_plan1d *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
  }
  p->w = (cdbl *)malloc(sizeof(cdbl) * (size_t)(n > 0 ? n : 1));
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n; ++k) {
    long double a = two_pi * (long double)k / (long double)n;
    p->w[k] = (double)cosl(a) - I * (double)sinl(a);
  }
  /* exact values on the axes */
  if (n % 4 == 0) {
    p->w[n / 4] = -I;
    p->w[n / 2] = -1.0;
    p->w[3 * n / 4] = I;
  } else if (n % 2 == 0) {
    p->w[n / 2] = -1.0;
  }
  if (n > 0) p->w[0] = 1.0;
}

static void plan1d_free(orc_plan1d *p) { free(p->w); }

/* One Stockham radix-r pass: x (length n) -> y. */
static void stockham_pass(const orc_plan1d *p, int r, int ns, const cdbl *x,
                          cdbl *y, int inverse) {
  const int n = p->n;
  const int nr = n / r;
  const int tstep = n / (ns * r);
  cdbl a[64], b[64];
  for (int j = 0; j < nr; ++j) {
    const int k = j % ns;
    for (int q = 0; q < r; ++q) {
      cdbl v = x[j + q * nr];
      if (q && k) {
        cdbl t = p->w[(long)q * k * tstep % n];
        if (inverse) t = conj(t);
        v = v * t;
      }
      a[q] = v;
    }
    if (r == 2) {
      b[0] = a[0] + a[1];
      b[1] = a[0] - a[1];
    } else if (r == 4) {
      cdbl t0 = a[0] + a[2], t1 = a[0] - a[2];
      cdbl t2 = a[1] + a[3], t3 = a[1] - a[3];
      /* forward: multiply t3 by -i ; inverse: by +i */
      cdbl t3r = inverse ? (-cimag(t3) + I * creal(t3))
                         : (cimag(t3) - I * creal(t3));
      b[0] = t0 + t2;
      b[2] = t0 - t2;
      b[1] = t1 + t3r;
      b[3] = t1 - t3r;
    } else {
      for (int s = 0; s < r; ++s) {
        cdbl acc = 0.0;
        for (int q = 0; q < r; ++q) {
          cdbl t = p->w[(long)q * s * nr % n];
          if (inverse) t = conj(t);
          acc += a[q] * t;
        }
        b[s] = acc;
      }
    }
    const int base = (j / ns) * ns * r + k;
    for (int s = 0; s < r; ++s) y[base + s * ns] = b[s];
  }
}

/* In-place transform of x (length p->n) with scratch of the same length. */
static void fft1d_exec(const orc_plan1d *p, cdbl *x, cdbl *scratch,
                       int inverse) {
  cdbl *src = x, *dst = scratch;
  int ns = 1;
  for (int f = 0; f < p->nf; ++f) {
    stockham_pass(p, p->fac[f], ns, src, dst, inverse);
    ns *= p->fac[f];
    cdbl *t = src;
    src = dst;
    dst = t;
  }
  if (src != x) memcpy(x, src, sizeof(cdbl) * (size_t)p->n);
}

/* Public 1-D complex transform, unnormalised; sign -1 forward, +1 inverse. */
int orc_fft1d(int n, double *data, int sign) {
  if (n < 1) return -1;
  orc_plan1d p;
  plan1d_init(&p, n);
  cdbl *s = (cdbl *)malloc(sizeof(cdbl) * (size_t)n);
  fft1d_exec(&p, (cdbl *)data, s, sign > 0);
  free(s);
  plan1d_free(&p);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* 3-D R2C / C2R in the reference layout (r2c along x, x fastest).           */
/* ------------------------------------------------------------------------ */

/* Real row of even length n -> half spectrum X[0..n/2], via an n/2-point
 * complex transform of z[m] = x[2m] + i x[2m+1]. */
static void r2c_row(const orc_plan1d *ph, const double *x, cdbl *X, cdbl *z,
                    cdbl *s, const cdbl *wn, int n) {
  const int M = n / 2;
  for (int m = 0; m < M; ++m) z[m] = x[2 * m] + I * x[2 * m + 1];
  fft1d_exec(ph, z, s, 0);
  for (int k = 0; k <= M; ++k) {
    cdbl zk = z[k % M];
    cdbl zc = conj(z[(M - k) % M]);
    cdbl e = 0.5 * (zk + zc);
    cdbl o = -0.5 * I * (zk - zc);
    X[k] = e + wn[k] * o;
  }
}

/* Half spectrum X[0..n/2] -> real row (unnormalised inverse DFT). The
 * imaginary parts of X[0] and X[n/2] are ignored, as a half-complex c2r
 * transform does. */
static void c2r_row(const orc_plan1d *ph, const cdbl *X, double *x, cdbl *z,
                    cdbl *s, const cdbl *wn, int n) {
  const int M = n / 2;
  for (int k = 0; k < M; ++k) {
    cdbl a = X[k], b = conj(X[M - k]);
    if (k == 0) {
      a = creal(X[0]);
      b = creal(X[M]);
    }
    cdbl e = a + b;
    cdbl o = (a - b) * conj(wn[k]);
    z[k] = e + I * o;
  }
  fft1d_exec(ph, z, s, 1);
  for (int m = 0; m < M; ++m) {
    x[2 * m] = creal(z[m]);
    x[2 * m + 1] = cimag(z[m]);
  }
}

/* Odd-length x rows: full complex transform of the real row. */
static void r2c_row_odd(const orc_plan1d *pn, const double *x, cdbl *X,
                        cdbl *z, cdbl *s, int n) {
  for (int i = 0; i < n; ++i) z[i] = x[i];
  fft1d_exec(pn, z, s, 0);
  for (int k = 0; k <= n / 2; ++k) X[k] = z[k];
}

static void c2r_row_odd(const orc_plan1d *pn, const cdbl *X, double *x,
                        cdbl *z, cdbl *s, int n) {
  z[0] = creal(X[0]);
  for (int k = 1; k <= n / 2; ++k) {
    z[k] = X[k];
    z[n - k] = conj(X[k]);
  }
  fft1d_exec(pn, z, s, 1);
  for (int i = 0; i < n; ++i) x[i] = creal(z[i]);
}

#define ORC_CB 8 /* columns gathered together for strided axes */

/* Complex transforms along y (axis 1) or z (axis 2) of a (w, ny, nz) array
 * with x fastest. */
static void c2c_axis(cdbl *a, int w, int ny, int nz, int axis, int inverse) {
  const int n = axis == 1 ? ny : nz;
  if (n == 1) return;
  orc_plan1d p;
  plan1d_init(&p, n);
  const long stride = axis == 1 ? (long)w : (long)w * ny;
  const int nouter = axis == 1 ? nz : ny;
  const long ostride = axis == 1 ? (long)w * ny : (long)w;
  const int ntiles = (w + ORC_CB - 1) / ORC_CB;
#pragma omp parallel
  {
    cdbl *buf = (cdbl *)malloc(sizeof(cdbl) * (size_t)n * ORC_CB);
    cdbl *s = (cdbl *)malloc(sizeof(cdbl) * (size_t)n);
#pragma omp for schedule(static)
    for (long t = 0; t < (long)ntiles * nouter; ++t) {
      const int o = (int)(t / ntiles);
      const int x0 = (int)(t % ntiles) * ORC_CB;
      const int cw = (x0 + ORC_CB <= w) ? ORC_CB : (w - x0);
      cdbl *base = a + o * ostride + x0;
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < cw; ++c) buf[c * n + i] = base[i * stride + c];
      for (int c = 0; c < cw; ++c) fft1d_exec(&p, buf + c * n, s, inverse);
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < cw; ++c) base[i * stride + c] = buf[c * n + i];
    }
    free(buf);
    free(s);
  }
  plan1d_free(&p);
}

static cdbl *make_wn(int n) {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  cdbl *wn = (cdbl *)malloc(sizeof(cdbl) * (size_t)(n / 2 + 1));
  for (int k = 0; k <= n / 2; ++k) {
    long double a = two_pi * (long double)k / (long double)n;
    wn[k] = (double)cosl(a) - I * (double)sinl(a);
  }
  return wn;
}

/* Forward R2C: in real (nx,ny,nz) x fastest; out complex (nx/2+1,ny,nz)
 * as interleaved doubles.  Unnormalised.  Follows the reference's
 * fft3d_r2c(r2c_direction=0) contract (fft.cpp:58-70,114-121). */
int orc_r2c3d(int nx, int ny, int nz, const double *in, double *out_) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  cdbl *out = (cdbl *)out_;
  const int w = nx / 2 + 1;
  const long rows = (long)ny * nz;
  const int even = (nx % 2 == 0) && nx >= 2;
  orc_plan1d p;
  plan1d_init(&p, even ? nx / 2 : nx);
  cdbl *wn = make_wn(nx);
#pragma omp parallel
  {
    cdbl *z = (cdbl *)malloc(sizeof(cdbl) * (size_t)nx);
    cdbl *s = (cdbl *)malloc(sizeof(cdbl) * (size_t)nx);
#pragma omp for schedule(static)
    for (long r = 0; r < rows; ++r) {
      if (even)
        r2c_row(&p, in + r * nx, out + r * w, z, s, wn, nx);
      else
        r2c_row_odd(&p, in + r * nx, out + r * w, z, s, nx);
    }
    free(z);
    free(s);
  }
  free(wn);
  plan1d_free(&p);
  c2c_axis(out, w, ny, nz, 1, 0);
  c2c_axis(out, w, ny, nz, 2, 0);
  return 0;
}

/* Backward C2R scaled by 1/(nx*ny*nz) (heffte::scale::full,
 * fft_heffte_backend.hpp:243).  `in` is not modified; `work` must hold
 * (nx/2+1)*ny*nz complex values (may be NULL: allocated internally). */
int orc_c2r3d(int nx, int ny, int nz, const double *in_, double *out,
              double *work_) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const int w = nx / 2 + 1;
  const size_t nc = (size_t)w * ny * nz;
  cdbl *work = (cdbl *)work_;
  int own = 0;
  if (!work) {
    work = (cdbl *)malloc(sizeof(cdbl) * nc);
    own = 1;
  }
  memcpy(work, in_, sizeof(cdbl) * nc);
  c2c_axis(work, w, ny, nz, 2, 1);
  c2c_axis(work, w, ny, nz, 1, 1);
  const long rows = (long)ny * nz;
  const int even = (nx % 2 == 0) && nx >= 2;
  orc_plan1d p;
  plan1d_init(&p, even ? nx / 2 : nx);
  cdbl *wn = make_wn(nx);
  const double scale = 1.0 / ((double)nx * (double)ny * (double)nz);
#pragma omp parallel
  {
    cdbl *z = (cdbl *)malloc(sizeof(cdbl) * (size_t)nx);
    cdbl *s = (cdbl *)malloc(sizeof(cdbl) * (size_t)nx);
#pragma omp for schedule(static)
    for (long r = 0; r < rows; ++r) {
      double *xr = out + r * nx;
      if (even)
        c2r_row(&p, work + r * w, xr, z, s, wn, nx);
      else
        c2r_row_odd(&p, work + r * w, xr, z, s, nx);
      for (int i = 0; i < nx; ++i) xr[i] *= scale;
    }
    free(z);
    free(s);
  }
  free(wn);
  plan1d_free(&p);
  if (own) free(work);
  return 0;
}

/* Brute-force O(N^2) real-to-half-spectrum DFT for tiny grids (checks). */
int orc_dft_r2c3d_brute(int nx, int ny, int nz, const double *in,
                        double *out_) {
  cdbl *out = (cdbl *)out_;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  const int w = nx / 2 + 1;
  for (int kz = 0; kz < nz; ++kz)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < w; ++kx) {
        long double re = 0, im = 0;
        for (int z = 0; z < nz; ++z)
          for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
              long m1 = (long)kx * x % nx, m2 = (long)ky * y % ny,
                   m3 = (long)kz * z % nz;
              long double a = two_pi * ((long double)m1 / nx +
                                        (long double)m2 / ny +
                                        (long double)m3 / nz);
              long double v = in[((long)z * ny + y) * nx + x];
              re += v * cosl(a);
              im -= v * sinl(a);
            }
        out[((long)kz * ny + ky) * w + kx] = (double)re + I * (double)im;
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* k-space helpers and ETD coefficients.                                     */
/* ------------------------------------------------------------------------ */

/* kspace.hpp:126-131 */
double orc_k_component(int index, int size, double freq_scale) {
  return (index <= size / 2) ? index * freq_scale : (index - size) * freq_scale;
}

/* spectral_exp_coefficients.hpp:68-82 */
void orc_spectral_exp_coeffs(double L, double dt, double thr, double *exp_Ldt,
                             double *phi1) {
  const double arg = L * dt;
  *exp_Ldt = exp(arg);
  if (fabs(L) < thr)
    *phi1 = dt + 0.5 * L * dt * dt;
  else
    *phi1 = expm1(arg) / L;
}

/* 10_complete_pfc_model.cpp:166-213.  Outputs over the outbox, k (z) outer,
 * i (x) inner. */
void orc_pfc_operators(int nx, int ny, int nz, double dx, double dy, double dz,
                       double eps, double dt, double *opL, double *opN) {
  const double two_pi = 6.283185307179586;
  const double fx = two_pi / (dx * nx), fy = two_pi / (dy * ny),
               fz = two_pi / (dz * nz);
  const int w = nx / 2 + 1;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < w; ++i) {
        const size_t idx = ((size_t)k * ny + j) * w + i;
        double kx = (i <= nx / 2) ? i * fx : (i - nx) * fx;
        double ky = (j <= ny / 2) ? j * fy : (j - ny) * fy;
        double kz = (k <= nz / 2) ? k * fz : (k - nz) * fz;
        double k2 = kx * kx + ky * ky + kz * kz;
        double k2_lap = -k2;
        double L = k2_lap * (eps + (1.0 - k2) * (1.0 - k2));
        double oL = exp(L * dt);
        opL[idx] = oL;
        if (fabs(L) > 1e-14)
          opN[idx] = (oL - 1.0) / L * k2_lap;
        else
          opN[idx] = dt * k2_lap;
      }
}

/* ETD combine psi_F = opL*psi_F + opN*N_F (10_complete_pfc_model.cpp:136-139;
 * tungsten_etd_workspace.hpp:137-148). */
static void etd_combine(cdbl *psiF, const cdbl *NF, const double *opL,
                        const double *opN, size_t n) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) {
    double re = opL[i] * creal(psiF[i]) + opN[i] * creal(NF[i]);
    double im = opL[i] * cimag(psiF[i]) + opN[i] * cimag(NF[i]);
    psiF[i] = re + I * im;
  }
}

/* Plain conserved PFC ETD1 step, nsteps times (10_complete_pfc_model.cpp:121-144).
 * work: NULL or 2*(nx/2+1)*ny*nz complex + nx*ny*nz real doubles. */
int orc_pfc_step(int nx, int ny, int nz, const double *opL, const double *opN,
                 double *psi, int nsteps) {
  const size_t nr = (size_t)nx * ny * nz;
  const size_t nc = (size_t)(nx / 2 + 1) * ny * nz;
  cdbl *psiF = (cdbl *)malloc(sizeof(cdbl) * nc);
  cdbl *NF = (cdbl *)malloc(sizeof(cdbl) * nc);
  double *psiN = (double *)malloc(sizeof(double) * nr);
  if (!psiF || !NF || !psiN) {
    free(psiF); free(NF); free(psiN);
    return -2;
  }
  for (int s = 0; s < nsteps; ++s) {
    double t0 = orc_wtime();
    orc_r2c3d(nx, ny, nz, psi, (double *)psiF);
    double t1 = orc_wtime();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nr; ++i) psiN[i] = psi[i] * psi[i] * psi[i];
    double t2 = orc_wtime();
    orc_r2c3d(nx, ny, nz, psiN, (double *)NF);
    double t3 = orc_wtime();
    etd_combine(psiF, NF, opL, opN, nc);
    double t4 = orc_wtime();
    orc_c2r3d(nx, ny, nz, (const double *)psiF, psi, (double *)NF);
    g_fft_seconds += (t1 - t0) + (t3 - t2) + (orc_wtime() - t4);
  }
  free(psiF);
  free(NF);
  free(psiN);
  return 0;
}

/* Seconds spent in the transforms of orc_pfc_step since the last reset (the
 * FFT share the reference's scalability driver reports,
 * apps/tungsten/src/tungsten_scalability.cpp:95-135). */
double orc_fft_seconds(int reset) {
  const double v = g_fft_seconds;
  if (reset) g_fft_seconds = 0.0;
  return v;
}

/* OpenMP threads of the following oracle calls (a CPU-baseline row on one
 * thread). */
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

/* Adjoint of one plain PFC step (no reference counterpart; the mathematics
 * of the step 10_complete_pfc_model.cpp:121-144 fixes it).  The step is
 * psi' = A psi + B psi^3 with A = F^-1 opL F, B = F^-1 opN F (opL, opN real
 * and even in k, so A and B are symmetric).  lambda <- A lambda +
 * 3 psi^2 (B lambda); with dr != NULL, *dr += sum_x lambda_in(x) g(x),
 * g = F^-1(dopL F(psi) + dopN F(psi^3)) the derivative of the step with
 * respect to eps, summed in real space.  Tables are unscaled (the C2R
 * applies 1/N). */
void orc_pfc_operator_derivs(int nx, int ny, int nz, double dx, double dy, double dz,
                             double eps, double dt, double *dopL, double *dopN) {
  const double two_pi = 6.283185307179586;
  const double fx = two_pi / (dx * nx), fy = two_pi / (dy * ny), fz = two_pi / (dz * nz);
  const int w = nx / 2 + 1;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < w; ++i) {
        const size_t idx = ((size_t)k * ny + j) * w + i;
        double kx = orc_k_component(i, nx, fx), ky = orc_k_component(j, ny, fy),
               kz = orc_k_component(k, nz, fz);
        double k2 = kx * kx + ky * ky + kz * kz;
        double L = -k2 * (eps + (1.0 - k2) * (1.0 - k2));
        double oL = exp(L * dt);
        dopL[idx] = dt * (-k2) * oL;
        dopN[idx] = fabs(L) > 1e-14 ? (k2 * k2) * (dt * oL * L - (oL - 1.0)) / (L * L) : 0.0;
      }
}

int orc_pfc_adjoint_step(int nx, int ny, int nz, const double *opL, const double *opN,
                         const double *dopL, const double *dopN, const double *psi,
                         double *lambda, double *dr) {
  const size_t nr = (size_t)nx * ny * nz;
  const size_t nc = (size_t)(nx / 2 + 1) * ny * nz;
  cdbl *L = (cdbl *)malloc(sizeof(cdbl) * nc);
  cdbl *X = (cdbl *)malloc(sizeof(cdbl) * nc);
  cdbl *Y = (cdbl *)malloc(sizeof(cdbl) * nc);
  double *a = (double *)malloc(sizeof(double) * nr);
  double *b = (double *)malloc(sizeof(double) * nr);
  if (!L || !X || !Y || !a || !b) {
    free(L); free(X); free(Y); free(a); free(b);
    return -2;
  }
  if (dr) {
    orc_r2c3d(nx, ny, nz, psi, (double *)X);
    for (size_t i = 0; i < nr; ++i) a[i] = psi[i] * psi[i] * psi[i];
    orc_r2c3d(nx, ny, nz, a, (double *)Y);
    for (size_t i = 0; i < nc; ++i) X[i] = dopL[i] * X[i] + dopN[i] * Y[i];
    orc_c2r3d(nx, ny, nz, (const double *)X, b, (double *)Y);
    double s = 0;
    for (size_t i = 0; i < nr; ++i) s += lambda[i] * b[i];
    *dr += s;
  }
  orc_r2c3d(nx, ny, nz, lambda, (double *)L);
  for (size_t i = 0; i < nc; ++i) {
    X[i] = opL[i] * L[i];
    Y[i] = opN[i] * L[i];
  }
  orc_c2r3d(nx, ny, nz, (const double *)X, a, (double *)L);
  orc_c2r3d(nx, ny, nz, (const double *)Y, b, (double *)L);
  for (size_t i = 0; i < nr; ++i) lambda[i] = a[i] + 3.0 * psi[i] * psi[i] * b[i];
  free(L); free(X); free(Y); free(a); free(b);
  return 0;
}

/* Free energy of the plain PFC model (no reference counterpart: the
 * functional whose conserved gradient flow is the step of
 * 10_complete_pfc_model.cpp:121-144):
 *   F = dV [ 1/(2N) sum_k |Psi_k|^2 (eps + (1-k^2)^2) + 1/4 sum_x psi^4 ],
 * the k sum over the half spectrum with weight 2 off the kx = 0 and
 * kx = nx/2 planes, accumulated in long double. */
double orc_pfc_free_energy(int nx, int ny, int nz, double dx, double dy, double dz, double eps,
                           const double *psi) {
  const size_t nr = (size_t)nx * ny * nz;
  const int w = nx / 2 + 1;
  cdbl *X = (cdbl *)malloc(sizeof(cdbl) * (size_t)w * ny * nz);
  if (!X) return NAN;
  orc_r2c3d(nx, ny, nz, psi, (double *)X);
  const double two_pi = 6.283185307179586;
  const double fx = two_pi / (dx * nx), fy = two_pi / (dy * ny), fz = two_pi / (dz * nz);
  long double spec = 0, quart = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < w; ++i) {
        const double kx = orc_k_component(i, nx, fx), ky = orc_k_component(j, ny, fy);
        const double kz = orc_k_component(k, nz, fz);
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double om = 1.0 - k2;
        const cdbl v = X[((size_t)k * ny + j) * w + i];
        const double wt = (i == 0 || 2 * i == nx) ? 1.0 : 2.0;
        spec += (long double)wt * ((long double)creal(v) * creal(v) +
                                   (long double)cimag(v) * cimag(v)) * (eps + om * om);
      }
  for (size_t i = 0; i < nr; ++i) {
    const long double u2 = (long double)psi[i] * psi[i];
    quart += u2 * u2;
  }
  free(X);
  return (double)((long double)dx * dy * dz * (spec / (2.0L * (long double)nr) + 0.25L * quart));
}

/* ------------------------------------------------------------------------ */
/* Tungsten mean-field PFC.                                                  */
/* ------------------------------------------------------------------------ */

/* tungsten_spectral.hpp:83-104.  p = {alpha2, lambda2, alpha_farTol,
 * alpha_highOrd, Bx, T, T0, stabP, p2_bar, q2_bar}. */
void orc_tungsten_physics_for_mode(double k_lap, const double *p,
                                   double *filterMF, double *opCk) {
  const double alpha2 = p[0], lambda2 = p[1], farTol = p[2];
  const int highOrd = (int)p[3];
  const double Bx = p[4], T = p[5], T0 = p[6], stabP = p[7], p2 = p[8],
               q2 = p[9];
  double fMF = exp(k_lap / lambda2);
  double k_val = sqrt(-k_lap) - 1.0;
  double k2 = k_val * k_val;
  double rTol = -alpha2 * log(farTol) - 1.0;
  double g1;
  if (highOrd == 0)
    g1 = exp(-k2 / alpha2);
  else
    g1 = exp(-(k2 + rTol * pow(k_val, highOrd)) / alpha2);
  double g2 = 1.0 - 1.0 / alpha2 * k2;
  double gf = (k_val < 0.0) ? g1 : g2;
  double opPeak = Bx * exp(-T / T0) * gf;
  *opCk = stabP + p2 - opPeak + q2 * fMF;
  *filterMF = fMF;
}

/* cpu/tungsten_model.hpp:223-279 + tungsten_etd_workspace.hpp:99-130:
 * filterMF, exp(L dt) and n_weight = k_lap * phi1(L), L = k_lap * opCk. */
void orc_tungsten_operators(int nx, int ny, int nz, double dx, double dy,
                            double dz, const double *p, double dt,
                            double *filterMF, double *expL, double *nW) {
  const double two_pi = 6.283185307179586;
  const double fx = two_pi / (dx * nx), fy = two_pi / (dy * ny),
               fz = two_pi / (dz * nz);
  const int w = nx / 2 + 1;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < w; ++i) {
        const size_t idx = ((size_t)k * ny + j) * w + i;
        double ki = orc_k_component(i, nx, fx);
        double kj = orc_k_component(j, ny, fy);
        double kk = orc_k_component(k, nz, fz);
        double kLap = -((ki * ki) + (kj * kj) + (kk * kk));
        double fMF, opCk;
        orc_tungsten_physics_for_mode(kLap, p, &fMF, &opCk);
        double L = kLap * opCk, e, phi1;
        orc_spectral_exp_coeffs(L, dt, 1e-12, &e, &phi1);
        filterMF[idx] = fMF;
        expL[idx] = e;
        nW[idx] = kLap * phi1;
      }
}

/* cpu/tungsten_model.hpp:306-366.  q = {p3_bar, p4_bar, q3_bar, q4_bar, stabP}. */
int orc_tungsten_step(int nx, int ny, int nz, const double *filterMF,
                      const double *expL, const double *nW, const double *q,
                      double *psi, double *psiMF_out, int nsteps) {
  const size_t nr = (size_t)nx * ny * nz;
  const size_t nc = (size_t)(nx / 2 + 1) * ny * nz;
  cdbl *psiF = (cdbl *)malloc(sizeof(cdbl) * nc);
  cdbl *tF = (cdbl *)malloc(sizeof(cdbl) * nc);
  cdbl *wk = (cdbl *)malloc(sizeof(cdbl) * nc);
  double *psiMF = (double *)malloc(sizeof(double) * nr);
  double *psiN = (double *)malloc(sizeof(double) * nr);
  const double p3 = q[0], p4 = q[1], q3 = q[2], q4 = q[3], stabP = q[4];
  for (int s = 0; s < nsteps; ++s) {
    orc_r2c3d(nx, ny, nz, psi, (double *)psiF);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nc; ++i) tF[i] = filterMF[i] * psiF[i];
    orc_c2r3d(nx, ny, nz, (const double *)tF, psiMF, (double *)wk);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nr; ++i) {
      double u = psi[i], v = psiMF[i];
      double u2 = u * u, u3 = u * u * u, v2 = v * v, v3 = v * v * v;
      psiN[i] = p3 * u2 + p4 * u3 + q3 * v2 + q4 * v3;
    }
    if (stabP != 0.0) {
#pragma omp parallel for schedule(static)
      for (long i = 0; i < (long)nr; ++i) psiN[i] = psiN[i] - stabP * psi[i];
    }
    orc_r2c3d(nx, ny, nz, psiN, (double *)tF);
    etd_combine(psiF, tF, expL, nW, nc);
    orc_c2r3d(nx, ny, nz, (const double *)psiF, psi, (double *)wk);
  }
  if (psiMF_out) memcpy(psiMF_out, psiMF, sizeof(double) * nr);
  free(psiF); free(tF); free(wk); free(psiMF); free(psiN);
  return 0;
}

/* FixedBC (include/openpfc/kernel/simulation/boundary_conditions/fixed_bc.hpp):
 * xpos = Lx*dx - xwidth; where |x - xpos| < xwidth (x = x0 + i*dx) the field
 * becomes rho_low*S + rho_high*(1 - S), S = 1/(1 + exp(-alpha*(x - xpos))). */
void orc_bc_fixed(int nx, int ny, int nz, double x0, double dx, double rho_low,
                  double rho_high, double xwidth, double alpha, double *psi) {
  const double xpos = ((double)nx * dx) - xwidth;
  size_t idx = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++idx) {
        const double x = x0 + (double)i * dx;
        if (fabs(x - xpos) < xwidth) {
          const double S = 1.0 / (1.0 + exp(-alpha * (x - xpos)));
          psi[idx] = (rho_low * S) + (rho_high * (1.0 - S));
        }
      }
}

/* ------------------------------------------------------------------------ */
/* Initial conditions (field::apply order: k outer, i inner).               */
/* ------------------------------------------------------------------------ */

/* constant.hpp:82-89 */
void orc_ic_constant(double *psi, size_t n, double value) {
  for (size_t i = 0; i < n; ++i) psi[i] = value;
}

/* single_seed.hpp:51-83: bcc one-mode seed with the 6 <110>/sqrt(2) wave
 * vectors inside x^2+y^2+z^2 < radius^2; writes 0.0 outside (the reference
 * hard-codes radius = 64.0 physical units). */
void orc_ic_single_seed(int nx, int ny, int nz, double x0, double y0,
                        double z0, double dx, double dy, double dz,
                        double rho, double amp, double radius, double *psi) {
  const double s = 1.0 / sqrt(2.0);
  const double q[6][3] = {{s, s, 0}, {s, 0, s},  {0, s, s},
                          {s, 0, -s}, {s, -s, 0}, {0, s, -s}};
  const double r2 = pow(radius, 2);
  size_t idx = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double x = x0 + i * dx, y = y0 + j * dy, z = z0 + k * dz;
        double u;
        if (x * x + y * y + z * z >= r2) {
          u = 0.0;
        } else {
          u = rho;
          for (int qi = 0; qi < 6; ++qi)
            u += 2.0 * amp *
                 cos((q[qi][0] * x) + (q[qi][1] * y) + (q[qi][2] * z));
        }
        psi[idx++] = u;
      }
}

/* Counter-based noise (no reference counterpart): psi[i] += lo + (hi-lo)*u_i,
 * u_i = (splitmix64(splitmix64(seed) + i) >> 11) * 2^-53. */
static uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* The mixing function itself, exported so the tests can pin it against the
 * published splitmix64 known answers. */
uint64_t orc_splitmix64(uint64_t z) { return splitmix64(z); }

void orc_add_counter_noise(double *psi, size_t n, uint64_t seed, double lo,
                           double hi) {
  const uint64_t s0 = splitmix64(seed);
  const double w = hi - lo;
  for (size_t i = 0; i < n; ++i) {
    double u = (double)(splitmix64(s0 + (uint64_t)i) >> 11) * 0x1.0p-53;
    psi[i] = psi[i] + (lo + w * u);
  }
}

/* SingleSeed with the outside behaviour selectable: keep_outside = 0 writes
 * 0.0 outside (single_seed.hpp:73-74), 1 leaves psi unchanged there. */
void orc_ic_single_seed_masked(int nx, int ny, int nz, double x0, double y0,
                               double z0, double dx, double dy, double dz,
                               double rho, double amp, double radius,
                               int keep_outside, double *psi) {
  const double s = 1.0 / sqrt(2.0);
  const double q[6][3] = {{s, s, 0}, {s, 0, s},  {0, s, s},
                          {s, 0, -s}, {s, -s, 0}, {0, s, -s}};
  const double r2 = radius * radius;
  size_t idx = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++idx) {
        const double x = x0 + i * dx, y = y0 + j * dy, z = z0 + k * dz;
        if (x * x + y * y + z * z >= r2) {
          if (!keep_outside) psi[idx] = 0.0;
          continue;
        }
        double u = rho;
        for (int qi = 0; qi < 6; ++qi)
          u += 2.0 * amp *
               cos((q[qi][0] * x) + (q[qi][1] * y) + (q[qi][2] * z));
        psi[idx] = u;
      }
}

/* ------------------------------------------------------------------------ */
/* Rotated seeds, SeedGrid, RandomSeeds and MovingBC.                        */
/* ------------------------------------------------------------------------ */

/* std::uniform_real_distribution<double>(a, b) over std::mt19937_64 as
 * libstdc++ evaluates it: generate_canonical<double, 53> takes
 * k = max(1, ceil(53 / 64)) = 1 engine output x and returns x / 2^64 (x
 * rounded to double first), pulled below 1 when the rounding reaches 1
 * ([rand.util.canonical]); the distribution returns u * (b - a) + a. */
double orc_uniform_real(orc_mt64 *g, double a, double b) {
  double u = (double)orc_mt64_next(g) / 18446744073709551616.0;
  if (u >= 1.0) u = nextafter(1.0, 0.0);
  return (u * (b - a)) + a;
}

/* The first n draws of uniform_real_distribution(a, b) on mt19937_64(seed). */
void orc_uniform_real_seq(uint64_t seed, double a, double b, int n, double *out) {
  orc_mt64 g;
  orc_mt64_seed(&g, seed);
  for (int i = 0; i < n; ++i) out[i] = orc_uniform_real(&g, a, b);
}

/* Seed (initial_conditions/seed.hpp:39-143).  A seed is stored as 24
 * doubles: location (3), the six rotated wave vectors (18), rho, radius,
 * amplitude.  R = yaw(o0) * (pitch(o1) * roll(o2)); q_i = R q_i with the
 * <110>/sqrt(2) vectors in the reference's order; every 3x3 product sums
 * from 0.0 over k = 0, 1, 2 as mult3 does. */
static void mat3_mul(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double c = 0.0;
      for (int k = 0; k < 3; k++) c += A[i][k] * B[k][j];
      C[i][j] = c;
    }
}

void orc_seed_make(const double loc[3], const double orient[3], double radius, double rho,
                   double amp, double out[24]) {
  const double ca = cos(orient[0]), sa = sin(orient[0]);
  const double cb = cos(orient[1]), sb = sin(orient[1]);
  const double cc = cos(orient[2]), sc = sin(orient[2]);
  const double Ra[3][3] = {{ca, -sa, 0.0}, {sa, ca, 0.0}, {0.0, 0.0, 1.0}};
  const double Rb[3][3] = {{cb, 0.0, sb}, {0.0, 1.0, 0.0}, {-sb, 0.0, cb}};
  const double Rc[3][3] = {{1.0, 0.0, 0.0}, {0.0, cc, -sc}, {0.0, sc, cc}};
  double Rbc[3][3], R[3][3];
  mat3_mul(Rb, Rc, Rbc);
  mat3_mul(Ra, Rbc, R);
  const double s = 1.0 / sqrt(2.0);
  const double q[6][3] = {{s, s, 0}, {s, 0, s},  {0, s, s},
                          {s, 0, -s}, {s, -s, 0}, {0, s, -s}};
  for (int d = 0; d < 3; ++d) out[d] = loc[d];
  for (int qi = 0; qi < 6; ++qi)
    for (int i = 0; i < 3; ++i) {
      double c = 0.0;
      for (int j = 0; j < 3; ++j) c += R[i][j] * q[qi][j];
      out[3 + 3 * qi + i] = c;
    }
  out[21] = rho;
  out[22] = radius;
  out[23] = amp;
}

/* SeedGrid::apply (initial_conditions/seed_grid.hpp:76-119): Ny*Nz seeds,
 * j outer, k inner; location {X0 + rt, Y0 + Dy*j + rt, Z0 + Dz*k + rt} and
 * orientation {rr, rr, rr} drawn left to right from mt19937_64(42), with
 * rt ~ U(-0.2 radius, 0.2 radius), rr ~ U(0, 8 atan(1)),
 * Dy = dy*Ly/Ny, Dz = dz*Lz/Nz, Y0 = Dy/2, Z0 = Dz/2.  Writes Ny*Nz seeds. */
void orc_seed_grid(int Ly, int Lz, double dy, double dz, int Ny, int Nz, double X0,
                   double radius, double rho, double amp, double *out) {
  const double Dy = dy * (double)Ly / (double)Ny;
  const double Dz = dz * (double)Lz / (double)Nz;
  const double Y0 = Dy / 2.0, Z0 = Dz / 2.0;
  orc_mt64 g;
  orc_mt64_seed(&g, 42);
  const double ta = -0.2 * radius, tb = 0.2 * radius;
  const double rb = 8.0 * atan(1.0);
  int n = 0;
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++) {
      double loc[3], o[3];
      loc[0] = X0 + orc_uniform_real(&g, ta, tb);
      loc[1] = Y0 + (Dy * j) + orc_uniform_real(&g, ta, tb);
      loc[2] = Z0 + (Dz * k) + orc_uniform_real(&g, ta, tb);
      for (int d = 0; d < 3; ++d) o[d] = orc_uniform_real(&g, 0.0, rb);
      orc_seed_make(loc, o, radius, rho, amp, out + 24 * n++);
    }
}

/* RandomSeeds::apply (initial_conditions/random_seeds.hpp:54-90): 150 seeds
 * of radius 20; location {U(-108, -68), U(-128, 128), U(-128, 128)} and
 * orientation {U(0, 2 pi)} x 3 from mt19937_64(42).  Writes 150 seeds. */
void orc_random_seeds(double rho, double amp, double *out) {
  const double radius = 20.0;
  const double lx = -128.0 + radius, ux = -128.0 + (3 * radius);
  const double rb = 8.0 * atan(1.0);
  orc_mt64 g;
  orc_mt64_seed(&g, 42);
  for (int i = 0; i < 150; i++) {
    double loc[3], o[3];
    loc[0] = orc_uniform_real(&g, lx, ux);
    loc[1] = orc_uniform_real(&g, -128.0, 128.0);
    loc[2] = orc_uniform_real(&g, -128.0, 128.0);
    for (int d = 0; d < 3; ++d) o[d] = orc_uniform_real(&g, 0.0, rb);
    orc_seed_make(loc, o, radius, rho, amp, out + 24 * i);
  }
}

/* The field of a seed list (seed_grid.hpp:108-118, random_seeds.hpp:80-89):
 * at X = origin + i*spacing the first seed with |X - location|^2 < radius^2
 * gives rho + sum_i 2*amp*cos(q_i . X) (seed.hpp:124-141, X absolute);
 * outside every seed 0.0, or the field unchanged with keep_outside. */
void orc_ic_seeds(int nx, int ny, int nz, double x0, double y0, double z0, double dx,
                  double dy, double dz, const double *seeds, int nseeds, int keep_outside,
                  double *psi) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t idx = ((size_t)k * ny + j) * nx + i;
        const double X = x0 + i * dx, Y = y0 + j * dy, Z = z0 + k * dz;
        int hit = -1;
        for (int s = 0; s < nseeds && hit < 0; ++s) {
          const double *sd = seeds + 24 * s;
          const double x = X - sd[0], y = Y - sd[1], z = Z - sd[2];
          const double r = sd[22];
          if (x * x + y * y + z * z < r * r) hit = s;
        }
        if (hit < 0) {
          if (!keep_outside) psi[idx] = 0.0;
          continue;
        }
        const double *sd = seeds + 24 * hit;
        double u = sd[21];
        const double a = sd[23];
        for (int qi = 0; qi < 6; ++qi) {
          const double *q = sd + 3 + 3 * qi;
          u += 2.0 * a * cos((q[0] * X) + (q[1] * Y) + (q[2] * Z));
        }
        psi[idx] = u;
      }
}

/* MovingBC's column profile (boundary_conditions/moving_bc.hpp:120-127):
 * out[i] = max over (j, k) of psi, starting from the lowest double. */
void orc_column_max(int nx, int ny, int nz, const double *psi, double *out) {
  for (int i = 0; i < nx; ++i) out[i] = -DBL_MAX;
  size_t idx = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++idx)
        if (psi[idx] > out[i]) out[i] = psi[idx];
}

/* MovingBC::apply and fill_bc (moving_bc.hpp:107-219) on one rank.  State:
 * state_i = {first, idx}, state_d = {xpos}.  On the first call idx is the
 * last column whose maximum exceeds threshold (unchanged if none); later
 * calls advance idx while column idx % nx exceeds it, at most nx columns.
 * xpos = max(x0 + idx*dx + disp, xpos).  Then, with l = nx*dx and
 * p = fmod(xpos, l), every point whose x lies within xwidth of p (or of
 * p -+ l where the band wraps) becomes rho_low*S + rho_high*(1-S),
 * S = 1/(1 + exp(-alpha*d)).  The column maxima go to xline (nx). */
void orc_bc_moving(int nx, int ny, int nz, double x0, double dx, double rho_low,
                   double rho_high, double xwidth, double alpha, double disp,
                   double threshold, int *state_i, double *state_d, double *xline,
                   double *psi) {
  orc_column_max(nx, ny, nz, psi, xline);
  int idx_ = state_i[1];
  if (state_i[0]) {
    for (int i = nx - 1; i >= 0; i--)
      if (xline[i] > threshold) {
        idx_ = i;
        break;
      }
  } else {
    int scanned = 0;
    while (xline[idx_ % nx] > threshold && scanned < nx) {
      idx_ += 1;
      scanned += 1;
    }
  }
  const double new_xpos = x0 + (idx_ * dx) + disp;
  const double old = state_d[0];
  state_d[0] = (new_xpos < old) ? old : new_xpos;
  state_i[0] = 0;
  state_i[1] = idx_;
  const double Lx = (double)nx;
  const double l = Lx * dx;
  const double xpos = fmod(state_d[0], l);
  size_t n = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++n) {
        const double x = x0 + i * dx;
        const double dist = x - xpos;
        double d;
        if (fabs(dist) < xwidth)
          d = dist;
        else if (xpos < xwidth && fabs(dist - l) < xwidth)
          d = dist - l;
        else if (xpos > l - xwidth && fabs(dist + l) < xwidth)
          d = dist + l;
        else
          continue;
        const double S = 1.0 / (1.0 + exp(-alpha * d));
        psi[n] = (rho_low * S) + (rho_high * (1.0 - S));
      }
}

/* ------------------------------------------------------------------------ */
/* Reductions.                                                               */
/* ------------------------------------------------------------------------ */
static double pairwise_sum(const double *a, size_t n, int square) {
  if (n <= 64) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += square ? a[i] * a[i] : a[i];
    return s;
  }
  size_t h = n / 2;
  return pairwise_sum(a, h, square) + pairwise_sum(a + h, n - h, square);
}

double orc_sum(const double *a, size_t n) { return pairwise_sum(a, n, 0); }
double orc_sum_sq(const double *a, size_t n) { return pairwise_sum(a, n, 1); }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
