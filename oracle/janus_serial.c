/*
 * C restatement of janus's serial Lippmann-Schwinger path (FP64), the timed
 * CPU baseline.  TEST INFRASTRUCTURE ONLY: loaded by tests/ and by the
 * cpu_baseline leg of bench.py through ctypes, never by janus_amd.
 *
 * janus (sbrisard/janus) itself cannot be built here: its FFT path needs
 * FFTW3, which the image lacks.  This file follows janus's algorithm step by
 * step, single-threaded by default as janus is (no FFTW threads, GIL held):
 *
 *   js_green_apply      DiscreteGreenOperator2D/3D.c_apply
 *                       (janus/green.pyx:182-216, 353-390): r2c of every
 *                       component into a spectrum scratch, then the loop over
 *                       the half spectrum calling c_set_frequency (truncated
 *                       green.pyx:432-445, 505-523; filtered :607-694, 730-910)
 *                       and c_apply_by_freq on the real and imaginary parts,
 *                       then c2r of every component scaled by 1/N
 *                       (fft/serial/_serial_fft.pyx:152-177, 296-325);
 *   gamma0_2d/3d        continuous Gamma0(k) of
 *                       material/elastic/linear/isotropic.pyx:145-156,
 *                       173-200, 242-308 (0 at k = 0), in the operation order
 *                       of the numpy oracle (oracle/janus_oracle.py gamma0);
 *   iso4_apply          FourthRankIsotropicTensor2D/3D.c_apply
 *                       (operators.pyx:292-299, 315-324, 352-365) per voxel
 *                       of a phase map (BlockDiagonalOperator, :655-664,
 *                       706-716);
 *   js_basic_iterations the loop of sphinx/tutorials/square_basic/
 *                       square_basic.py:83-92 with its residual
 *                       ||eps_new - eps|| / (sqrt(N) ||E||).
 *
 * The FFT arithmetic lives in FFTW in janus (third-party, version not pinned;
 * the docs mention 3.3.4, sphinx/installation.rst:72).  Here it is a batched
 * Stockham transform of our own (radices 4, 2, 3, 5 and a direct DFT for
 * other primes), vectorised over the batch of sequences.  Real transforms
 * pack two real rows as one complex sequence.
 *
 * Parity: pinned against the numpy oracle (itself pinned to janus's 18 Green
 * fixtures and the published C_1212) and against the same fixtures, in
 * tests/test_oracle_c.py.
 *
 * nthreads > 1 spreads the same loops over OpenMP threads (the all-cores
 * baseline); the arithmetic of every value is unchanged.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double re, im;
} cplx;

/* ---------------------------------------------------------------- FFT plan */
#define JS_MAX_STAGES 40

typedef struct {
  int n;
  int nst;
  int radix[JS_MAX_STAGES];
  int ns[JS_MAX_STAGES];   /* span Ns before the stage */
  cplx *tw[JS_MAX_STAGES]; /* tw[s][(r - 1) * Ns + k] = exp(-2 pi i r k / (Ns R)) */
} fft_plan;

/* Returns 0, or -4 for a prime factor above 61 (the direct DFT's limit). */
static int plan_init(fft_plan *p, int n) {
  memset(p, 0, sizeof(*p));
  p->n = n;
  int m = n, ns = 1;
  while (m > 1) {
    int r;
    if (m % 4 == 0) r = 4;
    else if (m % 2 == 0) r = 2;
    else if (m % 3 == 0) r = 3;
    else if (m % 5 == 0) r = 5;
    else {
      r = 7;
      while (m % r) r += 2;
    }
    if (r > 61 || p->nst >= JS_MAX_STAGES) return -4;
    int s = p->nst++;
    p->radix[s] = r;
    p->ns[s] = ns;
    p->tw[s] = (cplx *)malloc(sizeof(cplx) * (size_t)(r > 1 ? r - 1 : 1) * ns);
    for (int q = 1; q < r; ++q)
      for (int k = 0; k < ns; ++k) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L *
                        (long double)(q * k) / (long double)(ns * r);
        p->tw[s][(q - 1) * ns + k].re = (double)cosl(a);
        p->tw[s][(q - 1) * ns + k].im = (double)sinl(a);
      }
    ns *= r;
    m /= r;
  }
  return 0;
}

static void plan_free(fft_plan *p) {
  for (int s = 0; s < p->nst; ++s) free(p->tw[s]);
}

/* One Stockham stage on a batch: sequences of n complex values, element i of
 * sequence b at x[i * B + b].  Inputs x[(j + q n / R) B + b], twiddle
 * w^(q (j mod Ns)) (conjugated when sgn > 0), outputs
 * y[((j / Ns) Ns R + j mod Ns + q Ns) B + b]. */
static void stage(const cplx *x, cplx *y, int n, int B, int R, int Ns,
                  const cplx *tw, int sgn) {
  const int m = n / R;
  cplx v[64], t[64];
  for (int j = 0; j < m; ++j) {
    const int k = j % Ns;
    const int o = (j / Ns) * Ns * R + k;
    cplx w[64];
    for (int q = 1; q < R; ++q) {
      w[q] = tw[(q - 1) * Ns + k];
      if (sgn > 0) w[q].im = -w[q].im;
    }
    if (R == 4) {
      const cplx *x0 = x + (size_t)j * B, *x1 = x + (size_t)(j + m) * B;
      const cplx *x2 = x + (size_t)(j + 2 * m) * B, *x3 = x + (size_t)(j + 3 * m) * B;
      cplx *y0 = y + (size_t)o * B, *y1 = y + (size_t)(o + Ns) * B;
      cplx *y2 = y + (size_t)(o + 2 * Ns) * B, *y3 = y + (size_t)(o + 3 * Ns) * B;
      const double s = (double)sgn;
      for (int b = 0; b < B; ++b) {
        cplx a0 = x0[b], a1, a2, a3;
        a1.re = x1[b].re * w[1].re - x1[b].im * w[1].im;
        a1.im = x1[b].re * w[1].im + x1[b].im * w[1].re;
        a2.re = x2[b].re * w[2].re - x2[b].im * w[2].im;
        a2.im = x2[b].re * w[2].im + x2[b].im * w[2].re;
        a3.re = x3[b].re * w[3].re - x3[b].im * w[3].im;
        a3.im = x3[b].re * w[3].im + x3[b].im * w[3].re;
        double s02r = a0.re + a2.re, s02i = a0.im + a2.im;
        double d02r = a0.re - a2.re, d02i = a0.im - a2.im;
        double s13r = a1.re + a3.re, s13i = a1.im + a3.im;
        double d13r = a1.re - a3.re, d13i = a1.im - a3.im;
        /* X1 = d02 + sgn i d13 (sgn -1: forward, -i) */
        y0[b].re = s02r + s13r;
        y0[b].im = s02i + s13i;
        y2[b].re = s02r - s13r;
        y2[b].im = s02i - s13i;
        y1[b].re = d02r - s * d13i;
        y1[b].im = d02i + s * d13r;
        y3[b].re = d02r + s * d13i;
        y3[b].im = d02i - s * d13r;
      }
    } else if (R == 2) {
      const cplx *x0 = x + (size_t)j * B, *x1 = x + (size_t)(j + m) * B;
      cplx *y0 = y + (size_t)o * B, *y1 = y + (size_t)(o + Ns) * B;
      for (int b = 0; b < B; ++b) {
        cplx a1;
        a1.re = x1[b].re * w[1].re - x1[b].im * w[1].im;
        a1.im = x1[b].re * w[1].im + x1[b].im * w[1].re;
        y0[b].re = x0[b].re + a1.re;
        y0[b].im = x0[b].im + a1.im;
        y1[b].re = x0[b].re - a1.re;
        y1[b].im = x0[b].im - a1.im;
      }
    } else {
      /* direct R-point DFT, roots exp(sgn 2 pi i q r / R) */
      cplx root[64];
      for (int q = 0; q < R; ++q) {
        long double a = (long double)sgn * 2.0L *
                        3.14159265358979323846264338327950288L * q / R;
        root[q].re = (double)cosl(a);
        root[q].im = (double)sinl(a);
      }
      for (int b = 0; b < B; ++b) {
        for (int q = 0; q < R; ++q) {
          cplx a = x[(size_t)(j + q * m) * B + b];
          if (q) {
            cplx u;
            u.re = a.re * w[q].re - a.im * w[q].im;
            u.im = a.re * w[q].im + a.im * w[q].re;
            a = u;
          }
          v[q] = a;
        }
        for (int r = 0; r < R; ++r) {
          double sr = 0., si = 0.;
          for (int q = 0; q < R; ++q) {
            cplx e = root[(q * r) % R];
            sr += v[q].re * e.re - v[q].im * e.im;
            si += v[q].re * e.im + v[q].im * e.re;
          }
          t[r].re = sr;
          t[r].im = si;
        }
        for (int r = 0; r < R; ++r) y[(size_t)(o + r * Ns) * B + b] = t[r];
      }
    }
  }
}

/* Transform of B sequences in a (n, B) array a (in place; w: scratch of the
 * same size).  sgn = -1 forward, +1 inverse (unscaled). */
static void fft_batch(const fft_plan *p, cplx *a, cplx *w, int B, int sgn) {
  cplx *src = a, *dst = w;
  for (int s = 0; s < p->nst; ++s) {
    stage(src, dst, p->n, B, p->radix[s], p->ns[s], p->tw[s], sgn);
    cplx *t = src;
    src = dst;
    dst = t;
  }
  if (src != a) memcpy(a, src, sizeof(cplx) * (size_t)p->n * B);
}

/* ---------------------------------------------------------------- real FFT
 * Grid (n0, n1[, n2]) in 3D, or (n0, n1) in 2D with n0 playing no role:
 * treated as 3D with a leading axis of 1.  Spectrum of one component:
 * (n0, n1, nh) complex, nh = n2 / 2 + 1. */
typedef struct {
  int n0, n1, n2, nh;
  fft_plan p0, p1, p2;
} grid_fft;

static int grid_init(grid_fft *g, int dim, const int64_t *shape) {
  g->n0 = dim == 3 ? (int)shape[0] : 1;
  g->n1 = (int)shape[dim - 2];
  g->n2 = (int)shape[dim - 1];
  g->nh = g->n2 / 2 + 1;
  int rc = plan_init(&g->p0, g->n0);
  if (!rc) rc = plan_init(&g->p1, g->n1);
  if (!rc) rc = plan_init(&g->p2, g->n2);
  return rc;
}

static void grid_free(grid_fft *g) {
  plan_free(&g->p0);
  plan_free(&g->p1);
  plan_free(&g->p2);
}

static int max3(int a, int b, int c) {
  int m = a > b ? a : b;
  return m > c ? m : c;
}

/* r2c of component c of the real field x (voxel stride sym) into spec
 * (n0, n1, nh).  Rows r and r + 1 of a plane form one complex sequence. */
static void r2c(const grid_fft *g, const double *x, int sym, int c, cplx *spec,
                int nthreads) {
  const int n0 = g->n0, n1 = g->n1, n2 = g->n2, nh = g->nh;
  const int np = (n1 + 1) / 2;
  (void)nthreads;
#pragma omp parallel num_threads(nthreads)
  {
    const int big = max3(n2 * np, n1 * nh, 1);
    cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)big);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)big);
#pragma omp for schedule(static)
    for (int i0 = 0; i0 < n0; ++i0) {
      const double *xp = x + (size_t)i0 * n1 * n2 * sym;
      /* pack: a[j][pair] = x[2 pair][j] + i x[2 pair + 1][j] */
      for (int j = 0; j < n2; ++j)
        for (int q = 0; q < np; ++q) {
          int r0 = 2 * q, r1 = 2 * q + 1;
          a[(size_t)j * np + q].re = xp[((size_t)r0 * n2 + j) * sym + c];
          a[(size_t)j * np + q].im =
              r1 < n1 ? xp[((size_t)r1 * n2 + j) * sym + c] : 0.;
        }
      fft_batch(&g->p2, a, w, np, -1);
      /* separate: A[k] = (Z[k] + conj Z[n-k]) / 2, B[k] = (Z[k] - conj
       * Z[n-k]) / 2i, into the plane (n1, nh) */
      cplx *pl = spec + (size_t)i0 * n1 * nh;
      for (int q = 0; q < np; ++q)
        for (int k = 0; k < nh; ++k) {
          cplx zk = a[(size_t)k * np + q];
          cplx zm = a[(size_t)((n2 - k) % n2) * np + q];
          cplx A, Bv;
          A.re = 0.5 * (zk.re + zm.re);
          A.im = 0.5 * (zk.im - zm.im);
          Bv.re = 0.5 * (zk.im + zm.im);
          Bv.im = -0.5 * (zk.re - zm.re);
          pl[(size_t)(2 * q) * nh + k] = A;
          if (2 * q + 1 < n1) pl[(size_t)(2 * q + 1) * nh + k] = Bv;
        }
      /* axis 1: the plane is (n1, nh) = sequences along axis 1, batch nh */
      if (n1 > 1) fft_batch(&g->p1, pl, w, nh, -1);
    }
    free(a);
    free(w);
  }
  if (n0 > 1) {
#pragma omp parallel num_threads(nthreads)
    {
      cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)n0 * nh);
      cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n0 * nh);
#pragma omp for schedule(static)
      for (int i1 = 0; i1 < n1; ++i1) {
        for (int i0 = 0; i0 < n0; ++i0)
          memcpy(a + (size_t)i0 * nh, spec + ((size_t)i0 * n1 + i1) * nh,
                 sizeof(cplx) * nh);
        fft_batch(&g->p0, a, w, nh, -1);
        for (int i0 = 0; i0 < n0; ++i0)
          memcpy(spec + ((size_t)i0 * n1 + i1) * nh, a + (size_t)i0 * nh,
                 sizeof(cplx) * nh);
      }
      free(a);
      free(w);
    }
  }
}

/* c2r of spec (n0, n1, nh; overwritten) into component c of y, scaled by
 * scale.  The imaginary parts of the last axis's DC and Nyquist entries are
 * dropped, as FFTW and pocketfft do. */
static void c2r(const grid_fft *g, cplx *spec, double *y, int sym, int c,
                double scale, int nthreads) {
  const int n0 = g->n0, n1 = g->n1, n2 = g->n2, nh = g->nh;
  const int np = (n1 + 1) / 2;
  (void)nthreads;
  if (n0 > 1) {
#pragma omp parallel num_threads(nthreads)
    {
      cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)n0 * nh);
      cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n0 * nh);
#pragma omp for schedule(static)
      for (int i1 = 0; i1 < n1; ++i1) {
        for (int i0 = 0; i0 < n0; ++i0)
          memcpy(a + (size_t)i0 * nh, spec + ((size_t)i0 * n1 + i1) * nh,
                 sizeof(cplx) * nh);
        fft_batch(&g->p0, a, w, nh, +1);
        for (int i0 = 0; i0 < n0; ++i0)
          memcpy(spec + ((size_t)i0 * n1 + i1) * nh, a + (size_t)i0 * nh,
                 sizeof(cplx) * nh);
      }
      free(a);
      free(w);
    }
  }
#pragma omp parallel num_threads(nthreads)
  {
    const int big = max3(n2 * np, n1 * nh, 1);
    cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)big);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)big);
#pragma omp for schedule(static)
    for (int i0 = 0; i0 < n0; ++i0) {
      cplx *pl = spec + (size_t)i0 * n1 * nh;
      if (n1 > 1) fft_batch(&g->p1, pl, w, nh, +1);
      /* Z[k] = A[k] + i B[k] over the full length from the half spectra */
      for (int q = 0; q < np; ++q) {
        const cplx *A = pl + (size_t)(2 * q) * nh;
        const cplx *Bv = 2 * q + 1 < n1 ? pl + (size_t)(2 * q + 1) * nh : NULL;
        for (int k = 0; k < nh; ++k) {
          cplx av = A[k], bv = Bv ? Bv[k] : (cplx){0., 0.};
          if (k == 0 || 2 * k == n2) {
            a[(size_t)k * np + q].re = av.re;
            a[(size_t)k * np + q].im = bv.re;
          } else {
            a[(size_t)k * np + q].re = av.re - bv.im;
            a[(size_t)k * np + q].im = av.im + bv.re;
            a[(size_t)(n2 - k) * np + q].re = av.re + bv.im;
            a[(size_t)(n2 - k) * np + q].im = bv.re - av.im;
          }
        }
      }
      fft_batch(&g->p2, a, w, np, +1);
      double *yp = y + (size_t)i0 * n1 * n2 * sym;
      for (int j = 0; j < n2; ++j)
        for (int q = 0; q < np; ++q) {
          yp[((size_t)(2 * q) * n2 + j) * sym + c] = scale * a[(size_t)j * np + q].re;
          if (2 * q + 1 < n1)
            yp[((size_t)(2 * q + 1) * n2 + j) * sym + c] =
                scale * a[(size_t)j * np + q].im;
        }
    }
    free(a);
    free(w);
  }
}

/* ------------------------------------------------------ Gamma0 and Gamma-hat */
static const double SQRT2 = 1.41421356237309504880;

/* 2D: g[3][3] (isotropic.pyx:173-200). */
static void gamma0_2d(double k0, double k1, const double *d, double *g) {
  double k0k0 = k0 * k0, k1k1 = k1 * k1;
  double kk = k0k0 + k1k1;
  if (kk == 0.) {
    for (int i = 0; i < 9; ++i) g[i] = 0.;
    return;
  }
  double s = 1.0 / kk;
  k0k0 = k0k0 * s;
  k1k1 = k1k1 * s;
  double k0k1 = s * k0 * k1;
  double g00 = k0k0 * (d[0] - d[1] * k0k0);
  double g11 = k1k1 * (d[0] - d[1] * k1k1);
  double dummy = -d[1] * k0k0 * k1k1;
  double g22 = 2 * (d[2] + dummy);
  double g02 = SQRT2 * k0k1 * (d[3] - d[1] * k0k0);
  double g12 = SQRT2 * k0k1 * (d[3] - d[1] * k1k1);
  g[0] = g00; g[1] = dummy; g[2] = g02;
  g[3] = dummy; g[4] = g11; g[5] = g12;
  g[6] = g02; g[7] = g12; g[8] = g22;
}

/* 3D: g[6][6] (isotropic.pyx:242-308); d = daux1..4. */
static void gamma0_3d(double k0, double k1, double k2, const double *d,
                      double *g) {
  double k0k0 = k0 * k0, k1k1 = k1 * k1, k2k2 = k2 * k2;
  double kk = k0k0 + k1k1 + k2k2;
  if (kk == 0.) {
    for (int i = 0; i < 36; ++i) g[i] = 0.;
    return;
  }
  double s = 1.0 / kk;
  k0k0 = k0k0 * s;
  k1k1 = k1k1 * s;
  k2k2 = k2k2 * s;
  double k1k2 = s * k1 * k2, k2k0 = s * k2 * k0, k0k1 = s * k0 * k1;
  const double d1 = d[0], d2 = d[1], d3 = d[2], d4 = d[3];
  double c[6][6];
  c[0][0] = k0k0 * (d1 - d2 * k0k0);
  c[1][1] = k1k1 * (d1 - d2 * k1k1);
  c[2][2] = k2k2 * (d1 - d2 * k2k2);
  c[3][3] = 2. * (d3 * (k1k1 + k2k2) - d2 * k1k1 * k2k2);
  c[4][4] = 2. * (d3 * (k2k2 + k0k0) - d2 * k2k2 * k0k0);
  c[5][5] = 2. * (d3 * (k0k0 + k1k1) - d2 * k0k0 * k1k1);
  c[0][1] = -d2 * k0k0 * k1k1;
  c[0][2] = -d2 * k0k0 * k2k2;
  c[0][3] = -SQRT2 * d2 * k0k0 * k1k2;
  c[0][4] = SQRT2 * k2k0 * (d4 - d2 * k0k0);
  c[0][5] = SQRT2 * k0k1 * (d4 - d2 * k0k0);
  c[1][2] = -d2 * k1k1 * k2k2;
  c[1][3] = SQRT2 * k1k2 * (d4 - d2 * k1k1);
  c[1][4] = -SQRT2 * d2 * k1k1 * k2k0;
  c[1][5] = SQRT2 * k0k1 * (d4 - d2 * k1k1);
  c[2][3] = SQRT2 * k1k2 * (d4 - d2 * k2k2);
  c[2][4] = SQRT2 * k2k0 * (d4 - d2 * k2k2);
  c[2][5] = -SQRT2 * d2 * k2k2 * k0k1;
  c[3][4] = 2 * k0k1 * (d3 - d2 * k2k2);
  c[3][5] = 2 * k2k0 * (d3 - d2 * k1k1);
  c[4][5] = 2 * k1k2 * (d3 - d2 * k0k0);
  for (int i = 0; i < 6; ++i) {
    g[i * 6 + i] = c[i][i];
    for (int j = i + 1; j < 6; ++j) g[i * 6 + j] = g[j * 6 + i] = c[i][j];
  }
}

/* Per-axis tables: truncated k (green.pyx:432-445); filtered aliases
 * k- = s (b - n), k+ = s b with weights cos^2(h k / 4) (:740-800). */
typedef struct {
  double *kt, *km, *kp, *wm, *wp;
} axis_tab;

static void axis_init(axis_tab *t, int n, double h) {
  t->kt = (double *)malloc(sizeof(double) * 5 * (size_t)n);
  t->km = t->kt + n;
  t->kp = t->km + n;
  t->wm = t->kp + n;
  t->wp = t->wm + n;
  const double s = 2. * M_PI / (h * (double)n); /* green.pyx:140-141, 305-307 */
  for (int b = 0; b < n; ++b) {
    t->kt[b] = (2 * b > n) ? s * (double)(b - n) : s * (double)b;
    t->km[b] = s * (double)(b - n);
    t->kp[b] = s * (double)b;
    double w = cos(0.25 * h * t->km[b]);
    t->wm[b] = w * w;
    w = cos(0.25 * h * t->kp[b]);
    t->wp[b] = w * w;
  }
}

/* Gamma-hat(b) (sym x sym, row major) of the truncated (kind 0) or filtered
 * (kind 1) operator.  Filtered: the sum over the 2^dim alias combinations in
 * janus's order (the first axis varies slowest, the minus alias first),
 * each term w Gamma0(k) with w the product of the per-axis weights in axis
 * order, accumulated term after term. */
static void gamma_hat(int dim, int kind, const axis_tab *ax, const int *b,
                      const double *d, double *g) {
  const int sym = dim == 2 ? 3 : 6, nc = sym * sym;
  if (kind == 0) {
    if (dim == 2) gamma0_2d(ax[0].kt[b[0]], ax[1].kt[b[1]], d, g);
    else gamma0_3d(ax[0].kt[b[0]], ax[1].kt[b[1]], ax[2].kt[b[2]], d, g);
    return;
  }
  double t[36];
  for (int combo = 0; combo < (1 << dim); ++combo) {
    double k[3] = {0., 0., 0.}, w = 1.;
    for (int i = 0; i < dim; ++i) {
      int plus = (combo >> (dim - 1 - i)) & 1;
      k[i] = plus ? ax[i].kp[b[i]] : ax[i].km[b[i]];
      double wi = plus ? ax[i].wp[b[i]] : ax[i].wm[b[i]];
      w = i == 0 ? wi : w * wi;
    }
    if (dim == 2) gamma0_2d(k[0], k[1], d, t);
    else gamma0_3d(k[0], k[1], k[2], d, t);
    if (combo == 0)
      for (int i = 0; i < nc; ++i) g[i] = w * t[i];
    else
      for (int i = 0; i < nc; ++i) g[i] = g[i] + w * t[i];
  }
}

/* ---------------------------------------------------------------- apply */
typedef struct {
  int dim, sym, kind;
  int64_t n[3];
  int64_t N;
  double h, d[4];
  grid_fft gf;
  axis_tab ax[3];
  cplx *spec; /* (n0, n1, nh, sym) */
  cplx *work; /* (n0, n1, nh) */
} green_t;

static int green_init(green_t *G, int dim, const int64_t *shape, int kind,
                      double g0, double nu0, double h) {
  memset(G, 0, sizeof(*G));
  if (dim != 2 && dim != 3) return -1;
  if (kind != 0 && kind != 1) return -1;
  G->dim = dim;
  G->sym = dim == 2 ? 3 : 6;
  G->kind = kind;
  G->N = 1;
  for (int i = 0; i < dim; ++i) {
    if (shape[i] <= 0) return -1;
    G->n[i] = shape[i];
    G->N *= shape[i];
  }
  G->h = h;
  /* isotropic.pyx:145-156 */
  G->d[0] = 1.0 / g0;
  G->d[1] = 0.5 / (g0 * (1.0 - nu0));
  G->d[2] = 0.25 / g0;
  G->d[3] = 0.5 / g0;
  int rc = grid_init(&G->gf, dim, shape);
  for (int i = 0; i < dim; ++i) axis_init(&G->ax[i], (int)shape[i], h);
  if (rc) return rc;
  size_t nspec = (size_t)G->gf.n0 * G->gf.n1 * G->gf.nh;
  G->work = (cplx *)malloc(sizeof(cplx) * nspec);
  G->spec = (cplx *)malloc(sizeof(cplx) * nspec * G->sym);
  return (G->work && G->spec) ? 0 : -3;
}

static void green_free(green_t *G) {
  grid_free(&G->gf);
  for (int i = 0; i < G->dim; ++i) free(G->ax[i].kt);
  free(G->work);
  free(G->spec);
}

static void green_apply(green_t *G, const double *x, double *y, int nthreads) {
  const int sym = G->sym, dim = G->dim;
  const int n0 = G->gf.n0, n1 = G->gf.n1, nh = G->gf.nh;
  const size_t np = (size_t)n0 * n1 * nh;
  /* r2c of every component (green.pyx:357-360) */
  for (int c = 0; c < sym; ++c) {
    r2c(&G->gf, x, sym, c, G->work, nthreads);
    for (size_t i = 0; i < np; ++i) G->spec[i * sym + c] = G->work[i];
  }
  /* frequency loop (green.pyx:362-386): Gamma-hat(b) applied to the real
   * and imaginary parts */
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int i0 = 0; i0 < n0; ++i0) {
    double g[36], xr[6], xi[6];
    int b[3];
    for (int i1 = 0; i1 < n1; ++i1)
      for (int j = 0; j < nh; ++j) {
        if (dim == 3) {
          b[0] = i0;
          b[1] = i1;
          b[2] = j;
        } else {
          b[0] = i1;
          b[1] = j;
        }
        gamma_hat(dim, G->kind, G->ax, b, G->d, g);
        cplx *z = G->spec + (((size_t)i0 * n1 + i1) * nh + j) * sym;
        for (int c = 0; c < sym; ++c) {
          xr[c] = z[c].re;
          xi[c] = z[c].im;
        }
        for (int r = 0; r < sym; ++r) {
          double ar = 0., ai = 0.;
          for (int c = 0; c < sym; ++c) {
            ar += g[r * sym + c] * xr[c];
            ai += g[r * sym + c] * xi[c];
          }
          z[r].re = ar;
          z[r].im = ai;
        }
      }
  }
  /* c2r of every component, scaled by 1/N (green.pyx:387-390) */
  const double scale = 1.0 / (double)G->N;
  for (int c = 0; c < sym; ++c) {
    for (size_t i = 0; i < np; ++i) G->work[i] = G->spec[i * sym + c];
    c2r(&G->gf, G->work, y, sym, c, scale, nthreads);
  }
}

/* y = Gamma0[x]; x, y: (n..., sym) C order, may be the same array.  kind 0
 * truncated, 1 filtered.  Returns 0, or < 0 for invalid arguments / out of
 * memory. */
int js_green_apply(int dim, const int64_t *shape, int kind, double g0,
                   double nu0, double h, const double *x, double *y,
                   int nthreads) {
  green_t G;
  int rc = green_init(&G, dim, shape, kind, g0, nu0, h);
  if (rc == 0) green_apply(&G, x, y, nthreads < 1 ? 1 : nthreads);
  green_free(&G);
  return rc;
}

/* (sph J + dev K) : x per voxel, with (sph, dev) = table[phase[v]]
 * (operators.pyx:292-299 tr = (sph - dev) / dim, 315-324, 352-365). */
static void iso4_phases(int dim, int64_t N, const uint8_t *phase,
                        const double *table, const double *x, double *y,
                        int nthreads) {
  const int sym = dim == 2 ? 3 : 6;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t v = 0; v < N; ++v) {
    const double sph = table[2 * phase[v]], dev = table[2 * phase[v] + 1];
    const double tr = (sph - dev) / (double)dim;
    const double *xv = x + v * sym;
    double *yv = y + v * sym;
    double aux = dim == 2 ? tr * (xv[0] + xv[1]) : tr * (xv[0] + xv[1] + xv[2]);
    for (int i = 0; i < dim; ++i) yv[i] = aux + dev * xv[i];
    for (int i = dim; i < sym; ++i) yv[i] = dev * xv[i];
  }
}

/* iters iterations of the basic scheme (square_basic.py:83-92) from eps
 * (input: the initial iterate; output: the last one): tau = dC : eps,
 * eps_new = E - Gamma0[tau], res[k] = ||eps_new - eps|| / (sqrt(N) ||E||).
 * phase: uint8 map of the grid; table: nphase (sph, dev) pairs of
 * dC = C - C0.  Returns 0, or < 0 for invalid arguments / out of memory. */
int js_basic_iterations(int dim, const int64_t *shape, int kind, double g0,
                        double nu0, double h, const uint8_t *phase,
                        const double *table, const double *E, double *eps,
                        int iters, double *res, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  green_t G;
  int rc = green_init(&G, dim, shape, kind, g0, nu0, h);
  if (rc) {
    green_free(&G);
    return rc;
  }
  const int sym = G.sym;
  const int64_t N = G.N;
  double *tau = (double *)malloc(sizeof(double) * (size_t)N * sym);
  if (!tau) {
    green_free(&G);
    return -3;
  }
  double nE = 0.;
  for (int c = 0; c < sym; ++c) nE += E[c] * E[c];
  nE = sqrt(nE);
  const double norm = 1. / sqrt((double)N) / (nE > 0 ? nE : 1.);
  for (int it = 0; it < iters; ++it) {
    iso4_phases(dim, N, phase, table, eps, tau, nthreads);
    green_apply(&G, tau, tau, nthreads);
    double acc = 0.;
#pragma omp parallel for num_threads(nthreads) reduction(+ : acc) schedule(static)
    for (int64_t v = 0; v < N; ++v)
      for (int c = 0; c < sym; ++c) {
        double nv = E[c] - tau[v * sym + c];
        double dd = nv - eps[v * sym + c];
        acc += dd * dd;
        eps[v * sym + c] = nv;
      }
    if (res) res[it] = norm * sqrt(acc);
  }
  free(tau);
  green_free(&G);
  return 0;
}
