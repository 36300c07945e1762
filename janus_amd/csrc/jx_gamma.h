// Green operator for strains of an isotropic reference material, Mandel
// notation, computed from the wave vector (never stored).  Shared by the
// device kernels and the host-side per-frequency API.
//
// Gamma0(k) follows janus/material/elastic/linear/isotropic.pyx:133-156
// (daux1 = 1/g, daux2 = 1/(2 g (1 - nu)), daux3 = 1/(4 g), daux4 = 1/(2 g)),
// :173-200 (2D) and :242-308 (3D), keeping janus's operation order; it is
// exactly zero at k = 0.  The coefficients are stored as the upper triangle,
// row major: 2D g00 g01 g02 g11 g12 g22, 3D g00 .. g05 g11 .. g15 g22 .. g55.
#pragma once

#include "jx_internal.h"

namespace jx {

#define JX_HD __host__ __device__ __forceinline__

constexpr double kSqrt2 = 1.41421356237309504880;

JX_HD void gamma0_2d(double k0, double k1, double daux1, double daux2,
                     double daux3, double daux4, double g[6]) {
  double k0k0 = k0 * k0;
  double k1k1 = k1 * k1;
  if (k0k0 + k1k1 == 0.) {
    for (int i = 0; i < 6; ++i) g[i] = 0.;
    return;
  }
  double s = 1.0 / (k0k0 + k1k1);
  k0k0 *= s;
  k1k1 *= s;
  double k0k1 = s * k0 * k1;
  g[0] = k0k0 * (daux1 - daux2 * k0k0);                      // g00
  g[3] = k1k1 * (daux1 - daux2 * k1k1);                      // g11
  double dummy = -daux2 * k0k0 * k1k1;
  g[1] = dummy;                                              // g01
  g[5] = 2 * (daux3 + dummy);                                // g22
  g[2] = kSqrt2 * k0k1 * (daux4 - daux2 * k0k0);             // g02
  g[4] = kSqrt2 * k0k1 * (daux4 - daux2 * k1k1);             // g12
}

JX_HD void gamma0_3d(double k0, double k1, double k2, double daux1,
                     double daux2, double daux3, double daux4, double g[21]) {
  double k0k0 = k0 * k0;
  double k1k1 = k1 * k1;
  double k2k2 = k2 * k2;
  if (k0k0 + k1k1 + k2k2 == 0.) {
    for (int i = 0; i < 21; ++i) g[i] = 0.;
    return;
  }
  double s = 1.0 / (k0k0 + k1k1 + k2k2);
  k0k0 *= s;
  k1k1 *= s;
  k2k2 *= s;
  double k1k2 = s * k1 * k2;
  double k2k0 = s * k2 * k0;
  double k0k1 = s * k0 * k1;
  // index map of the upper triangle
  // 0:g00 1:g01 2:g02 3:g03 4:g04 5:g05 6:g11 7:g12 8:g13 9:g14 10:g15
  // 11:g22 12:g23 13:g24 14:g25 15:g33 16:g34 17:g35 18:g44 19:g45 20:g55
  g[0] = k0k0 * (daux1 - daux2 * k0k0);
  g[6] = k1k1 * (daux1 - daux2 * k1k1);
  g[11] = k2k2 * (daux1 - daux2 * k2k2);
  g[15] = 2. * (daux3 * (k1k1 + k2k2) - daux2 * k1k1 * k2k2);
  g[18] = 2. * (daux3 * (k2k2 + k0k0) - daux2 * k2k2 * k0k0);
  g[20] = 2. * (daux3 * (k0k0 + k1k1) - daux2 * k0k0 * k1k1);
  g[1] = -daux2 * k0k0 * k1k1;
  g[2] = -daux2 * k0k0 * k2k2;
  g[3] = -kSqrt2 * daux2 * k0k0 * k1k2;
  g[4] = kSqrt2 * k2k0 * (daux4 - daux2 * k0k0);
  g[5] = kSqrt2 * k0k1 * (daux4 - daux2 * k0k0);
  g[7] = -daux2 * k1k1 * k2k2;
  g[8] = kSqrt2 * k1k2 * (daux4 - daux2 * k1k1);
  g[9] = -kSqrt2 * daux2 * k1k1 * k2k0;
  g[10] = kSqrt2 * k0k1 * (daux4 - daux2 * k1k1);
  g[12] = kSqrt2 * k1k2 * (daux4 - daux2 * k2k2);
  g[13] = kSqrt2 * k2k0 * (daux4 - daux2 * k2k2);
  g[14] = -kSqrt2 * daux2 * k2k2 * k0k1;
  g[16] = 2 * k0k1 * (daux3 - daux2 * k2k2);
  g[17] = 2 * k2k0 * (daux3 - daux2 * k1k1);
  g[19] = 2 * k1k2 * (daux3 - daux2 * k0k0);
}

// Upper-triangle position of (i, j), i <= j, for sym = 3 or 6.
JX_HD int tri_index(int sym, int i, int j) {
  return i * sym - (i * (i - 1)) / 2 + (j - i);
}

// Discrete Gamma-hat at integer frequency b of the grid, from the per-axis
// tables.  Truncated: Gamma0 of the aliased wave vector (green.pyx:432-451,
// 505-529).  Filtered: sum over the 2^dim alias combinations of
// prod_i cos^2(h k_i / 4) Gamma0(k), accumulated in janus's order -- axis 0
// slowest, the "minus" alias first (green.pyx:607-694, 730-910).
// Willot's modified wave vector k_i = sin(phi_i) prod_{j != i} cos(phi_j)
// (green.pyx:984-1023), from the per-axis sin/cos tables.
template <int DIM>
JX_HD void willot_k(const GreenDev &G, const int *b, double *k) {
  if (DIM == 2) {
    double s0 = G.ax[0].sw[b[0]], c0 = G.ax[0].cw[b[0]];
    double s1 = G.ax[1].sw[b[1]], c1 = G.ax[1].cw[b[1]];
    k[0] = s0 * c1;
    k[1] = c0 * s1;
  } else {
    double s0 = G.ax[0].sw[b[0]], c0 = G.ax[0].cw[b[0]];
    double s1 = G.ax[1].sw[b[1]], c1 = G.ax[1].cw[b[1]];
    double s2 = G.ax[2].sw[b[2]], c2 = G.ax[2].cw[b[2]];
    k[0] = s0 * c1 * c2;
    k[1] = c0 * s1 * c2;
    k[DIM - 1] = c0 * c1 * s2;
  }
}

template <int DIM>
JX_HD void gamma_hat(const GreenDev &G, const int *b, double *g) {
  constexpr int NC = DIM == 2 ? 6 : 21;
  if (G.kind == JX_GREEN_WILLOT) {
    double k[DIM];
    willot_k<DIM>(G, b, k);
    if (DIM == 2)
      gamma0_2d(k[0], k[1], G.daux1, G.daux2, G.daux3, G.daux4, g);
    else
      gamma0_3d(k[0], k[1], k[DIM - 1], G.daux1, G.daux2, G.daux3, G.daux4, g);
    return;
  }
  if (G.kind == JX_GREEN_TRUNCATED) {
    if (DIM == 2)
      gamma0_2d(G.ax[0].kt[b[0]], G.ax[1].kt[b[1]], G.daux1, G.daux2, G.daux3,
                G.daux4, g);
    else
      gamma0_3d(G.ax[0].kt[b[0]], G.ax[1].kt[b[1]], G.ax[2].kt[b[2]], G.daux1,
                G.daux2, G.daux3, G.daux4, g);
    return;
  }
  double km[DIM], kp[DIM], wm[DIM], wp[DIM];
  for (int i = 0; i < DIM; ++i) {
    km[i] = G.ax[i].km[b[i]];
    kp[i] = G.ax[i].kp[b[i]];
    wm[i] = G.ax[i].wm[b[i]];
    wp[i] = G.ax[i].wp[b[i]];
  }
  for (int i = 0; i < NC; ++i) g[i] = 0.;
  double t[NC];
  for (int c = 0; c < (1 << DIM); ++c) {
    double k[DIM];
    double w = 1.0;
    for (int i = 0; i < DIM; ++i) {
      int plus = (c >> (DIM - 1 - i)) & 1;
      k[i] = plus ? kp[i] : km[i];
      double wi = plus ? wp[i] : wm[i];
      w = (i == 0) ? wi : w * wi;
    }
    if (DIM == 2)
      gamma0_2d(k[0], k[1], G.daux1, G.daux2, G.daux3, G.daux4, t);
    else
      gamma0_3d(k[0], k[1], k[DIM - 1], G.daux1, G.daux2, G.daux3, G.daux4, t);
    for (int i = 0; i < NC; ++i) g[i] += w * t[i];
  }
}

// Device evaluation of Gamma-hat through the two moment tensors of the
// aliases, A_ij = sum_c w_c n_i n_j and B_ijkl = sum_c w_c n_i n_j n_k n_l
// (n = k / |k| of alias c, weight w_c; one alias of weight 1 for the
// truncated operator).  Gamma0_ijkl(n) = (d_ik n_j n_l + d_il n_j n_k +
// d_jk n_i n_l + d_jl n_i n_k) / (4 g) - n_i n_j n_k n_l / (2 g (1 - nu))
// (the closed form of janus's tests/test_greenop.py:21-35) is linear in
// n (x) n and n (x) n (x) n (x) n, so the weighted sum of janus's
// green.pyx:730-910 equals the Mandel assembly below, at about 40 percent
// of the operations of summing eight full Gamma0 matrices.  Aliases with
// k = 0 contribute 0 (isotropic.pyx:253-275).
__device__ __forceinline__ void moments_add2(const double *k, double w,
                                             double *A, double *B) {
  double kk = k[0] * k[0] + k[1] * k[1];
  if (kk == 0.) return;
  double s = 1.0 / kk;
  double u[3], t[3];
  u[0] = s * k[0] * k[0];
  u[1] = s * k[0] * k[1];
  u[2] = s * k[1] * k[1];
  A[0] += w * u[0];
  A[1] += w * u[1];
  A[2] += w * u[2];
  t[0] = w * u[0];
  t[1] = w * u[1];
  t[2] = w * u[2];
  B[0] += t[0] * u[0];
  B[1] += t[0] * u[1];
  B[2] += t[0] * u[2];
  B[3] += t[1] * u[2];
  B[4] += t[2] * u[2];
}

__device__ __forceinline__ void moments_add3(const double *k, double w,
                                             double *A, double *B) {
  double kk = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
  if (kk == 0.) return;
  double s = 1.0 / kk;
  double u[6], t[6];
  u[0] = s * k[0] * k[0];
  u[1] = s * k[0] * k[1];
  u[2] = s * k[0] * k[2];
  u[3] = s * k[1] * k[1];
  u[4] = s * k[1] * k[2];
  u[5] = s * k[2] * k[2];
  A[0] += w * u[0];
  A[1] += w * u[1];
  A[2] += w * u[2];
  A[3] += w * u[3];
  A[4] += w * u[4];
  A[5] += w * u[5];
  t[0] = w * u[0];
  t[1] = w * u[1];
  t[2] = w * u[2];
  t[3] = w * u[3];
  t[4] = w * u[4];
  t[5] = w * u[5];
  B[0] += t[0] * u[0];
  B[1] += t[0] * u[1];
  B[2] += t[0] * u[2];
  B[3] += t[0] * u[3];
  B[4] += t[0] * u[4];
  B[5] += t[0] * u[5];
  B[6] += t[1] * u[3];
  B[7] += t[1] * u[4];
  B[8] += t[1] * u[5];
  B[9] += t[2] * u[5];
  B[10] += t[3] * u[3];
  B[11] += t[3] * u[4];
  B[12] += t[3] * u[5];
  B[13] += t[4] * u[5];
  B[14] += t[5] * u[5];
}

template <int DIM>
__device__ __forceinline__ void gamma_hat_moments(const GreenDev &G,
                                                  const int *b, double *g) {
  constexpr int NA = DIM == 2 ? 3 : 6;
  constexpr int NB = DIM == 2 ? 5 : 15;
  double A[NA], B[NB];
#pragma unroll
  for (int i = 0; i < NA; ++i) A[i] = 0.;
#pragma unroll
  for (int i = 0; i < NB; ++i) B[i] = 0.;
  if (G.kind == JX_GREEN_TRUNCATED || G.kind == JX_GREEN_WILLOT) {
    double k[DIM];
    if (G.kind == JX_GREEN_WILLOT) {
      willot_k<DIM>(G, b, k);
    } else {
#pragma unroll
      for (int i = 0; i < DIM; ++i) k[i] = G.ax[i].kt[b[i]];
    }
    if constexpr (DIM == 2) moments_add2(k, 1.0, A, B);
    else moments_add3(k, 1.0, A, B);
  } else {
    double km[DIM], kp[DIM], wm[DIM], wp[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      km[i] = G.ax[i].km[b[i]];
      kp[i] = G.ax[i].kp[b[i]];
      wm[i] = G.ax[i].wm[b[i]];
      wp[i] = G.ax[i].wp[b[i]];
    }
#pragma unroll
    for (int c = 0; c < (1 << DIM); ++c) {
      double k[DIM];
      double w = 1.0;
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        int plus = (c >> (DIM - 1 - i)) & 1;
        k[i] = plus ? kp[i] : km[i];
        w = (i == 0) ? (plus ? wp[i] : wm[i]) : w * (plus ? wp[i] : wm[i]);
      }
      if constexpr (DIM == 2) moments_add2(k, w, A, B);
      else moments_add3(k, w, A, B);
    }
  }
  const double d2 = G.daux2, d3 = G.daux3;
  if constexpr (DIM == 2) {
    g[0] = (d3 * (A[0] + A[0] + A[0] + A[0]) - d2 * B[0]);
    g[1] = (-d2 * B[2]);
    g[2] = kSqrt2 * (d3 * (A[1] + A[1]) - d2 * B[1]);
    g[3] = (d3 * (A[2] + A[2] + A[2] + A[2]) - d2 * B[4]);
    g[4] = kSqrt2 * (d3 * (A[1] + A[1]) - d2 * B[3]);
    g[5] = 2. * (d3 * (A[2] + A[0]) - d2 * B[2]);
  } else {
    g[0] = (d3 * (A[0] + A[0] + A[0] + A[0]) - d2 * B[0]);
    g[1] = (-d2 * B[3]);
    g[2] = (-d2 * B[5]);
    g[3] = kSqrt2 * (-d2 * B[4]);
    g[4] = kSqrt2 * (d3 * (A[2] + A[2]) - d2 * B[2]);
    g[5] = kSqrt2 * (d3 * (A[1] + A[1]) - d2 * B[1]);
    g[6] = (d3 * (A[3] + A[3] + A[3] + A[3]) - d2 * B[10]);
    g[7] = (-d2 * B[12]);
    g[8] = kSqrt2 * (d3 * (A[4] + A[4]) - d2 * B[11]);
    g[9] = kSqrt2 * (-d2 * B[7]);
    g[10] = kSqrt2 * (d3 * (A[1] + A[1]) - d2 * B[6]);
    g[11] = (d3 * (A[5] + A[5] + A[5] + A[5]) - d2 * B[14]);
    g[12] = kSqrt2 * (d3 * (A[4] + A[4]) - d2 * B[13]);
    g[13] = kSqrt2 * (d3 * (A[2] + A[2]) - d2 * B[9]);
    g[14] = kSqrt2 * (-d2 * B[8]);
    g[15] = 2. * (d3 * (A[5] + A[3]) - d2 * B[12]);
    g[16] = 2. * (d3 * (A[1]) - d2 * B[8]);
    g[17] = 2. * (d3 * (A[2]) - d2 * B[7]);
    g[18] = 2. * (d3 * (A[0] + A[5]) - d2 * B[5]);
    g[19] = 2. * (d3 * (A[4]) - d2 * B[4]);
    g[20] = 2. * (d3 * (A[3] + A[0]) - d2 * B[3]);
  }
}

// y = G x for the symmetric matrix stored as an upper triangle.
template <int SYM>
JX_HD void sym_matvec(const double *g, const double *x, double *y) {
  for (int i = 0; i < SYM; ++i) {
    double acc = 0.;
    for (int j = 0; j < SYM; ++j) {
      int a = i < j ? i : j, bb = i < j ? j : i;
      acc += g[tri_index(SYM, a, bb)] * x[j];
    }
    y[i] = acc;
  }
}

}  // namespace jx
