// Mixed-radix Stockham FFT of several complex sequences held in LDS.
//
// Sequences are stored sequence-major (seq * Ls + element) as double2
// {re, im}.  Each stage of radix R reads the L/R butterflies of the source
// buffer and writes the destination buffer (ping-pong), following the
// self-sorting Stockham recurrence: with Ns the product of the radices of the
// previous stages, butterfly j (k = j mod Ns) reads src[j + r L/R], applies
// the twiddles w^(r k) with w = exp(-+2 pi i / (Ns R)), and its R-point DFT
// output r goes to dst[(j / Ns) Ns R + k + r Ns].  The output is in natural
// order.  Twiddles come from the table W[m] = exp(-2 pi i m / L) (inverse:
// conjugate), so every factor is one correctly rounded table entry.
// Radices 2, 4 and 8 have register butterflies; any other radix uses a
// direct DFT through the same table.
#pragma once

#include "jx_internal.h"

namespace jx {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
// multiply by -i (DIR = -1) or +i (DIR = +1)
template <int DIR>
__device__ __forceinline__ double2 mul_i(double2 a) {
  return DIR < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}
template <int DIR>
__device__ __forceinline__ double2 twiddle(const double2 *W, int m) {
  double2 w = W[m];
  if (DIR > 0) w.y = -w.y;
  return w;
}

template <int DIR>
__device__ __forceinline__ void dft2(double2 &a, double2 &b) {
  double2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(double2 *v) {
  double2 a0 = cadd(v[0], v[2]);
  double2 a1 = csub(v[0], v[2]);
  double2 b0 = cadd(v[1], v[3]);
  double2 b1 = mul_i<DIR>(csub(v[1], v[3]));
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, b1);
  v[3] = csub(a1, b1);
}

template <int DIR>
__device__ __forceinline__ void dft8(double2 *v) {
  constexpr double h = 0.70710678118654752440;  // sqrt(2) / 2
  // radix-2 split: even / odd
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // o[k] *= w8^k, w8 = exp(DIR * 2 pi i / 8)
  double2 t1 = o[1], t3 = o[3];
  o[1] = make_double2(h * (t1.x - DIR * t1.y), h * (t1.y + DIR * t1.x));
  o[2] = mul_i<DIR>(o[2]);
  o[3] = make_double2(h * (-t3.x - DIR * t3.y), h * (-t3.y + DIR * t3.x));
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}

template <int R, int DIR>
__device__ __forceinline__ void stage_fixed(const double2 *__restrict__ src,
                                            double2 *__restrict__ dst, int j,
                                            int k, int Lr, int Ns, int tws,
                                            const double2 *W) {
  double2 v[R];
  v[0] = src[j];
#pragma unroll
  for (int r = 1; r < R; ++r)
    v[r] = cmul(src[j + r * Lr], twiddle<DIR>(W, r * k * tws));
  if (R == 2) dft2<DIR>(v[0], v[1]);
  if (R == 4) dft4<DIR>(v);
  if (R == 8) dft8<DIR>(v);
  int o = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[o + r * Ns] = v[r];
}

// Direct DFT of any radix R (O(R^2)); the R-th roots of unity are the table
// entries at multiples of L/R.
template <int DIR>
__device__ void stage_generic(const double2 *__restrict__ src,
                              double2 *__restrict__ dst, int R, int j, int k,
                              int Lr, int Ns, int tws, const double2 *W) {
  int o = (j / Ns) * Ns * R + k;
  for (int q = 0; q < R; ++q) {
    double2 acc = make_double2(0., 0.);
    for (int r = 0; r < R; ++r) {
      double2 v = src[j + r * Lr];
      if (r) v = cmul(v, twiddle<DIR>(W, r * k * tws));
      int m = ((q * r) % R) * Lr;
      acc = cadd(acc, cmul(v, twiddle<DIR>(W, m)));
    }
    dst[o + q * Ns] = acc;
  }
}

// Transforms nseq sequences of length P.L stored with stride Ls >= L; a and
// b are LDS buffers of nseq * Ls entries, data in a on entry.  Returns the
// buffer holding the result.  All threads of the block must call it.
template <int DIR>
__device__ double2 *lds_fft(double2 *a, double2 *b, int nseq, const FftLen &P,
                            int Ls, int tid, int nth) {
  const int L = P.L;
  int Ns = 1;
  for (int s = 0; s < P.nstages; ++s) {
    const int R = P.radix[s];
    const int Lr = L / R;
    const int tws = L / (Ns * R);
    const int total = nseq * Lr;
    for (int bf = tid; bf < total; bf += nth) {
      int seq = bf / Lr;
      int j = bf - seq * Lr;
      int k = j % Ns;
      const double2 *src = a + seq * Ls;
      double2 *dst = b + seq * Ls;
      switch (R) {
        case 2: stage_fixed<2, DIR>(src, dst, j, k, Lr, Ns, tws, P.W); break;
        case 4: stage_fixed<4, DIR>(src, dst, j, k, Lr, Ns, tws, P.W); break;
        case 8: stage_fixed<8, DIR>(src, dst, j, k, Lr, Ns, tws, P.W); break;
        default: stage_generic<DIR>(src, dst, R, j, k, Lr, Ns, tws, P.W);
      }
    }
    __syncthreads();
    double2 *t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

}  // namespace jx
