// Register-resident FFT of power-of-two length L = 16 ... 1024 for the hot
// passes.
//
// Each transform of length L is done by TPP = L / 16 threads, each holding
// 16 complex values.  Stages have radix 16 except possibly the last one
// (L = 2^a 16^b): 32 = 16 x 2, 64 = 16 x 4, 128 = 16 x 8, 256 = 16 x 16,
// 512 = 16 x 16 x 2, 1024 = 16 x 16 x 4.  In a stage of radix R a thread
// runs 16 / R butterflies j = t + TPP q (q < 16 / R) of the Stockham
// recurrence (inputs src[j + r L/R], twiddles w^(r k), k = j mod Ns,
// outputs dst[(j / Ns) Ns R + k + r Ns]); consecutive stages exchange data
// through LDS.  The first stage reads x[t + TPP r] (r < 16) and the last
// one leaves X[t + TPP q + r L/R_last] in registers, so loads and stores
// are plain strided patterns.  Stage twiddles come from the table
// W[m] = exp(-2 pi i m / L) (one correctly rounded entry per factor), read
// from global memory or from a quarter table staged in LDS.
#pragma once

#include "jx_fft_lds.h"

namespace jx {

template <int L>
struct RegPlan;
template <> struct RegPlan<16> { static constexpr int n = 1, r0 = 16, r1 = 1, r2 = 1; };
template <> struct RegPlan<32> { static constexpr int n = 2, r0 = 16, r1 = 2, r2 = 1; };
template <> struct RegPlan<64> { static constexpr int n = 2, r0 = 16, r1 = 4, r2 = 1; };
template <> struct RegPlan<128> { static constexpr int n = 2, r0 = 16, r1 = 8, r2 = 1; };
template <> struct RegPlan<256> { static constexpr int n = 2, r0 = 16, r1 = 16, r2 = 1; };
template <> struct RegPlan<512> { static constexpr int n = 3, r0 = 16, r1 = 16, r2 = 2; };
template <> struct RegPlan<1024> { static constexpr int n = 3, r0 = 16, r1 = 16, r2 = 4; };

// Lengths the register kernels are instantiated for.
constexpr bool reg_fft_length(int L) {
  return L == 16 || L == 32 || L == 64 || L == 128 || L == 256 || L == 512;
}

// (a + i b) * (c + i d) with compile-time constants
__device__ __forceinline__ double2 cmulc(double2 a, double c, double d) {
  return make_double2(a.x * c - a.y * d, a.x * d + a.y * c);
}

// 16-point DFT in registers, 4 x 4 decomposition:
// n = 4 n1 + n2, k = k1 + 4 k2; X[k] = sum_n2 W4^(n2 k2) W16^(n2 k1)
// sum_n1 x[4 n1 + n2] W4^(n1 k1).
template <int DIR>
__device__ __forceinline__ void dft16(double2 *v) {
  constexpr double c1 = 0.92387953251128673848;   // cos(pi/8)
  constexpr double s1 = 0.38268343236508978178;   // sin(pi/8)
  constexpr double h = 0.70710678118654757274;    // sqrt(2)/2
  constexpr double sg = DIR < 0 ? -1.0 : 1.0;     // sign of the exponent
  double2 y[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    double2 a[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
    dft4<DIR>(a);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[4 * n2 + k1] = a[k1];
  }
  // twiddles W16^(n2 k1) = cos(m pi/8) + i sg sin(m pi/8), m = n2 k1
  y[4 * 1 + 1] = cmulc(y[5], c1, sg * s1);          // m = 1
  y[4 * 1 + 2] = cmulc(y[6], h, sg * h);            // m = 2
  y[4 * 1 + 3] = cmulc(y[7], s1, sg * c1);          // m = 3
  y[4 * 2 + 1] = cmulc(y[9], h, sg * h);            // m = 2
  y[4 * 2 + 2] = mul_i<DIR>(y[10]);                 // m = 4
  y[4 * 2 + 3] = cmulc(y[11], -h, sg * h);          // m = 6
  y[4 * 3 + 1] = cmulc(y[13], s1, sg * c1);         // m = 3
  y[4 * 3 + 2] = cmulc(y[14], -h, sg * h);          // m = 6
  y[4 * 3 + 3] = cmulc(y[15], -c1, -sg * s1);       // m = 9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    double2 b[4] = {y[k1], y[4 + k1], y[8 + k1], y[12 + k1]};
    dft4<DIR>(b);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
  }
}

template <int R, int DIR>
__device__ __forceinline__ void dft_r(double2 *v) {
  if constexpr (R == 16) dft16<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 2) dft2<DIR>(v[0], v[1]);
}

// Twiddle source of the register kernels.  TQ: the kernel keeps the first
// quarter of the table, W[m] for m < L/4, in LDS and rebuilds
// W[m] = (-i)^(m div L/4) W[m mod L/4] (exact: a swap and sign changes);
// otherwise W is the global table.
template <int L, int DIR, bool TQ>
__device__ __forceinline__ double2 reg_twiddle(const double2 *W, int m) {
  if constexpr (TQ) {
    constexpr int Q = L / 4;
    int quad = m / Q;
    double2 w = W[m - quad * Q];
    double2 r;
    switch (quad & 3) {
      case 0: r = w; break;
      case 1: r = make_double2(w.y, -w.x); break;
      case 2: r = make_double2(-w.x, -w.y); break;
      default: r = make_double2(-w.y, w.x); break;
    }
    if (DIR > 0) r.y = -r.y;
    return r;
  } else {
    return twiddle<DIR>(W, m);
  }
}

// Entries of the twiddle table a register kernel stages into LDS.
template <int L, bool TQ>
constexpr int reg_twiddle_entries() {
  return TQ ? L / 4 : 0;
}

// One stage on the registers: twiddle, then 16 / R butterflies of radix R.
template <int L, int R, int NSP, int DIR, bool TQ>
__device__ __forceinline__ void reg_stage(double2 *v, int t, const double2 *W) {
  constexpr int TPP = L / 16;
  constexpr int tws = L / (NSP * R);
#pragma unroll
  for (int q = 0; q < 16 / R; ++q) {
    if constexpr (NSP > 1) {
      int j = t + TPP * q;
      int k = j % NSP;
#pragma unroll
      for (int r = 1; r < R; ++r)
        v[q * R + r] = cmul(v[q * R + r], reg_twiddle<L, DIR, TQ>(W, r * k * tws));
    }
    dft_r<R, DIR>(&v[q * R]);
  }
}

// LDS position of element p of a sequence: one pad entry every 16 entries
// spreads the stride-16 accesses of the radix-16 stages over the banks.
__device__ __forceinline__ int lp(int p) { return p + (p >> 4); }
// LDS entries per sequence for length L (odd, so that sequences start on
// different banks).
constexpr int reg_lds_stride(int L) { return L + L / 16 + 1; }

// Write the outputs of a stage (radix R, NSP) to LDS sequence `seq`.
template <int L, int R, int NSP>
__device__ __forceinline__ void reg_store_stage(const double2 *v, int t,
                                                double2 *seq) {
  constexpr int TPP = L / 16;
#pragma unroll
  for (int q = 0; q < 16 / R; ++q) {
    int j = t + TPP * q;
    int o = (j / NSP) * NSP * R + (j % NSP);
#pragma unroll
    for (int r = 0; r < R; ++r) seq[lp(o + r * NSP)] = v[q * R + r];
  }
}

// Read the inputs of a stage of radix R from LDS sequence `seq`.
template <int L, int R>
__device__ __forceinline__ void reg_load_stage(double2 *v, int t,
                                               const double2 *seq) {
  constexpr int TPP = L / 16;
#pragma unroll
  for (int q = 0; q < 16 / R; ++q) {
    int j = t + TPP * q;
#pragma unroll
    for (int r = 0; r < R; ++r) v[q * R + r] = seq[lp(j + r * (L / R))];
  }
}

// Position of register slot i after the last stage.
template <int L>
__device__ __forceinline__ int reg_out_pos(int t, int i) {
  using P = RegPlan<L>;
  constexpr int TPP = L / 16;
  constexpr int RL = P::n == 1 ? P::r0 : (P::n == 2 ? P::r1 : P::r2);
  int q = i / RL, r = i - q * RL;
  return t + TPP * q + r * (L / RL);
}

// Position of register slot i before the first stage.
template <int L>
__device__ __forceinline__ int reg_in_pos(int t, int i) {
  return t + (L / 16) * i;
}

// Full transform.  On entry v[i] = x[reg_in_pos(t, i)]; on exit
// v[i] = X[reg_out_pos(t, i)].  `seq` is this sequence's LDS scratch of
// L entries; the caller's block must reach the barriers inside (every thread
// of the block calls this function the same number of times).
template <int L, int DIR, bool TQ>
__device__ __forceinline__ void reg_fft(double2 *v, int t, double2 *seq,
                                        const double2 *W) {
  using P = RegPlan<L>;
  reg_stage<L, P::r0, 1, DIR, TQ>(v, t, W);
  if constexpr (P::n >= 2) {
    reg_store_stage<L, P::r0, 1>(v, t, seq);
    __syncthreads();
    reg_load_stage<L, P::r1>(v, t, seq);
    reg_stage<L, P::r1, P::r0, DIR, TQ>(v, t, W);
  }
  if constexpr (P::n >= 3) {
    __syncthreads();
    reg_store_stage<L, P::r1, P::r0>(v, t, seq);
    __syncthreads();
    reg_load_stage<L, P::r2>(v, t, seq);
    reg_stage<L, P::r2, P::r0 * P::r1, DIR, TQ>(v, t, W);
  }
}

}  // namespace jx
