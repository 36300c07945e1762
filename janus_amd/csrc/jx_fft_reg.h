// Register-resident FFT for the hot lengths.
//
// A transform of length L is done by TPP = L / V threads, each holding V
// complex values; every stage radix divides V, so in a stage of radix R a
// thread runs V / R butterflies j = t + TPP q (q < V / R) of the Stockham
// recurrence (inputs src[j + r L/R], twiddles w^(r k), k = j mod Ns,
// outputs dst[(j / Ns) Ns R + k + r Ns]); consecutive stages exchange data
// through LDS.  Plans (V; radices), variant 0 and, where listed, 1:
//   16 (16; 16)   32 (16; 16 2)   64 (16; 16 4) | (8; 8 8)
//   128 (16; 16 8) | (8; 8 4 4)   256 (16; 16 16) | (8; 8 8 4)
//   512 (16; 16 16 2) | (8; 8 8 8)   1024 (16; 16 16 4)
//   80 (20; 5 4 4) | (10; 10 2 2 2)   96 (12; 6 4 4) | (6; 6 2 2 2 2)
//   160 (20; 10 4 4) | (10; 10 2 2 2 2)
// Variant 1 holds half the values per thread: fewer registers and more
// waves per SIMD for one more LDS exchange; each pass picks its variant.
// The first stage reads x[t + TPP (i / R0) + (i mod R0) L / R0] and the last
// one leaves X[t + TPP q + r L / R_last] (i = q R_last + r) in registers, so
// loads and stores are plain strided patterns.  Stage twiddles come from the
// table W[m] = exp(-2 pi i m / L) (one correctly rounded entry per factor),
// read from global memory or from a quarter table staged in LDS.
#pragma once

#include "jx_fft_lds.h"

namespace jx {

// A plan: length L, V values per thread, n stages of radices r0 .. r4
// (unused ones 1); rl is the last stage's radix.
template <int L_, int V_, int N_, int R0_, int R1_, int R2_, int R3_ = 1,
          int R4_ = 1>
struct Plan {
  static constexpr int L = L_, V = V_, n = N_, r0 = R0_, r1 = R1_, r2 = R2_,
                       r3 = R3_, r4 = R4_;
  static constexpr int radix(int s) {
    return s == 0 ? r0 : s == 1 ? r1 : s == 2 ? r2 : s == 3 ? r3 : r4;
  }
  static constexpr int rl = radix(N_ - 1);
};

template <int L, int K>
struct RegPlan;
#define JX_REG_PLAN(LL, KK, VV, N, ...)                                     \
  template <>                                                              \
  struct RegPlan<LL, KK> {                                                 \
    using type = Plan<LL, VV, N, __VA_ARGS__>;                             \
  };
JX_REG_PLAN(16, 0, 16, 1, 16, 1, 1)
JX_REG_PLAN(32, 0, 16, 2, 16, 2, 1)
JX_REG_PLAN(64, 0, 16, 2, 16, 4, 1)
JX_REG_PLAN(64, 1, 8, 2, 8, 8, 1)
JX_REG_PLAN(128, 0, 16, 2, 16, 8, 1)
JX_REG_PLAN(128, 1, 8, 3, 8, 4, 4)
JX_REG_PLAN(256, 0, 16, 2, 16, 16, 1)
JX_REG_PLAN(256, 1, 8, 3, 8, 8, 4)
JX_REG_PLAN(512, 0, 16, 3, 16, 16, 2)
JX_REG_PLAN(512, 1, 8, 3, 8, 8, 8)
JX_REG_PLAN(1024, 0, 16, 3, 16, 16, 4)
JX_REG_PLAN(80, 0, 20, 3, 5, 4, 4)
JX_REG_PLAN(96, 0, 12, 3, 6, 4, 4)
JX_REG_PLAN(160, 0, 20, 3, 10, 4, 4)
// Ten values per thread for the lengths with a factor 5: half the registers
// and half the LDS per wave of the 20-value plans (a 20-value plan holds
// 20 KB of LDS per wave, which caps its blocks at 6 waves per CU), for two
// more LDS exchanges of radix 2.
JX_REG_PLAN(80, 1, 10, 4, 10, 2, 2, 2)
JX_REG_PLAN(160, 1, 10, 5, 10, 2, 2, 2, 2)
// Six values per thread for 96 points, the same trade.
JX_REG_PLAN(96, 1, 6, 5, 6, 2, 2, 2, 2)
#undef JX_REG_PLAN

// Lengths the register kernels are instantiated for.
constexpr bool reg_fft_length(int L) {
  return L == 16 || L == 32 || L == 64 || L == 128 || L == 256 || L == 512 ||
         L == 1024 ||
         L == 80 || L == 96 || L == 160;
}

// Plan variants compiled for a length.
constexpr int reg_variants(int L) {
  return (L == 64 || L == 128 || L == 256 || L == 512 || L == 80 ||
          L == 96 || L == 160) ? 2 : 1;
}

// Values held per thread, threads per transform.
template <class P>
constexpr int reg_v() { return P::V; }
template <class P>
constexpr int reg_tpp() { return P::L / P::V; }

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

// 3-point DFT: y1,2 = v0 - (v1 + v2)/2 +- DIR i (sqrt3/2)(v1 - v2)
template <int DIR>
__device__ __forceinline__ void dft3(double2 *v) {
  constexpr double s3 = 0.86602540378443864676;  // sqrt(3)/2
  double2 t1 = cadd(v[1], v[2]);
  double2 t2 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
  double2 t3 = make_double2(s3 * (v[1].x - v[2].x), s3 * (v[1].y - v[2].y));
  t3 = mul_i<DIR>(t3);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

// 5-point DFT with c_k = cos(2 pi k / 5), s_k = sin(2 pi k / 5):
// a1 = v1 + v4, a2 = v2 + v3, b1 = v1 - v4, b2 = v2 - v3,
// y1,4 = v0 + c1 a1 + c2 a2 +- DIR i (s1 b1 + s2 b2),
// y2,3 = v0 + c2 a1 + c1 a2 +- DIR i (s2 b1 - s1 b2).
template <int DIR>
__device__ __forceinline__ void dft5(double2 *v) {
  constexpr double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
  constexpr double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
  double2 a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
  double2 b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
  double2 t1 = make_double2(v[0].x + c1 * a1.x + c2 * a2.x,
                            v[0].y + c1 * a1.y + c2 * a2.y);
  double2 t2 = make_double2(v[0].x + c2 * a1.x + c1 * a2.x,
                            v[0].y + c2 * a1.y + c1 * a2.y);
  double2 u1 = mul_i<DIR>(make_double2(s1 * b1.x + s2 * b2.x,
                                       s1 * b1.y + s2 * b2.y));
  double2 u2 = mul_i<DIR>(make_double2(s2 * b1.x - s1 * b2.x,
                                       s2 * b1.y - s1 * b2.y));
  v[0] = cadd(v[0], cadd(a1, a2));
  v[1] = cadd(t1, u1);
  v[4] = csub(t1, u1);
  v[2] = cadd(t2, u2);
  v[3] = csub(t2, u2);
}

// 2 R2-point DFT (R2 = 3, 5): n = R2 n1 + n2, k = k1 + 2 k2; pairs over n1,
// twiddles W_{2 R2}^(n2 k1), then R2-point DFTs over n2.  cw/sw hold
// cos/sin(2 pi m / (2 R2)) for m = 0 .. R2 - 1.
template <int R2, int DIR>
__device__ __forceinline__ void dft2xr(double2 *v, const double *cw,
                                       const double *sw) {
  double2 a[R2], b[R2];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    a[n2] = cadd(v[n2], v[R2 + n2]);
    double2 d = csub(v[n2], v[R2 + n2]);
    b[n2] = n2 ? cmulc(d, cw[n2], DIR * sw[n2]) : d;
  }
  if constexpr (R2 == 3) {
    dft3<DIR>(a);
    dft3<DIR>(b);
  } else {
    dft5<DIR>(a);
    dft5<DIR>(b);
  }
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) {
    v[2 * k2] = a[k2];
    v[2 * k2 + 1] = b[k2];
  }
}

template <int DIR>
__device__ __forceinline__ void dft6(double2 *v) {
  constexpr double cw[3] = {1.0, 0.5, -0.5};
  constexpr double sw[3] = {0.0, 0.86602540378443864676, 0.86602540378443864676};
  dft2xr<3, DIR>(v, cw, sw);
}

template <int DIR>
__device__ __forceinline__ void dft10(double2 *v) {
  constexpr double cw[5] = {1.0, 0.80901699437494742410, 0.30901699437494742410,
                            -0.30901699437494742410, -0.80901699437494742410};
  constexpr double sw[5] = {0.0, 0.58778525229247312917, 0.95105651629515357212,
                            0.95105651629515357212, 0.58778525229247312917};
  dft2xr<5, DIR>(v, cw, sw);
}

template <int R, int DIR>
__device__ __forceinline__ void dft_r(double2 *v) {
  if constexpr (R == 16) dft16<DIR>(v);
  else if constexpr (R == 10) dft10<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 6) dft6<DIR>(v);
  else if constexpr (R == 5) dft5<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 3) dft3<DIR>(v);
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

// One stage on the registers: twiddle, then V / R butterflies of radix R.
template <class P, int R, int NSP, int DIR, bool TQ>
__device__ __forceinline__ void reg_stage(double2 *v, int t, const double2 *W) {
  constexpr int L = P::L;
  constexpr int V = reg_v<P>();
  constexpr int TPP = reg_tpp<P>();
  constexpr int tws = L / (NSP * R);
#pragma unroll
  for (int q = 0; q < V / R; ++q) {
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
// LDS entries per sequence for length L.  The 3 past L + L/16 (odd, so that
// sequences start on different banks) places consecutive sequences three
// 16-byte bank groups apart: at 256^3 the fused update pass, the Gamma-hat
// pass and pass A measured 26-31% fewer LDS bank-conflict cycles
// (SQ_LDS_BANK_CONFLICT) than with 1, and 0.3% more iterations/s; 9 pushes
// the 12-sequence Gamma-hat block past three blocks per CU.
#ifndef JX_LDS_SEQ_PAD
#define JX_LDS_SEQ_PAD 3
#endif
constexpr int reg_lds_stride(int L) { return L + L / 16 + JX_LDS_SEQ_PAD; }

// Write the outputs of a stage (radix R, NSP) to LDS sequence `seq`.
template <class P, int R, int NSP>
__device__ __forceinline__ void reg_store_stage(const double2 *v, int t,
                                                double2 *seq) {
  constexpr int V = reg_v<P>();
  constexpr int TPP = reg_tpp<P>();
#pragma unroll
  for (int q = 0; q < V / R; ++q) {
    int j = t + TPP * q;
    int o = (j / NSP) * NSP * R + (j % NSP);
#pragma unroll
    for (int r = 0; r < R; ++r) seq[lp(o + r * NSP)] = v[q * R + r];
  }
}

// Read the inputs of a stage of radix R from LDS sequence `seq`.
template <class P, int R>
__device__ __forceinline__ void reg_load_stage(double2 *v, int t,
                                               const double2 *seq) {
  constexpr int L = P::L;
  constexpr int V = reg_v<P>();
  constexpr int TPP = reg_tpp<P>();
#pragma unroll
  for (int q = 0; q < V / R; ++q) {
    int j = t + TPP * q;
#pragma unroll
    for (int r = 0; r < R; ++r) v[q * R + r] = seq[lp(j + r * (L / R))];
  }
}

// Position of register slot i after the last stage.
template <class P>
__device__ __forceinline__ int reg_out_pos(int t, int i) {
  constexpr int L = P::L;
  constexpr int TPP = reg_tpp<P>();
  constexpr int RL = P::rl;
  int q = i / RL, r = i - q * RL;
  return t + TPP * q + r * (L / RL);
}

// Position of register slot i before the first stage.
template <class P>
__device__ __forceinline__ int reg_in_pos(int t, int i) {
  constexpr int L = P::L;
  constexpr int TPP = reg_tpp<P>();
  constexpr int R0 = P::r0;
  int q = i / R0, r = i - q * R0;
  return t + TPP * q + r * (L / R0);
}

// Full transform.  On entry v[i] = x[reg_in_pos(t, i)] (i < V); on exit
// v[i] = X[reg_out_pos(t, i)].  `seq` is this sequence's LDS scratch; the
// caller's block must reach the barriers inside (every thread of the block
// calls this function the same number of times).
// Stages S .. n-1 of a transform; NSP is the product of the radices of the
// stages before S, whose outputs are in registers.
template <class P, int S, int NSP, int DIR, bool TQ>
__device__ __forceinline__ void reg_fft_stages(double2 *v, int t, double2 *seq,
                                               const double2 *W) {
  if constexpr (S < P::n) {
    constexpr int RP = P::radix(S - 1), R = P::radix(S);
    if constexpr (S >= 2) __syncthreads();
    reg_store_stage<P, RP, NSP / RP>(v, t, seq);
    __syncthreads();
    reg_load_stage<P, R>(v, t, seq);
    reg_stage<P, R, NSP, DIR, TQ>(v, t, W);
    reg_fft_stages<P, S + 1, NSP * R, DIR, TQ>(v, t, seq, W);
  }
}

template <class P, int DIR, bool TQ>
__device__ __forceinline__ void reg_fft(double2 *v, int t, double2 *seq,
                                        const double2 *W) {
  reg_stage<P, P::r0, 1, DIR, TQ>(v, t, W);
  reg_fft_stages<P, 1, P::r0, DIR, TQ>(v, t, seq, W);
}

// Calls f(P{}) with the plan type of variant k of length L (variant 0 when
// L has no variant k); returns JX_EUNSUPPORTED for other lengths.
template <int LL, typename F>
int with_variant(int k, F &&f) {
  if constexpr (reg_variants(LL) > 1) {
    if (k == 1) return f(typename RegPlan<LL, 1>::type());
  }
  return f(typename RegPlan<LL, 0>::type());
}

template <typename F>
int with_plan(int L, int k, F &&f) {
  switch (L) {
    case 16: return with_variant<16>(k, f);
    case 32: return with_variant<32>(k, f);
    case 64: return with_variant<64>(k, f);
    case 128: return with_variant<128>(k, f);
    case 256: return with_variant<256>(k, f);
    case 512: return with_variant<512>(k, f);
    case 1024: return with_variant<1024>(k, f);
    case 80: return with_variant<80>(k, f);
    case 96: return with_variant<96>(k, f);
    case 160: return with_variant<160>(k, f);
    default: return fail(JX_EUNSUPPORTED, "no register FFT for this length");
  }
}

}  // namespace jx
