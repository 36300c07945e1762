// Device kernels of the Lippmann-Schwinger hot path and their launchers.
//
// Data layouts (FP64, C order):
//   real field  (rows, L, C)        rows = n0 [* n1], L = n_last, C comps
//   spectrum    (rows, nh, C, 2)    nh = L / 2 + 1; complex values, the C
//                                   components of a frequency side by side
//                                   (for C = 1: janus's r2c layout,
//                                   _serial_fft.pyx:57-74, 181-200)
//
// Pass structure of one application of Gamma0 (3D; 2D drops pass B/D):
//   A  rows_forward   last-axis r2c (+ optional local operator)
//   B  axis(1, fwd)
//   C  axis(0, fwd -> Gamma-hat -> inv)
//   D  axis(1, inv)
//   E  rows_inverse   last-axis c2r, 1/N (+ optional fixed-point update)
// Real sequences are paired two by two as the real and imaginary parts of
// one complex sequence, which is then transformed once.

#include <cmath>

#include "jx_fft_lds.h"
#include "jx_fft_reg.h"
#include "jx_gamma.h"
#include "jx_internal.h"

namespace jx {

// LDS row stride of a sequence: one complex pad for even lengths breaks the
// power-of-two strides between sequences.
static inline int lds_stride(int L) { return (L % 2 == 0) ? L + 1 : L; }

// ------------------------------------------------------------ local operator
// (sph J + dev K) : e in Mandel notation, janus/operators.pyx:315-324 (2D)
// and :352-365 (3D), tr = (sph - dev) / dim (:299).  Contraction into FMA is
// disabled so that the result is the one of the separate products and sums.
template <int C>
__device__ __forceinline__ void iso4(double tr, double dev, const double *e,
                                     double *t) {
#pragma clang fp contract(off)
  if (C == 3) {
    double aux = tr * (e[0] + e[1]);
    t[0] = aux + dev * e[0];
    t[1] = aux + dev * e[1];
    t[2] = dev * e[2];
  } else {
    double aux = tr * (e[0] + e[1] + e[2]);
    t[0] = aux + dev * e[0];
    t[1] = aux + dev * e[1];
    t[2] = aux + dev * e[2];
    t[3] = dev * e[3];
    t[4] = dev * e[4];
    t[5] = dev * e[5];
  }
}

// Phase tables hold (tr, dev) per phase; field mode computes tr per voxel.
template <int C>
__device__ __forceinline__ void local_voxel(const LocalDev &loc, int64_t vox,
                                            const double *e, double *t) {
  double tr, dev;
  if (loc.mode == LOCAL_PHASES) {
    int p = loc.phase[vox];
    tr = loc.table[2 * p];
    dev = loc.table[2 * p + 1];
  } else {
    double sph = loc.sph[vox];
    dev = loc.dev[vox];
    tr = (sph - dev) / (C == 3 ? 2.0 : 3.0);
  }
  iso4<C>(tr, dev, e, t);
}

// ------------------------------------------------------------ update
// Fixed-point update of one component: (E + rhs) - scale * y, evaluated the
// same way (no contraction into FMA) in every kernel, so that all schedules
// produce the same iterates bit for bit.
__device__ __forceinline__ double ls_new(double e, double y, double scale) {
#pragma clang fp contract(off)
  double val = scale * y;
  return e - val;
}
__device__ __forceinline__ double ls_new_rhs(double e, double h, double y,
                                             double scale) {
#pragma clang fp contract(off)
  double val = scale * y;
  return (e + h) - val;
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Deterministic block sum (fixed order); result valid in thread 0.
__device__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) s += red[w];
  }
  return s;
}

struct RowsArgs {
  int64_t nrows;
  int L, nh, Ls, RB;
  double scale;
};

__device__ __forceinline__ bool solve_done(const SolveState *st) {
  return st != nullptr && *((volatile const int *)&st->done) != 0;
}

// ------------------------------------------------------------ pass A
template <int C>
__global__ void __launch_bounds__(kThreads)
    k_rows_forward(const double *__restrict__ x, double *__restrict__ spec,
                   RowsArgs a, FftLen P, LocalDev loc, const SolveState *st) {
  if (solve_done(st)) return;
  extern __shared__ double2 smem[];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int64_t row0 = (int64_t)blockIdx.x * a.RB;
  const int rb = (int)min((int64_t)a.RB, a.nrows - row0);
  const int Q = rb * C;
  const int nseq = (Q + 1) / 2;
  const int nseq_max = (a.RB * C + 1) / 2;
  double2 *bufA = smem;
  double2 *bufB = smem + (size_t)nseq_max * a.Ls;
  const int L = a.L;

  for (int v = tid; v < rb * L; v += nth) {
    int r = v / L, j = v - r * L;
    int64_t vox = (row0 + r) * L + j;
    double e[C], t[C];
#pragma unroll
    for (int c = 0; c < C; ++c) e[c] = x[vox * C + c];
    if (loc.mode != LOCAL_NONE) {
      local_voxel<C>(loc, vox, e, t);
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) t[c] = e[c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      int q = r * C + c;
      double *slot = (double *)&bufA[(q >> 1) * a.Ls + j];
      slot[q & 1] = t[c];
    }
    if ((Q & 1) && r == rb - 1) bufA[((Q - 1) >> 1) * a.Ls + j].y = 0.;
  }
  __syncthreads();
  double2 *res = lds_fft<-1>(bufA, bufB, nseq, P, a.Ls, tid, nth);

  const int nh = a.nh;
  for (int v = tid; v < rb * nh * C; v += nth) {
    int c = v % C;
    int rest = v / C;
    int k = rest % nh;
    int r = rest / nh;
    int q = r * C + c;
    const double2 *Z = res + (q >> 1) * a.Ls;
    double2 zk = Z[k];
    double2 zm = Z[k == 0 ? 0 : L - k];
    double re, im;
    if (!(q & 1)) {
      re = 0.5 * (zk.x + zm.x);
      im = 0.5 * (zk.y - zm.y);
    } else {
      re = 0.5 * (zk.y + zm.y);
      im = -0.5 * (zk.x - zm.x);
    }
    int64_t pt = (row0 + r) * nh + k;
    spec[pt * 2 * C + 2 * c] = re;
    spec[pt * 2 * C + 2 * c + 1] = im;
  }
}

// ------------------------------------------------------------ passes B, C, D
struct AxisArgs {
  int L, Ls, T, ntiles;
  int64_t stride_ax, n_outer, stride_outer, n_inner;
  int nh;  // last-axis half length, decodes the inner index in 3D pass C
};

template <int C, int MODE>
__global__ void __launch_bounds__(kThreads)
    k_axis(const double *in, double *out, AxisArgs a, FftLen P, GreenDev G,
           const SolveState *st) {
  if (solve_done(st)) return;
  extern __shared__ double2 smem[];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int64_t o = blockIdx.x / a.ntiles;
  const int64_t m0 = (int64_t)(blockIdx.x % a.ntiles) * a.T;
  const int tc = (int)min((int64_t)a.T, a.n_inner - m0);
  const int W2 = tc * 2 * C;
  const int nseq = tc * C;
  const int64_t base = o * a.stride_outer + m0 * 2 * C;
  double2 *bufA = smem;
  double2 *bufB = smem + (size_t)a.T * C * a.Ls;
  const int L = a.L;

  for (int v = tid; v < L * W2; v += nth) {
    int e = v / W2;
    int w = v - e * W2;
    int t = w / (2 * C);
    int rem = w - t * 2 * C;
    int part = rem & 1;
    int c = rem >> 1;
    double val = in[base + (int64_t)e * a.stride_ax + w];
    ((double *)&bufA[(t * C + c) * a.Ls + e])[part] = val;
  }
  __syncthreads();

  double2 *res;
  if (MODE == 0) {
    res = lds_fft<-1>(bufA, bufB, nseq, P, a.Ls, tid, nth);
    if constexpr (C == 3 || C == 6) {
      constexpr int DIM = C == 3 ? 2 : 3;
      constexpr int NC = C == 3 ? 6 : 21;
      for (int v = tid; v < tc * L; v += nth) {
        int t = v / L, e = v - t * L;
        int64_t m = m0 + t;
        int b[3];
        b[0] = e;
        if (DIM == 2) {
          b[1] = (int)m;
        } else {
          b[1] = (int)(m / a.nh);
          b[2] = (int)(m - (int64_t)b[1] * a.nh);
        }
        double g[NC];
        gamma_hat_moments<DIM>(G, b, g);
        double xr[C], xi[C], yr[C], yi[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          double2 z = res[(t * C + c) * a.Ls + e];
          xr[c] = z.x;
          xi[c] = z.y;
        }
        sym_matvec<C>(g, xr, yr);
        sym_matvec<C>(g, xi, yi);
#pragma unroll
        for (int c = 0; c < C; ++c)
          res[(t * C + c) * a.Ls + e] = make_double2(yr[c], yi[c]);
      }
    }
    __syncthreads();
    double2 *other = (res == bufA) ? bufB : bufA;
    res = lds_fft<+1>(res, other, nseq, P, a.Ls, tid, nth);
  } else {
    res = lds_fft<MODE>(bufA, bufB, nseq, P, a.Ls, tid, nth);
  }

  for (int v = tid; v < L * W2; v += nth) {
    int e = v / W2;
    int w = v - e * W2;
    int t = w / (2 * C);
    int rem = w - t * 2 * C;
    int part = rem & 1;
    int c = rem >> 1;
    out[base + (int64_t)e * a.stride_ax + w] =
        ((const double *)&res[(t * C + c) * a.Ls + e])[part];
  }
}

// Gamma-hat at frequency (b0 = e, inner index m) applied to the C complex
// values z[c * zs] (real and imaginary parts separately).
template <int C>
__device__ __forceinline__ void gamma_point(const GreenDev &G, int e, int64_t m,
                                            int nh, double2 *z, int zs) {
  constexpr int DIM = C == 3 ? 2 : 3;
  constexpr int NC = C == 3 ? 6 : 21;
  int b[3];
  b[0] = e;
  if (DIM == 2) {
    b[1] = (int)m;
  } else {
    b[1] = (int)(m / nh);
    b[2] = (int)(m - (int64_t)b[1] * nh);
  }
  double g[NC];
  gamma_hat_moments<DIM>(G, b, g);
  double xr[C], xi[C], yr[C], yi[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    double2 v = z[c * zs];
    xr[c] = v.x;
    xi[c] = v.y;
  }
  sym_matvec<C>(g, xr, yr);
  sym_matvec<C>(g, xi, yi);
#pragma unroll
  for (int c = 0; c < C; ++c) z[c * zs] = make_double2(yr[c], yi[c]);
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so that tiles sharing cache lines meet in the
// same L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
  const int64_t nx = 8;
  int64_t q = nb / nx, r = nb - q * nx;
  int64_t x = b - (b / nx) * nx, slot = b / nx;
  return x * q + (x < r ? x : r) + slot;
}

// Stage the twiddle entries a register kernel reads into LDS (none when the
// global table is used).  The caller synchronizes before the first use.
template <int L, int NT, bool TQ>
__device__ __forceinline__ const double2 *stage_twiddles(const double2 *W,
                                                         double2 *tw) {
  constexpr int NTW = reg_twiddle_entries<L, TQ>();
  if constexpr (NTW == 0) {
    return W;
  } else {
    for (int i = threadIdx.x; i < NTW; i += NT) tw[i] = W[i];
    return tw;
  }
}

// Register-resident transform along a strided axis of the spectrum for
// power-of-two L.  NSEQ = T * C complex sequences per block (T points of the
// contiguous inner index, all C components), TPP = L / V threads each.
// Thread tid = t * NSEQ + s: consecutive lanes read consecutive doubles of
// the T * 2C doubles that one axis element of the tile spans.
// Minimum waves per SIMD requested for the Gamma-hat pass of the 16-value
// plans (register budget of the compiler).  3 (a few spilled registers)
// measured faster than the compiler's own choice of 2 waves at 256^3; the
// 20- and 24-value plans would spill heavily and keep the compiler's choice.
#ifndef JX_GAMMA_PASS_MIN_WAVES
#define JX_GAMMA_PASS_MIN_WAVES 3
#endif

template <int L, int C, int NSEQ, int MODE>
__global__ void __launch_bounds__(NSEQ * reg_tpp<L>(),
                                  MODE == 0 && reg_v<L>() == 16
                                      ? JX_GAMMA_PASS_MIN_WAVES
                                      : 1)
    k_axis_reg(const double *in, double *out, AxisArgs a, const double2 *W,
               GreenDev G, const SolveState *st) {
  if (solve_done(st)) return;
  constexpr int TPP = reg_tpp<L>();
  constexpr int V = reg_v<L>();
  constexpr int T = NSEQ / C;
  constexpr int LS = reg_lds_stride(L);
  extern __shared__ double2 smem[];
  const int tid = threadIdx.x;
  const int s = tid % NSEQ, t = tid / NSEQ;
  const int tp = s / C, c = s - tp * C;
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t o = blk / a.ntiles;
  const int64_t m0 = (blk - o * a.ntiles) * T;
  const int tc = (int)min((int64_t)T, a.n_inner - m0);
  const bool valid = tp < tc;
  const double *src = in + o * a.stride_outer + (m0 + tp) * 2 * C + 2 * c;
  double *dst = out + o * a.stride_outer + (m0 + tp) * 2 * C + 2 * c;
  double2 v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    if (valid) {
      v[i] = *(const double2 *)(src + (int64_t)reg_in_pos<L>(t, i) * a.stride_ax);
    } else {
      v[i] = make_double2(0., 0.);
    }
  }
  // the Gamma-hat pass reads the global table: its LDS budget is kept for
  // three blocks per CU
  constexpr bool TQ = MODE != 0;
  __shared__ double2 twl[TQ ? L / 4 : 1];
  const double2 *Wt = stage_twiddles<L, NSEQ * TPP, TQ>(W, twl);
  double2 *seq = smem + s * LS;
  reg_fft<L, (MODE > 0 ? 1 : -1), TQ>(v, t, seq, Wt);
  if constexpr (MODE == 0 && (C == 3 || C == 6)) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i) seq[lp(reg_out_pos<L>(t, i))] = v[i];
    __syncthreads();
    for (int w = tid; w < tc * L; w += NSEQ * TPP) {
      int p = w / L, e = w - p * L;
      gamma_point<C>(G, e, m0 + p, a.nh, smem + (p * C) * LS + lp(e), LS);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<L>(t, i))];
    __syncthreads();
    reg_fft<L, 1, TQ>(v, t, seq, Wt);
  }
  if (valid) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      *(double2 *)(dst + (int64_t)reg_out_pos<L>(t, i) * a.stride_ax) = v[i];
    }
  }
}

// ------------------------------------------------------------ pass E
struct LsArgs {
  const double *rhs;
  double *partials;
  double E[6];
};

template <int C, bool LS>
__global__ void __launch_bounds__(kThreads)
    k_rows_inverse(const double *__restrict__ spec, double *y, RowsArgs a,
                   FftLen P, LsArgs ls, const SolveState *st) {
  if (solve_done(st)) return;
  extern __shared__ double2 smem[];
  __shared__ double red[kThreads / 64];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int64_t row0 = (int64_t)blockIdx.x * a.RB;
  const int rb = (int)min((int64_t)a.RB, a.nrows - row0);
  const int Q = rb * C;
  const int nseq = (Q + 1) / 2;
  const int nseq_max = (a.RB * C + 1) / 2;
  double2 *bufA = smem;
  double2 *bufB = smem + (size_t)nseq_max * a.Ls;
  const int L = a.L, nh = a.nh;

  // Load the half spectra: pair p = (Xa, Xb).  Slot k holds Xa[k] and slot
  // L - k holds Xb[k] (0 < k < L/2); the DC and Nyquist slots hold
  // (Re Xa, Re Xb) -- their imaginary parts are dropped, as in FFTW's and
  // pocketfft's c2r.
  for (int v = tid; v < rb * nh * C; v += nth) {
    int c = v % C;
    int rest = v / C;
    int k = rest % nh;
    int r = rest / nh;
    int q = r * C + c;
    int part = q & 1;
    int64_t pt = (row0 + r) * nh + k;
    double re = spec[pt * 2 * C + 2 * c];
    double im = spec[pt * 2 * C + 2 * c + 1];
    double2 *Z = bufA + (q >> 1) * a.Ls;
    bool lonely = (!part) && (q + 1 == Q);
    if (k == 0 || 2 * k == L) {
      ((double *)&Z[k])[part] = re;
      if (lonely) Z[k].y = 0.;
    } else if (!part) {
      Z[k] = make_double2(re, im);
      if (lonely) Z[L - k] = make_double2(0., 0.);
    } else {
      Z[L - k] = make_double2(re, im);
    }
  }
  __syncthreads();
  // Z[k] = Xa[k] + i Xb[k], Z[L-k] = conj(Xa[k]) + i conj(Xb[k])
  const int half = (L - 1) / 2;
  for (int v = tid; v < nseq * half; v += nth) {
    int p = v / half;
    int k = 1 + (v - p * half);
    double2 *Z = bufA + p * a.Ls;
    double2 A = Z[k], B = Z[L - k];
    Z[k] = make_double2(A.x - B.y, A.y + B.x);
    Z[L - k] = make_double2(A.x + B.y, B.x - A.y);
  }
  __syncthreads();
  double2 *res = lds_fft<+1>(bufA, bufB, nseq, P, a.Ls, tid, nth);

  double acc = 0.;
  for (int v = tid; v < rb * L; v += nth) {
    int r = v / L, j = v - r * L;
    int64_t vox = (row0 + r) * L + j;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      int q = r * C + c;
      double2 z = res[(q >> 1) * a.Ls + j];
      double raw = (q & 1) ? z.y : z.x;
      double val = a.scale * raw;
      if (LS) {
        double nv = ls.rhs ? ls_new_rhs(ls.E[c], ls.rhs[vox * C + c], raw, a.scale)
                           : ls_new(ls.E[c], raw, a.scale);
        double d = nv - y[vox * C + c];
        acc += d * d;
        y[vox * C + c] = nv;
      } else {
        y[vox * C + c] = val;
      }
    }
  }
  if (LS) {
    double s = block_sum(acc, red);
    if (tid == 0) ls.partials[blockIdx.x] = s;
  }
}

// ------------------------------------------------------------ passes A, E
// Twiddle source of the last-axis kernels: quarter table in LDS (1) or the
// global table (0, fewer registers; measured as fast or faster).
#ifndef JX_ROWS_TQ
#define JX_ROWS_TQ 0
#endif

// Register-resident last-axis transforms for power-of-two L.  A block
// handles RB rows, i.e. 2 NP real sequences paired into NP complex ones
// (sequence q = r C + c; pair p = q / 2, part q % 2), TPP = L / V threads
// per pair.

template <int L, int C, int NP, bool LOCAL>
__global__ void __launch_bounds__(NP * reg_tpp<L>())
    k_rows_forward_reg(const double *__restrict__ x, double *__restrict__ spec,
                       int64_t nrows, const double2 *W, LocalDev loc,
                       const SolveState *st) {
  if (solve_done(st)) return;
  constexpr int TPP = reg_tpp<L>(), LS = reg_lds_stride(L), NH = L / 2 + 1;
  constexpr int V = reg_v<L>();
  constexpr int NT = NP * TPP;
  constexpr int RB = 2 * NP / C;
  static_assert((2 * NP) % C == 0, "whole rows per block");
  extern __shared__ double2 smem[];
  const int tid = threadIdx.x;
  const int64_t row0 = xcd_tile(blockIdx.x, gridDim.x) * RB;
  const int rb = (int)min((int64_t)RB, nrows - row0);
  __shared__ double2 twl[JX_ROWS_TQ ? L / 4 : 1];
  const double2 *Wt = stage_twiddles<L, NT, JX_ROWS_TQ>(W, twl);
  // stage the rows (local operator applied) into the pair sequences
  for (int w = tid; w < RB * L; w += NT) {
    int r = w / L, j = w - r * L;
    double t[C];
    if (r < rb) {
      int64_t vox = (row0 + r) * L + j;
      double e[C];
      if constexpr (C == 6) {
        const double2 *xv = (const double2 *)(x + vox * 6);
        double2 a = xv[0], b = xv[1], cc = xv[2];
        e[0] = a.x; e[1] = a.y; e[2] = b.x; e[3] = b.y; e[4] = cc.x; e[5] = cc.y;
      } else {
#pragma unroll
        for (int c = 0; c < C; ++c) e[c] = x[vox * C + c];
      }
      if (LOCAL) {
        local_voxel<C>(loc, vox, e, t);
      } else {
#pragma unroll
        for (int c = 0; c < C; ++c) t[c] = e[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) t[c] = 0.;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      int q = r * C + c;
      ((double *)&smem[(q >> 1) * LS + lp(j)])[q & 1] = t[c];
    }
  }
  __syncthreads();
  const int p = tid % NP, t = tid / NP;
  double2 *seq = smem + p * LS;
  double2 v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<L>(t, i))];
  __syncthreads();
  reg_fft<L, -1, JX_ROWS_TQ>(v, t, seq, Wt);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < V; ++i) seq[lp(reg_out_pos<L>(t, i))] = v[i];
  __syncthreads();
  // separate the pairs: Xa = (Z[k] + conj Z[-k]) / 2, Xb = (Z[k] - conj Z[-k]) / 2i
  if constexpr (C == 6) {
    // pair p of row r holds components (c0, c0 + 1), adjacent in memory:
    // two 16-byte stores of complex values
    for (int w = tid; w < rb * NH * 3; w += NT) {
      int cp = w % 3;
      int rest = w / 3;
      int k = rest % NH;
      int r = rest / NH;
      const double2 *Z = smem + (r * 3 + cp) * LS;
      double2 zk = Z[lp(k)];
      double2 zm = Z[lp(k == 0 ? 0 : L - k)];
      double2 xa = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      double2 xb = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
      double *sp = spec + ((row0 + r) * NH + k) * 12 + 4 * cp;
      *(double2 *)sp = xa;
      *(double2 *)(sp + 2) = xb;
    }
    return;
  }
  for (int w = tid; w < rb * NH * C; w += NT) {
    int c = w % C;
    int rest = w / C;
    int k = rest % NH;
    int r = rest / NH;
    int q = r * C + c;
    const double2 *Z = smem + (q >> 1) * LS;
    double2 zk = Z[lp(k)];
    double2 zm = Z[lp(k == 0 ? 0 : L - k)];
    double re, im;
    if (!(q & 1)) {
      re = 0.5 * (zk.x + zm.x);
      im = 0.5 * (zk.y - zm.y);
    } else {
      re = 0.5 * (zk.y + zm.y);
      im = -0.5 * (zk.x - zm.x);
    }
    int64_t pt = (row0 + r) * NH + k;
    spec[pt * 2 * C + 2 * c] = re;
    spec[pt * 2 * C + 2 * c + 1] = im;
  }
}

// MODE 0: y = c2r(spec); 1: y <- E - c2r(spec) with residual partials;
// 2: y <- (E + rhs) - c2r(spec) with residual partials.
template <int L, int C, int NP, int MODE>
__global__ void __launch_bounds__(NP * reg_tpp<L>())
    k_rows_inverse_reg(const double *__restrict__ spec, double *y,
                       int64_t nrows, double scale, const double2 *W,
                       LsArgs ls, const SolveState *st) {
  if (solve_done(st)) return;
  constexpr int TPP = reg_tpp<L>(), LS = reg_lds_stride(L), NH = L / 2 + 1;
  constexpr int V = reg_v<L>();
  constexpr int NT = NP * TPP;
  constexpr int RB = 2 * NP / C;
  constexpr bool LS_UPDATE = MODE > 0;
  extern __shared__ double2 smem[];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t row0 = blk * RB;
  const int rb = (int)min((int64_t)RB, nrows - row0);
  const int p = tid % NP, t = tid / NP;
  __shared__ double2 twl[JX_ROWS_TQ ? L / 4 : 1];
  const double2 *Wt = stage_twiddles<L, NT, JX_ROWS_TQ>(W, twl);
  // half spectra into the pair slots: slot k <- Xa[k], slot L-k <- Xb[k]
  // (0 < k < L/2); DC and Nyquist slots <- (Re Xa, Re Xb), their imaginary
  // parts dropped as FFTW's and pocketfft's c2r do
  if constexpr (C == 6) {
    for (int w = tid; w < RB * NH * 3; w += NT) {
      int cp = w % 3;
      int rest = w / 3;
      int k = rest % NH;
      int r = rest / NH;
      double2 xa = make_double2(0., 0.), xb = make_double2(0., 0.);
      if (r < rb) {
        const double *sp = spec + ((row0 + r) * NH + k) * 12 + 4 * cp;
        xa = *(const double2 *)sp;
        xb = *(const double2 *)(sp + 2);
      }
      double2 *Z = smem + (r * 3 + cp) * LS;
      if (k == 0 || 2 * k == L) {
        Z[lp(k)] = make_double2(xa.x, xb.x);
      } else {
        Z[lp(k)] = xa;
        Z[lp(L - k)] = xb;
      }
    }
  } else
  for (int w = tid; w < RB * NH * C; w += NT) {
    int c = w % C;
    int rest = w / C;
    int k = rest % NH;
    int r = rest / NH;
    int q = r * C + c;
    int part = q & 1;
    double re = 0., im = 0.;
    if (r < rb) {
      int64_t pt = (row0 + r) * NH + k;
      re = spec[pt * 2 * C + 2 * c];
      im = spec[pt * 2 * C + 2 * c + 1];
    }
    double2 *Z = smem + (q >> 1) * LS;
    if (k == 0 || 2 * k == L)
      ((double *)&Z[lp(k)])[part] = re;
    else if (!part)
      Z[lp(k)] = make_double2(re, im);
    else
      Z[lp(L - k)] = make_double2(re, im);
  }
  __syncthreads();
  constexpr int HALF = L / 2 - 1;
  if (HALF > 0) {
    for (int w = tid; w < NP * HALF; w += NT) {
      int pp = w / HALF;
      int k = 1 + (w - pp * HALF);
      double2 *Z = smem + pp * LS;
      double2 A = Z[lp(k)], B = Z[lp(L - k)];
      Z[lp(k)] = make_double2(A.x - B.y, A.y + B.x);
      Z[lp(L - k)] = make_double2(A.x + B.y, B.x - A.y);
    }
    __syncthreads();
  }
  double2 *seq = smem + p * LS;
  double2 v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<L>(t, i))];
  __syncthreads();
  reg_fft<L, 1, JX_ROWS_TQ>(v, t, seq, Wt);
  double acc = 0.;
  if constexpr (C == 6) {
    // pair p = components (c0, c0 + 1) of row r: 16-byte accesses
    const int r = p / 3, c0 = 2 * (p - r * 3);
    if (r < rb) {
      double2 *yr = (double2 *)(y + (row0 + r) * L * 6 + c0);
      const double2 *hr =
          MODE == 2 ? (const double2 *)(ls.rhs + (row0 + r) * L * 6 + c0)
                    : nullptr;
      const double e0 = ls.E[c0], e1 = ls.E[c0 + 1];
#pragma unroll
      for (int i = 0; i < V; ++i) {
        int64_t off = (int64_t)reg_out_pos<L>(t, i) * 3;
        double2 val = make_double2(scale * v[i].x, scale * v[i].y);
        if constexpr (LS_UPDATE) {
          double2 nv = make_double2(ls_new(e0, v[i].x, scale),
                                    ls_new(e1, v[i].y, scale));
          if constexpr (MODE == 2) {
            double2 h = hr[off];
            nv = make_double2(ls_new_rhs(e0, h.x, v[i].x, scale),
                              ls_new_rhs(e1, h.y, v[i].y, scale));
          }
          double2 o = yr[off];
          double dx = nv.x - o.x, dy = nv.y - o.y;
          acc += dx * dx;
          acc += dy * dy;
          yr[off] = nv;
        } else {
          yr[off] = val;
        }
      }
    }
  } else {
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const int q = 2 * p + part;
      const int r = q / C, c = q - (q / C) * C;
      if (r < rb) {
        double *yr = y + (row0 + r) * L * C + c;
        const double *hr = MODE == 2 ? ls.rhs + (row0 + r) * L * C + c : nullptr;
#pragma unroll
        for (int i = 0; i < V; ++i) {
          int64_t off = (int64_t)reg_out_pos<L>(t, i) * C;
          double raw = part ? v[i].y : v[i].x;
          double val = scale * raw;
          if constexpr (LS_UPDATE) {
            double nv = ls_new(ls.E[c], raw, scale);
            if constexpr (MODE == 2) nv = ls_new_rhs(ls.E[c], hr[off], raw, scale);
            double d = nv - yr[off];
            acc += d * d;
            yr[off] = nv;
          } else {
            yr[off] = val;
          }
        }
      }
    }
  }
  if (LS_UPDATE) {
    double sum = block_sum(acc, red);
    if (tid == 0) ls.partials[blk] = sum;
  }
}

// Fused pass E of iteration k and pass A of iteration k + 1: inverse
// last-axis transform, fixed-point update eps <- E (+ rhs) - y with the
// residual partial sums, then (C - C0) : eps and the forward last-axis
// transform of the result, written over the same spectrum rows.  The new
// iterate never travels back from memory.  MODE 1: E only; 2: E + rhs.
// Minimum waves per SIMD requested for the fused update pass: 3 (about 70
// registers spilled to scratch) measured faster than 2 at 256^3.
#ifndef JX_FUSED_MIN_WAVES
#define JX_FUSED_MIN_WAVES 3
#endif

template <int L, int C, int NP, int MODE>
__global__ void __launch_bounds__(NP * reg_tpp<L>(),
                                  reg_v<L>() == 16 ? JX_FUSED_MIN_WAVES : 1)
    k_rows_update_forward_reg(double *spec, double *y, int64_t nrows,
                              double scale, const double2 *W, LsArgs ls,
                              LocalDev loc, const SolveState *st) {
  if (solve_done(st)) return;
  constexpr int TPP = reg_tpp<L>(), LS = reg_lds_stride(L), NH = L / 2 + 1;
  constexpr int V = reg_v<L>();
  constexpr int NT = NP * TPP;
  constexpr int RB = 2 * NP / C;
  extern __shared__ double2 smem[];
  __shared__ double red[4];
  __shared__ double2 twl[JX_ROWS_TQ ? L / 4 : 1];
  const int tid = threadIdx.x;
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t row0 = blk * RB;
  const int rb = (int)min((int64_t)RB, nrows - row0);
  const int p = tid % NP, t = tid / NP;
  const double2 *Wt = stage_twiddles<L, NT, JX_ROWS_TQ>(W, twl);

  // ---- inverse: half spectra into the pair slots (see k_rows_inverse_reg)
  if constexpr (C == 6) {
    for (int w = tid; w < RB * NH * 3; w += NT) {
      int cp = w % 3;
      int rest = w / 3;
      int k = rest % NH;
      int r = rest / NH;
      double2 xa = make_double2(0., 0.), xb = make_double2(0., 0.);
      if (r < rb) {
        const double *sp = spec + ((row0 + r) * NH + k) * 12 + 4 * cp;
        xa = *(const double2 *)sp;
        xb = *(const double2 *)(sp + 2);
      }
      double2 *Z = smem + (r * 3 + cp) * LS;
      if (k == 0 || 2 * k == L) {
        Z[lp(k)] = make_double2(xa.x, xb.x);
      } else {
        Z[lp(k)] = xa;
        Z[lp(L - k)] = xb;
      }
    }
  } else {
    for (int w = tid; w < RB * NH * C; w += NT) {
      int c = w % C;
      int rest = w / C;
      int k = rest % NH;
      int r = rest / NH;
      int q = r * C + c;
      int part = q & 1;
      double re = 0., im = 0.;
      if (r < rb) {
        int64_t pt = (row0 + r) * NH + k;
        re = spec[pt * 2 * C + 2 * c];
        im = spec[pt * 2 * C + 2 * c + 1];
      }
      double2 *Z = smem + (q >> 1) * LS;
      if (k == 0 || 2 * k == L)
        ((double *)&Z[lp(k)])[part] = re;
      else if (!part)
        Z[lp(k)] = make_double2(re, im);
      else
        Z[lp(L - k)] = make_double2(re, im);
    }
  }
  __syncthreads();
  constexpr int HALF = L / 2 - 1;
  if (HALF > 0) {
    for (int w = tid; w < NP * HALF; w += NT) {
      int pp = w / HALF;
      int k = 1 + (w - pp * HALF);
      double2 *Z = smem + pp * LS;
      double2 A = Z[lp(k)], B = Z[lp(L - k)];
      Z[lp(k)] = make_double2(A.x - B.y, A.y + B.x);
      Z[lp(L - k)] = make_double2(A.x + B.y, B.x - A.y);
    }
    __syncthreads();
  }
  double2 *seq = smem + p * LS;
  double2 v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<L>(t, i))];
  __syncthreads();
  reg_fft<L, 1, JX_ROWS_TQ>(v, t, seq, Wt);

  // ---- update; v becomes the new iterate (part 0 in .x, part 1 in .y)
  double acc = 0.;
  if constexpr (C == 6) {
    const int r = p / 3, c0 = 2 * (p - r * 3);
    if (r < rb) {
      double2 *yr = (double2 *)(y + (row0 + r) * L * 6 + c0);
      const double2 *hr =
          MODE == 2 ? (const double2 *)(ls.rhs + (row0 + r) * L * 6 + c0)
                    : nullptr;
      const double e0 = ls.E[c0], e1 = ls.E[c0 + 1];
#pragma unroll
      for (int i = 0; i < V; ++i) {
        int64_t off = (int64_t)reg_out_pos<L>(t, i) * 3;
        double2 nv = make_double2(ls_new(e0, v[i].x, scale),
                                  ls_new(e1, v[i].y, scale));
        if constexpr (MODE == 2) {
          double2 h = hr[off];
          nv = make_double2(ls_new_rhs(e0, h.x, v[i].x, scale),
                            ls_new_rhs(e1, h.y, v[i].y, scale));
        }
        double2 o = yr[off];
        double dx = nv.x - o.x, dy = nv.y - o.y;
        acc += dx * dx;
        acc += dy * dy;
        yr[off] = nv;
        v[i] = nv;
      }
    } else {
#pragma unroll
      for (int i = 0; i < V; ++i) v[i] = make_double2(0., 0.);
    }
  } else {
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const int q = 2 * p + part;
      const int r = q / C, c = q - (q / C) * C;
      double *yr = y + (row0 + r) * L * C + c;
      const double *hr = MODE == 2 ? ls.rhs + (row0 + r) * L * C + c : nullptr;
#pragma unroll
      for (int i = 0; i < V; ++i) {
        double nv = 0.;
        if (r < rb) {
          int64_t off = (int64_t)reg_out_pos<L>(t, i) * C;
          double raw = part ? v[i].y : v[i].x;
          nv = ls_new(ls.E[c], raw, scale);
          if constexpr (MODE == 2) nv = ls_new_rhs(ls.E[c], hr[off], raw, scale);
          double d = nv - yr[off];
          acc += d * d;
          yr[off] = nv;
        }
        if (part) v[i].y = nv; else v[i].x = nv;
      }
    }
  }
  double sum = block_sum(acc, red);
  if (tid == 0) ls.partials[blk] = sum;

  // ---- the new iterate into LDS, local operator voxel by voxel
  __syncthreads();
#pragma unroll
  for (int i = 0; i < V; ++i) seq[lp(reg_out_pos<L>(t, i))] = v[i];
  __syncthreads();
  for (int w = tid; w < rb * L; w += NT) {
    int r = w / L, j = w - r * L;
    int64_t vox = (row0 + r) * L + j;
    double e[C], tt[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      int q = r * C + c;
      e[c] = ((const double *)&smem[(q >> 1) * LS + lp(j)])[q & 1];
    }
    local_voxel<C>(loc, vox, e, tt);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      int q = r * C + c;
      ((double *)&smem[(q >> 1) * LS + lp(j)])[q & 1] = tt[c];
    }
  }
  __syncthreads();

  // ---- forward transform of the polarization rows (see k_rows_forward_reg)
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<L>(t, i))];
  __syncthreads();
  reg_fft<L, -1, JX_ROWS_TQ>(v, t, seq, Wt);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < V; ++i) seq[lp(reg_out_pos<L>(t, i))] = v[i];
  __syncthreads();
  if constexpr (C == 6) {
    for (int w = tid; w < rb * NH * 3; w += NT) {
      int cp = w % 3;
      int rest = w / 3;
      int k = rest % NH;
      int r = rest / NH;
      const double2 *Z = smem + (r * 3 + cp) * LS;
      double2 zk = Z[lp(k)];
      double2 zm = Z[lp(k == 0 ? 0 : L - k)];
      double2 xa = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      double2 xb = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
      double *sp = spec + ((row0 + r) * NH + k) * 12 + 4 * cp;
      *(double2 *)sp = xa;
      *(double2 *)(sp + 2) = xb;
    }
  } else {
    for (int w = tid; w < rb * NH * C; w += NT) {
      int c = w % C;
      int rest = w / C;
      int k = rest % NH;
      int r = rest / NH;
      int q = r * C + c;
      const double2 *Z = smem + (q >> 1) * LS;
      double2 zk = Z[lp(k)];
      double2 zm = Z[lp(k == 0 ? 0 : L - k)];
      double re, im;
      if (!(q & 1)) {
        re = 0.5 * (zk.x + zm.x);
        im = 0.5 * (zk.y - zm.y);
      } else {
        re = 0.5 * (zk.y + zm.y);
        im = -0.5 * (zk.x - zm.x);
      }
      int64_t pt = (row0 + r) * NH + k;
      spec[pt * 2 * C + 2 * c] = re;
      spec[pt * 2 * C + 2 * c + 1] = im;
    }
  }
}

// ------------------------------------------------------------ residual
constexpr int kFinalizeThreads = 1024;

__global__ void __launch_bounds__(kFinalizeThreads)
    k_finalize(const double *partials, int nb, SolveState *st,
               double *res_hist) {
  if (solve_done(st)) return;
  __shared__ double red[kFinalizeThreads / 64];
  // eight independent accumulators keep eight loads in flight per thread;
  // the summation order is fixed, so the result is deterministic
  double acc[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  const int stride = kFinalizeThreads * 8;
  for (int i0 = 0; i0 < nb; i0 += stride) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      int i = i0 + u * kFinalizeThreads + threadIdx.x;
      if (i < nb) acc[u] += partials[i];
    }
  }
  double sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) +
               ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  double total = block_sum(sum, red);
  if (threadIdx.x == 0) {
    double res = sqrt(total) * st->res_scale;
    int it = st->iter;
    res_hist[it] = res;
    st->iter = it + 1;
    if (!isfinite(res))
      st->done = JX_SOLVE_NONFINITE;
    else if (res < st->tol)
      st->done = JX_SOLVE_CONVERGED;
    else if (it + 1 >= st->max_iter)
      st->done = JX_SOLVE_MAXITER;
  }
}

// ------------------------------------------------------------ local apply
template <int C>
__global__ void __launch_bounds__(kThreads)
    k_local_apply(const double *x, double *y, int64_t N, LocalDev loc) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double e[C], t[C];
#pragma unroll
    for (int c = 0; c < C; ++c) e[c] = x[v * C + c];
    local_voxel<C>(loc, v, e, t);
#pragma unroll
    for (int c = 0; c < C; ++c) y[v * C + c] = t[c];
  }
}

// Adjoint contraction of the implicit gradient (no janus counterpart): for
// dC = sph J + dev K at each voxel, dL/dsph = -(J eps) . mu and
// dL/ddev = -(K eps) . mu = -(eps . mu - (J eps) . mu), with
// (J eps) . mu = tr(eps) tr(mu) / dim in Mandel notation.
template <int C>
__global__ void __launch_bounds__(kThreads)
    k_adjoint_contract(const double *eps, const double *mu, int64_t N,
                       double *dsph, double *ddev) {
  constexpr int D = C == 3 ? 2 : 3;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double e[C], m[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      e[c] = eps[v * C + c];
      m[c] = mu[v * C + c];
    }
    double tre = 0., trm = 0., dot = 0.;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      tre += e[c];
      trm += m[c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) dot += e[c] * m[c];
    double j = tre * trm / (double)D;
    dsph[v] = -j;
    ddev[v] = -(dot - j);
  }
}

// per-block sums of local : x (columns 0..C-1) and x (columns C..2C-1)
template <int C>
__global__ void __launch_bounds__(kThreads)
    k_local_sums(const double *x, int64_t N, LocalDev loc, double *partials) {
  __shared__ double red[kThreads / 64];
  double acc[2 * C];
#pragma unroll
  for (int c = 0; c < 2 * C; ++c) acc[c] = 0.;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double e[C], t[C];
#pragma unroll
    for (int c = 0; c < C; ++c) e[c] = x[v * C + c];
    local_voxel<C>(loc, v, e, t);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      acc[c] += t[c];
      acc[C + c] += e[c];
    }
  }
  for (int c = 0; c < 2 * C; ++c) {
    double s = block_sum(acc[c], red);
    if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * 2 * C + c] = s;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kThreads)
    k_reduce_columns(const double *partials, int nb, int ncol, double *out) {
  __shared__ double red[kThreads / 64];
  for (int c = 0; c < ncol; ++c) {
    double acc = 0.;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
      acc += partials[(int64_t)i * ncol + c];
    double s = block_sum(acc, red);
    if (threadIdx.x == 0) out[c] = s;
    __syncthreads();
  }
}

// ------------------------------------------------------------ stream copy
// Plain 16-byte-per-lane copy: the reference kernel for the practical HBM
// bandwidth and for calibrating traffic counters.
__global__ void __launch_bounds__(kThreads)
    k_stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                  int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int launch_stream_copy(const void *src, void *dst, int64_t n16,
                       hipStream_t s) {
  int64_t nb = std::min<int64_t>((n16 + kThreads - 1) / kThreads, 1 << 20);
  if (nb < 1) return JX_OK;
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)nb), dim3(kThreads), 0, s,
                     (const double2 *)src, (double2 *)dst, n16);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

// ============================================================ launchers

static int check_lds(size_t bytes, const char *what) {
  if ((long long)bytes > lds_limit_bytes())
    return fail(JX_EUNSUPPORTED,
                std::string(what) + ": transform length too large for LDS (" +
                    std::to_string(bytes) + " bytes needed)");
  return JX_OK;
}

template <typename K>
static int set_lds_attr(K kernel, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute((const void *)kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
  return JX_OK;
}


// Rows per block of the register kernels (0 = not available).
template <int L>
constexpr int reg_np() {
  return 192 / reg_tpp<L>();
}

static int reg_rows_per_block(int L, int C) {
  if (!reg_fft_length(L) || (C != 1 && C != 3 && C != 6)) return 0;
  int np = 0;
  switch (L) {
    case 16: np = reg_np<16>(); break;
    case 32: np = reg_np<32>(); break;
    case 64: np = reg_np<64>(); break;
    case 128: np = reg_np<128>(); break;
    case 256: np = reg_np<256>(); break;
    case 512: np = reg_np<512>(); break;
    case 80: np = reg_np<80>(); break;
    case 96: np = reg_np<96>(); break;
    case 160: np = reg_np<160>(); break;
  }
  return 2 * np / C;
}

template <int L, int C>
static int launch_rows_forward_reg_lc(const Grid &g, const double *x,
                                      double *spec, const LocalDev &loc,
                                      const SolveState *st, const double2 *W,
                                      hipStream_t s) {
  constexpr int NP = reg_np<L>();
  constexpr int RB = 2 * NP / C;
  size_t lds = (size_t)NP * reg_lds_stride(L) * sizeof(double2);
  int rc = check_lds(lds, "rows_forward");
  if (rc) return rc;
  dim3 grid((unsigned)((g.rows + RB - 1) / RB)), block(NP * reg_tpp<L>());
  if (loc.mode != LOCAL_NONE) {
    if constexpr (C == 3 || C == 6) {
      set_lds_attr(k_rows_forward_reg<L, C, NP, true>, lds);
      hipLaunchKernelGGL((k_rows_forward_reg<L, C, NP, true>), grid, block, lds,
                         s, x, spec, g.rows, W, loc, st);
    } else {
      return fail(JX_EINVAL, "local operator needs sym components");
    }
  } else {
    set_lds_attr(k_rows_forward_reg<L, C, NP, false>, lds);
    hipLaunchKernelGGL((k_rows_forward_reg<L, C, NP, false>), grid, block, lds,
                       s, x, spec, g.rows, W, loc, st);
  }
  JX_HIP(hipGetLastError());
  return JX_OK;
}

template <int L, int C>
static int launch_rows_inverse_reg_lc(const Grid &g, const double *spec,
                                      double *y, const LsArgs &la, bool upd,
                                      const SolveState *st, const double2 *W,
                                      hipStream_t s) {
  constexpr int NP = reg_np<L>();
  constexpr int RB = 2 * NP / C;
  size_t lds = (size_t)NP * reg_lds_stride(L) * sizeof(double2);
  int rc = check_lds(lds, "rows_inverse");
  if (rc) return rc;
  dim3 grid((unsigned)((g.rows + RB - 1) / RB)), block(NP * reg_tpp<L>());
  double scale = 1.0 / (double)g.N;
  if (upd) {
    if constexpr (C == 3 || C == 6) {
      if (la.rhs) {
        set_lds_attr(k_rows_inverse_reg<L, C, NP, 2>, lds);
        hipLaunchKernelGGL((k_rows_inverse_reg<L, C, NP, 2>), grid, block,
                           lds, s, spec, y, g.rows, scale, W, la, st);
      } else {
        set_lds_attr(k_rows_inverse_reg<L, C, NP, 1>, lds);
        hipLaunchKernelGGL((k_rows_inverse_reg<L, C, NP, 1>), grid, block,
                           lds, s, spec, y, g.rows, scale, W, la, st);
      }
    } else {
      return fail(JX_EINVAL, "fixed point needs sym components");
    }
  } else {
    set_lds_attr(k_rows_inverse_reg<L, C, NP, 0>, lds);
    hipLaunchKernelGGL((k_rows_inverse_reg<L, C, NP, 0>), grid, block, lds,
                       s, spec, y, g.rows, scale, W, la, st);
  }
  JX_HIP(hipGetLastError());
  return JX_OK;
}

#define JX_DISPATCH_LC(FN, ...)                                            \
  switch (L) {                                                             \
    case 16: JX_DISPATCH_C(16, FN, __VA_ARGS__);                           \
    case 32: JX_DISPATCH_C(32, FN, __VA_ARGS__);                           \
    case 64: JX_DISPATCH_C(64, FN, __VA_ARGS__);                           \
    case 128: JX_DISPATCH_C(128, FN, __VA_ARGS__);                         \
    case 256: JX_DISPATCH_C(256, FN, __VA_ARGS__);                         \
    case 512: JX_DISPATCH_C(512, FN, __VA_ARGS__);                         \
    case 80: JX_DISPATCH_C(80, FN, __VA_ARGS__);                           \
    case 96: JX_DISPATCH_C(96, FN, __VA_ARGS__);                           \
    case 160: JX_DISPATCH_C(160, FN, __VA_ARGS__);                         \
    default: return fail(JX_EUNSUPPORTED, "no register FFT for length");   \
  }
#define JX_DISPATCH_C(LL, FN, ...)                                         \
  switch (C) {                                                             \
    case 1: return FN<LL, 1>(__VA_ARGS__);                                 \
    case 3: return FN<LL, 3>(__VA_ARGS__);                                 \
    case 6: return FN<LL, 6>(__VA_ARGS__);                                 \
    default: return fail(JX_EINVAL, "ncomp must be 1, 3 or 6");            \
  }

static int rows_forward_reg(const Grid &g, int C, const double *x, double *spec,
                            const LocalDev &loc, const SolveState *st,
                            const double2 *W, hipStream_t s) {
  const int L = (int)g.L;
  JX_DISPATCH_LC(launch_rows_forward_reg_lc, g, x, spec, loc, st, W, s)
}

static int rows_inverse_reg(const Grid &g, int C, const double *spec, double *y,
                            const LsArgs &la, bool upd, const SolveState *st,
                            const double2 *W, hipStream_t s) {
  const int L = (int)g.L;
  JX_DISPATCH_LC(launch_rows_inverse_reg_lc, g, spec, y, la, upd, st, W, s)
}

template <int L, int C>
static int launch_rows_update_forward_reg_lc(const Grid &g, double *spec,
                                             double *y, const LsArgs &la,
                                             const LocalDev &loc,
                                             const SolveState *st,
                                             const double2 *W, hipStream_t s) {
  if constexpr (C == 3 || C == 6) {
    constexpr int NP = reg_np<L>();
    constexpr int RB = 2 * NP / C;
    size_t lds = (size_t)NP * reg_lds_stride(L) * sizeof(double2);
    int rc = check_lds(lds, "rows_update_forward");
    if (rc) return rc;
    dim3 grid((unsigned)((g.rows + RB - 1) / RB)), block(NP * reg_tpp<L>());
    double scale = 1.0 / (double)g.N;
    if (la.rhs) {
      set_lds_attr(k_rows_update_forward_reg<L, C, NP, 2>, lds);
      hipLaunchKernelGGL((k_rows_update_forward_reg<L, C, NP, 2>), grid, block,
                         lds, s, spec, y, g.rows, scale, W, la, loc, st);
    } else {
      set_lds_attr(k_rows_update_forward_reg<L, C, NP, 1>, lds);
      hipLaunchKernelGGL((k_rows_update_forward_reg<L, C, NP, 1>), grid, block,
                         lds, s, spec, y, g.rows, scale, W, la, loc, st);
    }
    JX_HIP(hipGetLastError());
    return JX_OK;
  } else {
    return fail(JX_EINVAL, "fixed point needs sym components");
  }
}

static int rows_update_forward_reg(const Grid &g, int C, double *spec,
                                   double *y, const LsArgs &la,
                                   const LocalDev &loc, const SolveState *st,
                                   const double2 *W, hipStream_t s) {
  const int L = (int)g.L;
  JX_DISPATCH_LC(launch_rows_update_forward_reg_lc, g, spec, y, la, loc, st, W, s)
}

static int rows_per_block(int L, int C) {
  int Ls = lds_stride(L);
  // aim at about 1536 complex entries per ping-pong buffer
  int nseq = std::max(1, 1536 / Ls);
  int rb = std::max(1, (2 * nseq) / C);
  return rb;
}

static size_t rows_lds_bytes(int L, int C, int RB) {
  int nseq_max = (RB * C + 1) / 2;
  return 2 * (size_t)nseq_max * lds_stride(L) * sizeof(double2);
}

int launch_rows_forward(const Grid &g, int ncomp, const double *x, double *spec,
                        const LocalDev &loc, const SolveState *st,
                        hipStream_t s) {
  FftLen P;
  int rc = get_fft_len((int)g.L, &P);
  if (rc) return rc;
  if (reg_rows_per_block((int)g.L, ncomp))
    return rows_forward_reg(g, ncomp, x, spec, loc, st, P.W, s);
  RowsArgs a;
  a.nrows = g.rows;
  a.L = (int)g.L;
  a.nh = (int)g.nh;
  a.Ls = lds_stride(a.L);
  a.RB = rows_per_block(a.L, ncomp);
  a.scale = 1.0;
  size_t lds = rows_lds_bytes(a.L, ncomp, a.RB);
  if ((rc = check_lds(lds, "rows_forward"))) return rc;
  int64_t nb = (g.rows + a.RB - 1) / a.RB;
  dim3 grid((unsigned)nb), block(kThreads);
  switch (ncomp) {
    case 1:
      set_lds_attr(k_rows_forward<1>, lds);
      hipLaunchKernelGGL(k_rows_forward<1>, grid, block, lds, s, x, spec, a, P,
                         loc, st);
      break;
    case 3:
      set_lds_attr(k_rows_forward<3>, lds);
      hipLaunchKernelGGL(k_rows_forward<3>, grid, block, lds, s, x, spec, a, P,
                         loc, st);
      break;
    case 6:
      set_lds_attr(k_rows_forward<6>, lds);
      hipLaunchKernelGGL(k_rows_forward<6>, grid, block, lds, s, x, spec, a, P,
                         loc, st);
      break;
    default:
      return fail(JX_EINVAL, "rows_forward: ncomp must be 1, 3 or 6");
  }
  JX_HIP(hipGetLastError());
  return JX_OK;
}

template <int C>
static void launch_axis_c(int dir, dim3 grid, size_t lds, hipStream_t s,
                          const double *in, double *out, const AxisArgs &a,
                          const FftLen &P, const GreenDev &G,
                          const SolveState *st) {
  if (dir < 0) {
    set_lds_attr(k_axis<C, -1>, lds);
    hipLaunchKernelGGL((k_axis<C, -1>), grid, dim3(kThreads), lds, s, in, out,
                       a, P, G, st);
  } else if (dir > 0) {
    set_lds_attr(k_axis<C, 1>, lds);
    hipLaunchKernelGGL((k_axis<C, 1>), grid, dim3(kThreads), lds, s, in, out,
                       a, P, G, st);
  } else {
    set_lds_attr(k_axis<C, 0>, lds);
    hipLaunchKernelGGL((k_axis<C, 0>), grid, dim3(kThreads), lds, s, in, out,
                       a, P, G, st);
  }
}

template <int L, int C, int NSEQ>
static void launch_axis_reg_lc(int dir, dim3 grid, size_t lds, hipStream_t s,
                               const double *in, double *out,
                               const AxisArgs &a, const double2 *W,
                               const GreenDev &G, const SolveState *st) {
  dim3 block(NSEQ * reg_tpp<L>());
  if (dir < 0) {
    set_lds_attr(k_axis_reg<L, C, NSEQ, -1>, lds);
    hipLaunchKernelGGL((k_axis_reg<L, C, NSEQ, -1>), grid, block, lds, s, in,
                       out, a, W, G, st);
  } else if (dir > 0) {
    set_lds_attr(k_axis_reg<L, C, NSEQ, 1>, lds);
    hipLaunchKernelGGL((k_axis_reg<L, C, NSEQ, 1>), grid, block, lds, s, in,
                       out, a, W, G, st);
  } else {
    set_lds_attr(k_axis_reg<L, C, NSEQ, 0>, lds);
    hipLaunchKernelGGL((k_axis_reg<L, C, NSEQ, 0>), grid, block, lds, s, in,
                       out, a, W, G, st);
  }
}

// sequences per block: 192 threads for every length up to 512
template <int L>
constexpr int reg_nseq() {
  return 192 / reg_tpp<L>();
}

template <int L>
static int launch_axis_reg_l(const Grid &g, int ncomp, int axis, int dir,
                             const double *in, double *out,
                             const GreenDev *green, const SolveState *st,
                             const FftLen &P, hipStream_t s) {
  constexpr int NSEQ = reg_nseq<L>();
  static_assert(NSEQ % 6 == 0, "sequences per block must hold whole points");
  AxisArgs a;
  a.L = L;
  a.Ls = reg_lds_stride(L);
  a.nh = (int)g.nh;
  int64_t inner = g.nh;
  for (int d = axis + 1; d < g.dim - 1; ++d) inner *= g.n[d];
  int64_t outer = 1;
  for (int d = 0; d < axis; ++d) outer *= g.n[d];
  a.n_inner = inner;
  a.n_outer = outer;
  a.stride_ax = inner * 2 * ncomp;
  a.stride_outer = (int64_t)L * inner * 2 * ncomp;
  const int T = NSEQ / ncomp;
  a.T = T;
  a.ntiles = (int)((inner + T - 1) / T);
  size_t lds = (size_t)NSEQ * reg_lds_stride(L) * sizeof(double2);
  int rc = check_lds(lds, "axis");
  if (rc) return rc;
  dim3 grid((unsigned)(outer * a.ntiles));
  GreenDev G = green ? *green : GreenDev();
  switch (ncomp) {
    case 1: launch_axis_reg_lc<L, 1, NSEQ>(dir, grid, lds, s, in, out, a, P.W, G, st); break;
    case 3: launch_axis_reg_lc<L, 3, NSEQ>(dir, grid, lds, s, in, out, a, P.W, G, st); break;
    case 6: launch_axis_reg_lc<L, 6, NSEQ>(dir, grid, lds, s, in, out, a, P.W, G, st); break;
    default: return fail(JX_EINVAL, "launch_axis: ncomp must be 1, 3 or 6");
  }
  JX_HIP(hipGetLastError());
  return JX_OK;
}

int launch_axis_reg(const Grid &g, int ncomp, int axis, int dir,
                    const double *in, double *out, const GreenDev *green,
                    const SolveState *st, const FftLen &P, hipStream_t s) {
  switch (P.L) {
    case 16: return launch_axis_reg_l<16>(g, ncomp, axis, dir, in, out, green, st, P, s);
    case 32: return launch_axis_reg_l<32>(g, ncomp, axis, dir, in, out, green, st, P, s);
    case 64: return launch_axis_reg_l<64>(g, ncomp, axis, dir, in, out, green, st, P, s);
    case 128: return launch_axis_reg_l<128>(g, ncomp, axis, dir, in, out, green, st, P, s);
    case 256: return launch_axis_reg_l<256>(g, ncomp, axis, dir, in, out, green, st, P, s);
    case 512: return launch_axis_reg_l<512>(g, ncomp, axis, dir, in, out, green, st, P, s);
    case 80: return launch_axis_reg_l<80>(g, ncomp, axis, dir, in, out, green, st, P, s);
    case 96: return launch_axis_reg_l<96>(g, ncomp, axis, dir, in, out, green, st, P, s);
    case 160: return launch_axis_reg_l<160>(g, ncomp, axis, dir, in, out, green, st, P, s);
    default: return fail(JX_EUNSUPPORTED, "no register FFT for this length");
  }
}

int launch_axis(const Grid &g, int ncomp, int axis, int dir, const double *in,
                double *out, const GreenDev *green, const SolveState *st,
                hipStream_t s) {
  if (axis < 0 || axis >= g.dim - 1)
    return fail(JX_EINVAL, "launch_axis: axis out of range");
  if (dir == 0 && (axis != 0 || green == nullptr || ncomp == 1))
    return fail(JX_EINVAL, "launch_axis: Gamma-hat pass needs axis 0");
  FftLen P;
  int rc = get_fft_len((int)g.n[axis], &P);
  if (rc) return rc;
  if (reg_fft_length((int)g.n[axis]))
    return launch_axis_reg(g, ncomp, axis, dir, in, out, green, st, P, s);
  AxisArgs a;
  a.L = (int)g.n[axis];
  a.Ls = lds_stride(a.L);
  a.nh = (int)g.nh;
  // points after `axis` (all inner axes, last one halved) are contiguous
  int64_t inner = g.nh;
  for (int d = axis + 1; d < g.dim - 1; ++d) inner *= g.n[d];
  int64_t outer = 1;
  for (int d = 0; d < axis; ++d) outer *= g.n[d];
  a.n_inner = inner;
  a.n_outer = outer;
  a.stride_ax = inner * 2 * ncomp;
  a.stride_outer = (int64_t)a.L * inner * 2 * ncomp;
  int T = std::max(1, 1536 / (a.Ls * ncomp));
  if (T > inner) T = (int)inner;
  a.T = T;
  a.ntiles = (int)((inner + T - 1) / T);
  size_t lds = 2 * (size_t)T * ncomp * a.Ls * sizeof(double2);
  if ((rc = check_lds(lds, "axis"))) return rc;
  int64_t nb = outer * a.ntiles;
  dim3 grid((unsigned)nb);
  GreenDev G = green ? *green : GreenDev();
  switch (ncomp) {
    case 1: launch_axis_c<1>(dir, grid, lds, s, in, out, a, P, G, st); break;
    case 3: launch_axis_c<3>(dir, grid, lds, s, in, out, a, P, G, st); break;
    case 6: launch_axis_c<6>(dir, grid, lds, s, in, out, a, P, G, st); break;
    default: return fail(JX_EINVAL, "launch_axis: ncomp must be 1, 3 or 6");
  }
  JX_HIP(hipGetLastError());
  return JX_OK;
}

int rows_inverse_nblocks(const Grid &g, int ncomp) {
  int RB = reg_rows_per_block((int)g.L, ncomp);
  if (!RB) RB = rows_per_block((int)g.L, ncomp);
  return (int)((g.rows + RB - 1) / RB);
}

int launch_rows_inverse(const Grid &g, int ncomp, const double *spec, double *y,
                        const RowsInverseLs &ls, const SolveState *st,
                        int *nblocks_out, hipStream_t s) {
  FftLen P;
  int rc = get_fft_len((int)g.L, &P);
  if (rc) return rc;
  if (int RBr = reg_rows_per_block((int)g.L, ncomp)) {
    LsArgs la;
    la.rhs = ls.rhs;
    la.partials = ls.partials;
    for (int c = 0; c < 6; ++c) la.E[c] = ls.E[c];
    if (nblocks_out) *nblocks_out = (int)((g.rows + RBr - 1) / RBr);
    return rows_inverse_reg(g, ncomp, spec, y, la, ls.enabled, st, P.W, s);
  }
  RowsArgs a;
  a.nrows = g.rows;
  a.L = (int)g.L;
  a.nh = (int)g.nh;
  a.Ls = lds_stride(a.L);
  a.RB = rows_per_block(a.L, ncomp);
  a.scale = 1.0 / (double)g.N;
  size_t lds = rows_lds_bytes(a.L, ncomp, a.RB);
  if ((rc = check_lds(lds, "rows_inverse"))) return rc;
  int64_t nb = (g.rows + a.RB - 1) / a.RB;
  if (nblocks_out) *nblocks_out = (int)nb;
  LsArgs la;
  la.rhs = ls.rhs;
  la.partials = ls.partials;
  for (int c = 0; c < 6; ++c) la.E[c] = ls.E[c];
  dim3 grid((unsigned)nb), block(kThreads);
#define JX_RI(CC, LSV)                                                        \
  do {                                                                        \
    set_lds_attr(k_rows_inverse<CC, LSV>, lds);                               \
    hipLaunchKernelGGL((k_rows_inverse<CC, LSV>), grid, block, lds, s, spec, \
                       y, a, P, la, st);                                      \
  } while (0)
  if (ls.enabled) {
    if (ncomp == 3) JX_RI(3, true);
    else if (ncomp == 6) JX_RI(6, true);
    else return fail(JX_EINVAL, "rows_inverse: fixed point needs sym comps");
  } else {
    if (ncomp == 1) JX_RI(1, false);
    else if (ncomp == 3) JX_RI(3, false);
    else if (ncomp == 6) JX_RI(6, false);
    else return fail(JX_EINVAL, "rows_inverse: ncomp must be 1, 3 or 6");
  }
#undef JX_RI
  JX_HIP(hipGetLastError());
  return JX_OK;
}

bool rows_fusable(const Grid &g, int ncomp) {
  return (ncomp == 3 || ncomp == 6) && reg_rows_per_block((int)g.L, ncomp) > 0;
}

int launch_rows_update_forward(const Grid &g, int ncomp, double *spec,
                               double *y, const RowsInverseLs &ls,
                               const LocalDev &loc, const SolveState *st,
                               int *nblocks_out, hipStream_t s) {
  if (!rows_fusable(g, ncomp))
    return fail(JX_EUNSUPPORTED, "fused update pass needs a register FFT length");
  FftLen P;
  int rc = get_fft_len((int)g.L, &P);
  if (rc) return rc;
  LsArgs la;
  la.rhs = ls.rhs;
  la.partials = ls.partials;
  for (int c = 0; c < 6; ++c) la.E[c] = ls.E[c];
  int RB = reg_rows_per_block((int)g.L, ncomp);
  if (nblocks_out) *nblocks_out = (int)((g.rows + RB - 1) / RB);
  return rows_update_forward_reg(g, ncomp, spec, y, la, loc, st, P.W, s);
}

int launch_finalize(const double *partials, int nblocks, SolveState *st,
                    double *res_hist, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinalizeThreads), 0, s,
                     partials, nblocks, st, res_hist);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

int launch_local_apply(const Grid &g, int ncomp, const LocalDev &loc,
                       const double *x, double *y, hipStream_t s) {
  int64_t nb = std::min<int64_t>((g.N + kThreads - 1) / kThreads, 65536);
  if (ncomp == 3)
    hipLaunchKernelGGL(k_local_apply<3>, dim3((unsigned)nb), dim3(kThreads), 0,
                       s, x, y, g.N, loc);
  else if (ncomp == 6)
    hipLaunchKernelGGL(k_local_apply<6>, dim3((unsigned)nb), dim3(kThreads), 0,
                       s, x, y, g.N, loc);
  else
    return fail(JX_EINVAL, "local_apply: ncomp must be 3 or 6");
  JX_HIP(hipGetLastError());
  return JX_OK;
}

int launch_adjoint_contract(const Grid &g, int ncomp, const double *eps,
                            const double *mu, double *dsph, double *ddev,
                            hipStream_t s) {
  int64_t nb = std::min<int64_t>((g.N + kThreads - 1) / kThreads, 65536);
  if (ncomp == 3)
    hipLaunchKernelGGL(k_adjoint_contract<3>, dim3((unsigned)nb),
                       dim3(kThreads), 0, s, eps, mu, g.N, dsph, ddev);
  else if (ncomp == 6)
    hipLaunchKernelGGL(k_adjoint_contract<6>, dim3((unsigned)nb),
                       dim3(kThreads), 0, s, eps, mu, g.N, dsph, ddev);
  else
    return fail(JX_EINVAL, "adjoint_contract: ncomp must be 3 or 6");
  JX_HIP(hipGetLastError());
  return JX_OK;
}

int launch_local_sums(const Grid &g, int ncomp, const LocalDev &loc,
                      const double *x, double *partials, int *nblocks,
                      hipStream_t s) {
  int nb = (int)std::min<int64_t>((g.N + kThreads - 1) / kThreads, 1024);
  *nblocks = nb;
  if (ncomp == 3)
    hipLaunchKernelGGL(k_local_sums<3>, dim3(nb), dim3(kThreads), 0, s, x, g.N,
                       loc, partials);
  else if (ncomp == 6)
    hipLaunchKernelGGL(k_local_sums<6>, dim3(nb), dim3(kThreads), 0, s, x, g.N,
                       loc, partials);
  else
    return fail(JX_EINVAL, "local_sums: ncomp must be 3 or 6");
  JX_HIP(hipGetLastError());
  return JX_OK;
}

int launch_reduce_columns(const double *partials, int nblocks, int ncol,
                          double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_columns, dim3(1), dim3(kThreads), 0, s, partials,
                     nblocks, ncol, out);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

}  // namespace jx
