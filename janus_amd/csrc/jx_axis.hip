// Passes B, C and D of the Lippmann-Schwinger iteration: complex transforms
// along the strided axes of the spectrum, with Gamma-hat applied between the
// forward and inverse transforms of axis 0 (pass C).
//
// Spectrum layout (rows, nh, C, 2) as in pass A (jx_rows_fwd.hip).

#include "jx_device.h"

namespace jx {

// ------------------------------------------------------------ passes B, C, D
struct AxisArgs {
  int L, Ls, T, ntiles;
  int64_t stride_ax, n_outer, stride_outer, n_inner;
  int nh;  // last-axis half length, decodes the inner index in 3D pass C
  int axis;  // the transformed axis (0, or 1 for a 3D Gamma-hat pass on axis 1)
};

// Full frequency index of the Gamma-hat pass from the outer index o, the
// transformed-axis index e and the inner index m: axis 0 -- b = (e, m) in 2D,
// (e, m / nh, m mod nh) in 3D; axis 1 (3D) -- b = (o, e, m).
template <int DIM>
__device__ __forceinline__ void gamma_index(int axis, int64_t o, int e,
                                            int64_t m, int nh, int *b) {
  if (DIM == 2) {
    b[0] = e;
    b[1] = (int)m;
  } else if (axis == 0) {
    b[0] = e;
    b[1] = (int)(m / nh);
    b[2] = (int)(m - (int64_t)b[1] * nh);
  } else {
    b[0] = (int)o;
    b[1] = e;
    b[2] = (int)m;
  }
}

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
        gamma_index<DIM>(a.axis, o, e, m, a.nh, b);
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

// z <- D Gamma-hat D z for the C complex values z[c * zs] (real and
// imaginary parts separately), D = diag(-1 on the Mandel components whose
// bit is set in FLIP).  Negation is exact.
// Each output row is stored as soon as it is summed (the inputs are all in
// registers by then), which keeps the live set to g and the inputs.
template <int C, int FLIP>
__device__ __forceinline__ void gamma_apply_flip(const double *g, double2 *z,
                                                 int zs) {
  double xr[C], xi[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    double2 v = z[c * zs];
    const bool f = (FLIP >> c) & 1;
    xr[c] = f ? -v.x : v.x;
    xi[c] = f ? -v.y : v.y;
  }
#pragma unroll
  for (int i = 0; i < C; ++i) {
    double yr = 0., yi = 0.;
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const double gij = g[tri_index(C, i < j ? i : j, i < j ? j : i)];
      yr += gij * xr[j];
      yi += gij * xi[j];
    }
    const bool f = (FLIP >> i) & 1;
    z[i * zs] = f ? make_double2(-yr, -yi) : make_double2(yr, yi);
  }
}

// Sign pattern of Gamma-hat(s k) = D Gamma-hat(k) D, D_(ij) = s_i s_j over
// the Mandel components (3D: xx yy zz yz zx xy; 2D: xx yy xy): for k0
// negated zx and xy flip (2D: xy); for k1 negated yz and xy flip.
constexpr int kFlip0_3 = 0x30;
constexpr int kFlip0_2 = 0x4;
constexpr int kFlip1_3 = 0x28;

// Gamma-hat at frequencies e and L - e of the transformed axis (outer
// index o, inner index m), evaluated once.  The truncated wave vector and
// the filtered alias set of L - e are those of e with that axis's component
// negated (and the filtered weights swapped), so Gamma-hat(L - e) =
// S Gamma-hat(e) S with S = -1 on the Mandel components holding the axis
// once (axis 0: zx, xy in 3D, xy in 2D; axis 1: yz, xy) and +1 elsewhere
// (Willot's wave vector negates the other two components instead, the same
// Gamma0).  This holds for every e except 0 and L / 2 (even L), which are
// their own mirrors; the caller passes z2 = nullptr for those.  Applied to
// the C complex values at z0[c * zs] and, when z2 != nullptr, z2[c * zs].
template <int C>
__device__ __forceinline__ void gamma_point_pair(const GreenDev &G, int axis,
                                                 int64_t o, int e, int64_t m,
                                                 int nh, double2 *z0,
                                                 double2 *z2, int zs) {
  constexpr int DIM = C == 3 ? 2 : 3;
  constexpr int NC = C == 3 ? 6 : 21;
  int b[3];
  gamma_index<DIM>(axis, o, e, m, nh, b);
  double g[NC];
  gamma_hat_moments<DIM>(G, b, g);
  gamma_apply_flip<C, 0>(g, z0, zs);
  if (z2 != nullptr) {
    if (DIM == 2) gamma_apply_flip<C, kFlip0_2>(g, z2, zs);
    else if (axis == 0) gamma_apply_flip<C, kFlip0_3>(g, z2, zs);
    else gamma_apply_flip<C, kFlip1_3>(g, z2, zs);
  }
}

// Register-resident transform along a strided axis of the spectrum for the
// register FFT lengths.  NSEQ = T * C complex sequences per block (T points
// of the contiguous inner index, all C components), TPP = L / V threads
// each.  Thread tid = t * NSEQ + s: consecutive lanes read consecutive
// doubles of the T * 2C doubles that one axis element of the tile spans.
// Batched: the entry's spectrum (in place), Green constants and state come
// from its slot.
//
// Minimum waves per SIMD requested for the Gamma-hat pass of the 16-value
// plans (register budget of the compiler).  3 (a few spilled registers)
// measured faster than the compiler's own choice of 2 waves at 256^3; the
// 20- and 24-value plans would spill heavily and keep the compiler's choice.
#ifndef JX_GAMMA_PASS_MIN_WAVES
#define JX_GAMMA_PASS_MIN_WAVES 3
#endif

// The same for the 6-value 96-point plan (three spilled registers): 0.0374
// ms at 96^3 against 0.0488 ms at the compiler's 2 waves, which makes it
// faster than the 12-value plan's 0.0393 ms (the 10-value plans spill
// 20-36 registers at 3 waves and measured no faster).
#ifndef JX_GAMMA_PASS_V6_MIN_WAVES
#define JX_GAMMA_PASS_V6_MIN_WAVES 3
#endif

template <class P, int C, int NSEQ, int MODE>
__global__ void __launch_bounds__(NSEQ * reg_tpp<P>(),
                                  MODE == 0 && reg_v<P>() == 16
                                      ? JX_GAMMA_PASS_MIN_WAVES
                                  : MODE == 0 && reg_v<P>() == 6
                                      ? JX_GAMMA_PASS_V6_MIN_WAVES
                                      : 1)
    k_axis_reg(const double *in, double *out, AxisArgs a, const double2 *W,
               GreenDev G, const SolveState *st, BatchRef br) {
  constexpr int L = P::L;
  const Slot *sl;
  const int64_t blk = batch_block(br, sl);
  if (sl) {
    in = sl->spec;
    out = sl->spec;
    G.daux1 = sl->daux[0];
    G.daux2 = sl->daux[1];
    G.daux3 = sl->daux[2];
    G.daux4 = sl->daux[3];
    st = sl->st;
  }
  in = global_ptr(in);
  out = global_ptr(out);
  st = global_ptr(st);
  W = global_ptr(W);
  if constexpr (MODE == 0) {
    for (int d = 0; d < kMaxDim; ++d) {
      G.ax[d].kt = global_ptr(G.ax[d].kt);
      G.ax[d].km = global_ptr(G.ax[d].km);
      G.ax[d].kp = global_ptr(G.ax[d].kp);
      G.ax[d].wm = global_ptr(G.ax[d].wm);
      G.ax[d].wp = global_ptr(G.ax[d].wp);
      G.ax[d].sw = global_ptr(G.ax[d].sw);
      G.ax[d].cw = global_ptr(G.ax[d].cw);
    }
  }
  if (solve_done(st)) return;
  constexpr int TPP = reg_tpp<P>();
  constexpr int V = reg_v<P>();
  constexpr int T = NSEQ / C;
  constexpr int LS = reg_lds_stride(L);
  extern __shared__ double2 smem[];
  const int tid = threadIdx.x;
  // sequence-fast lanes: a wave instruction reads whole rows of the tile
  // (thread-fast lanes, free of LDS bank conflicts, measured 10-25% slower
  // in every axis pass)
  const int s = tid % NSEQ, t = tid / NSEQ;
  const int tp = s / C, c = s - tp * C;
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
      v[i] = *(const double2 *)(src + (int64_t)reg_in_pos<P>(t, i) * a.stride_ax);
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
  reg_fft<P, (MODE > 0 ? 1 : -1), TQ>(v, t, seq, Wt);
  if constexpr (MODE == 0 && (C == 3 || C == 6)) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i) seq[lp(reg_out_pos<P>(t, i))] = v[i];
    __syncthreads();
    // frequencies 0 ... L/2 of each point, each with its mirror L - e
    constexpr int LH = L / 2 + 1;
    for (int w = tid; w < tc * LH; w += NSEQ * TPP) {
      int p = w / LH, e = w - p * LH;
      double2 *col = smem + (p * C) * LS;
      double2 *z2 = (e == 0 || 2 * e == L) ? nullptr : col + lp(L - e);
      gamma_point_pair<C>(G, a.axis, o, e, m0 + p, a.nh, col + lp(e), z2, LS);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<P>(t, i))];
    __syncthreads();
    // the inverse transform reads the same twiddles as the forward one; an
    // opaque copy of the table pointer keeps the compiler from holding them
    // in registers across the Gamma-hat step (which spilled)
    const double2 *Wi = Wt;
    if constexpr (TQ) asm volatile("" : "+s"(Wi));
    else Wi = global_ptr(Wt);
    reg_fft<P, 1, TQ>(v, t, seq, Wi);
  }
  if (valid) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      *(double2 *)(dst + (int64_t)reg_out_pos<P>(t, i) * a.stride_ax) = v[i];
    }
  }
}

// ============================================================ launchers

// Strides of the padded 3D spectrum layout (spec_layout): axis 0 steps by
// whole padded slabs, axis 1 passes take one slab per outer index.
static void padded_strides(const Grid &g, int ncomp, int axis, AxisArgs &a) {
  if (g.dim != 3) return;
  const int64_t slab = g.n[1] * g.nh * 2 * ncomp + spec_layout(g, ncomp).pad;
  if (axis == 0) {
    a.stride_ax = slab;
    a.stride_outer = g.n[0] * slab;
  } else {
    a.stride_outer = slab;
  }
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

template <class P, int C, int NSEQ>
static void launch_axis_reg_lc(int dir, dim3 grid, size_t lds, hipStream_t s,
                               const double *in, double *out,
                               const AxisArgs &a, const double2 *W,
                               const GreenDev &G, const SolveState *st,
                               const BatchRef &br) {
  dim3 block(NSEQ * reg_tpp<P>());
  if (dir < 0) {
    set_lds_attr(k_axis_reg<P, C, NSEQ, -1>, lds);
    hipLaunchKernelGGL((k_axis_reg<P, C, NSEQ, -1>), grid, block, lds, s, in,
                       out, a, W, G, st, br);
  } else if (dir > 0) {
    set_lds_attr(k_axis_reg<P, C, NSEQ, 1>, lds);
    hipLaunchKernelGGL((k_axis_reg<P, C, NSEQ, 1>), grid, block, lds, s, in,
                       out, a, W, G, st, br);
  } else {
    set_lds_attr(k_axis_reg<P, C, NSEQ, 0>, lds);
    hipLaunchKernelGGL((k_axis_reg<P, C, NSEQ, 0>), grid, block, lds, s, in,
                       out, a, W, G, st, br);
  }
}

// Passes B and D of the 16-thread-per-sequence plans (256 points, 16
// values) take four points per block (384 threads, 105 KB of LDS, one block
// per CU): a tile then spans 384 bytes, three whole 128-byte lines, and the
// passes measured 0.313 against 0.324 ms at 256^3 with two points per block
// (and 0.349 with one, 0.328 with three).  The Gamma-hat pass keeps two:
// it needs the occupancy more (0.468 against 0.523 ms with four).  The
// 512-point plans keep one point per block: two (one 384-thread block per
// CU) measured slower at 512^3 for every pass.
#ifndef JX_AXIS_BD_POINTS4
#define JX_AXIS_BD_POINTS4 1
#endif

template <class P, int NSEQ>
static int launch_axis_reg_pn(const Grid &g, int ncomp, int axis, int dir,
                              const double *in, double *out,
                              const GreenDev *green, const SolveState *st,
                              const FftLen &FL, hipStream_t s,
                              const Batch *b) {
  constexpr int L = P::L;
  static_assert(NSEQ % 6 == 0, "sequences per block must hold whole points");
  AxisArgs a;
  a.L = L;
  a.Ls = reg_lds_stride(L);
  a.nh = (int)g.nh;
  a.axis = axis;
  int64_t inner = g.nh;
  for (int d = axis + 1; d < g.dim - 1; ++d) inner *= g.n[d];
  int64_t outer = 1;
  for (int d = 0; d < axis; ++d) outer *= g.n[d];
  a.n_inner = inner;
  a.n_outer = outer;
  a.stride_ax = inner * 2 * ncomp;
  a.stride_outer = (int64_t)L * inner * 2 * ncomp;
  padded_strides(g, ncomp, axis, a);
  const int T = NSEQ / ncomp;
  a.T = T;
  a.ntiles = (int)((inner + T - 1) / T);
  size_t lds = (size_t)NSEQ * reg_lds_stride(L) * sizeof(double2);
  int rc = check_lds(lds, "axis");
  if (rc) return rc;
  const int64_t nbe = outer * a.ntiles;
  const BatchRef br = batch_ref(b, nbe);
  dim3 grid((unsigned)(nbe * batch_count(b)));
  GreenDev G = green ? *green : GreenDev();
  switch (ncomp) {
    case 1: launch_axis_reg_lc<P, 1, NSEQ>(dir, grid, lds, s, in, out, a, FL.W, G, st, br); break;
    case 3: launch_axis_reg_lc<P, 3, NSEQ>(dir, grid, lds, s, in, out, a, FL.W, G, st, br); break;
    case 6: launch_axis_reg_lc<P, 6, NSEQ>(dir, grid, lds, s, in, out, a, FL.W, G, st, br); break;
    default: return fail(JX_EINVAL, "launch_axis: ncomp must be 1, 3 or 6");
  }
  JX_HIP(hipGetLastError());
  return JX_OK;
}

template <class P>
static int launch_axis_reg_p(const Grid &g, int ncomp, int axis, int dir,
                             const double *in, double *out,
                             const GreenDev *green, const SolveState *st,
                             const FftLen &FL, hipStream_t s, const Batch *b) {
  if constexpr (reg_tpp<P>() == 16 && reg_nseq<P>() == 12) {
    if (JX_AXIS_BD_POINTS4 && dir != 0)
      return launch_axis_reg_pn<P, 24>(g, ncomp, axis, dir, in, out, green,
                                       st, FL, s, b);
  }
  return launch_axis_reg_pn<P, reg_nseq<P>()>(g, ncomp, axis, dir, in, out,
                                              green, st, FL, s, b);
}

// The Gamma-hat pass and the plain axis passes choose their plan variants
// separately.
static int launch_axis_reg(const Grid &g, int ncomp, int axis, int dir,
                           const double *in, double *out,
                           const GreenDev *green, const SolveState *st,
                           const FftLen &FL, hipStream_t s, const Batch *b) {
  const int k = g.plan[dir == 0 ? PASS_GAMMA : PASS_AXIS];
  return with_plan(FL.L, k, [&](auto plan) -> int {
    return launch_axis_reg_p<decltype(plan)>(g, ncomp, axis, dir, in, out,
                                             green, st, FL, s, b);
  });
}

// ------------------------------------------------- access-pattern probe
// The memory traffic of a strided pass without its transforms: every block
// copies the L rows of one tile of T points (6 complex components each,
// 96 B per point and row) of the solver's padded 3D spectrum layout along
// `axis`, in the passes' tile order (xcd_tile) and lane order (consecutive
// lanes read consecutive 16-byte values of a tile row).  Its rate is what
// the memory system gives that pattern with no arithmetic in the way.
template <int T, int V>
__global__ void __launch_bounds__(6 * T * 16)
    k_tile_probe(const double2 *in, double2 *out,
                 int64_t step, int64_t outer_step, int64_t ntiles,
                 int64_t n_inner) {
  constexpr int NS = 6 * T;
  const int tid = threadIdx.x, s = tid % NS, t = tid / NS;
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t o = blk / ntiles;
  const int64_t m0 = (blk - o * ntiles) * T;
  if (m0 + s / 6 >= n_inner) return;
  const double2 *src = in + o * outer_step + m0 * 6 + s;
  double2 *dst = out + o * outer_step + m0 * 6 + s;
  // every row of the tile is loaded before the first store, as in the passes
  double vx[V], vy[V];
#pragma unroll
  for (int r = 0; r < V; ++r) {
    const double2 z = src[(int64_t)(t + 16 * r) * step];
    vx[r] = z.x;
    vy[r] = z.y;
  }
#pragma unroll
  for (int r = 0; r < V; ++r)
    dst[(int64_t)(t + 16 * r) * step] = make_double2(vx[r], vy[r]);
}

template <int T>
static int launch_tile_probe(int L, dim3 grid, hipStream_t s, const double *in,
                             double *out, const AxisArgs &a) {
  dim3 block(6 * T * 16);
  const double2 *i2 = (const double2 *)in;
  double2 *o2 = (double2 *)out;
  if (L == 256)
    hipLaunchKernelGGL((k_tile_probe<T, 16>), grid, block, 0, s, i2, o2,
                       a.stride_ax / 2, a.stride_outer / 2,
                       (int64_t)a.ntiles, a.n_inner);
  else
    hipLaunchKernelGGL((k_tile_probe<T, 32>), grid, block, 0, s, i2, o2,
                       a.stride_ax / 2, a.stride_outer / 2,
                       (int64_t)a.ntiles, a.n_inner);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

int tile_probe(const Grid &g, int axis, int points, const double *src,
               double *dst, hipStream_t s) {
  if (g.dim != 3) return fail(JX_EINVAL, "jx_tile_probe: the grid must be 3D");
  if (axis != 0 && axis != 1)
    return fail(JX_EINVAL, "jx_tile_probe: axis must be 0 or 1");
  const int L = (int)g.n[axis];
  if (L != 256 && L != 512)
    return fail(JX_EINVAL, "jx_tile_probe: the probed axis must have 256 or 512 points");
  if (points != 1 && points != 2 && points != 4)
    return fail(JX_EINVAL, "jx_tile_probe: points must be 1, 2 or 4");
  AxisArgs a;
  a.L = L;
  a.Ls = 0;
  a.nh = (int)g.nh;
  a.axis = axis;
  int64_t inner = g.nh;
  for (int d = axis + 1; d < g.dim - 1; ++d) inner *= g.n[d];
  int64_t outer = 1;
  for (int d = 0; d < axis; ++d) outer *= g.n[d];
  a.n_inner = inner;
  a.n_outer = outer;
  a.stride_ax = inner * 2 * 6;
  a.stride_outer = (int64_t)L * inner * 2 * 6;
  padded_strides(g, 6, axis, a);
  a.T = points;
  a.ntiles = (int)((inner + points - 1) / points);
  dim3 grid((unsigned)(outer * a.ntiles));
  if (points == 1) return launch_tile_probe<1>(L, grid, s, src, dst, a);
  if (points == 2) return launch_tile_probe<2>(L, grid, s, src, dst, a);
  return launch_tile_probe<4>(L, grid, s, src, dst, a);
}

int launch_axis(const Grid &g, int ncomp, int axis, int dir, const double *in,
                double *out, const GreenDev *green, const SolveState *st,
                hipStream_t s, const Batch *b) {
  if (axis < 0 || axis >= g.dim - 1)
    return fail(JX_EINVAL, "launch_axis: axis out of range");
  if (dir == 0 && (axis != pivot_axis(g) || green == nullptr || ncomp == 1))
    return fail(JX_EINVAL, "launch_axis: Gamma-hat pass needs the pivot axis");
  FftLen P;
  int rc = get_fft_len((int)g.n[axis], &P);
  if (rc) return rc;
  if (reg_fft_length((int)g.n[axis]))
    return launch_axis_reg(g, ncomp, axis, dir, in, out, green, st, P, s, b);
  if (b) return fail(JX_EUNSUPPORTED, "axis: batch needs a register FFT length");
  AxisArgs a;
  a.L = (int)g.n[axis];
  a.Ls = lds_stride(a.L);
  a.nh = (int)g.nh;
  a.axis = axis;
  // points after `axis` (all inner axes, last one halved) are contiguous
  int64_t inner = g.nh;
  for (int d = axis + 1; d < g.dim - 1; ++d) inner *= g.n[d];
  int64_t outer = 1;
  for (int d = 0; d < axis; ++d) outer *= g.n[d];
  a.n_inner = inner;
  a.n_outer = outer;
  a.stride_ax = inner * 2 * ncomp;
  a.stride_outer = (int64_t)a.L * inner * 2 * ncomp;
  padded_strides(g, ncomp, axis, a);
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

}  // namespace jx
