// Pass E of the Lippmann-Schwinger iteration: inverse real transform of the
// last axis (scaled by 1/N), optionally with the fixed-point update
// eps <- E (+ rhs) - y and the residual partial sums.  Layouts as in pass A
// (jx_rows_fwd.hip).

#include "jx_device.h"

namespace jx {

// ------------------------------------------------------------ pass E

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
    const double *sr = spec + spec_row_off(row0 + r, (int64_t)nh * 2 * C, a.sp) +
                       (int64_t)k * 2 * C;
    const double2 z = *(const double2 *)(sr + 2 * c);
    double re = z.x, im = z.y;
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
        double nv = ls.rhs ? ls_new_rhs(st->E[c], ls.rhs[vox * C + c], raw, a.scale)
                           : ls_new(st->E[c], raw, a.scale);
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

// Conjugate gradients: q = p + y at one voxel (y = c2r(spec) / N, the
// operations of the separate vector kernels); returns p : C0 : q.
template <int C>
__device__ __forceinline__ double cg_qp_voxel(const LsArgs &ls, double *q,
                                              int64_t vox, const double *yv) {
#pragma clang fp contract(off)
  double pv[C], qv[C];
  if constexpr (C == 6) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double2 a = ((const double2 *)(ls.cgp + vox * 6))[i];
      pv[2 * i] = a.x;
      pv[2 * i + 1] = a.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) pv[c] = ls.cgp[vox * C + c];
  }
#pragma unroll
  for (int c = 0; c < C; ++c) qv[c] = pv[c] + yv[c];
  if constexpr (C == 6) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
      ((double2 *)(q + vox * 6))[i] = make_double2(qv[2 * i], qv[2 * i + 1]);
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) q[vox * C + c] = qv[c];
  }
  return c0_dot<C>(pv, qv, ls.wd, ls.wt);
}

// Rows in reverse order within each XCD's range (xcd_tile_rev), as the
// fused pass: this pass follows the inverse axis-1 pass (256^3: Green apply
// 1.699 -> 1.690 ms, five-pass iteration 1.858 -> 1.842 ms, CG iteration
// 3.122 -> 3.111 ms).
#ifndef JX_ROWS_INV_REVERSE
#define JX_ROWS_INV_REVERSE 1
#endif

// Register-FFT variant for the register lengths.  MODE 0: y = c2r(spec);
// 1: y <- E - c2r(spec) with residual partials; 2: y <- (E + rhs) -
// c2r(spec) with residual partials; 3 (conjugate gradients): y <- p +
// c2r(spec) with the block sums of p : C0 : y in partials.  Batched: the
// entry's spectrum, iterate, rhs, partials and state come from its slot.
template <class P, int C, int NP, int MODE>
__global__ void __launch_bounds__(NP * reg_tpp<P>())
    k_rows_inverse_reg(const double *__restrict__ spec, double *y,
                       int64_t nrows, SpecPad lay, double scale, const double2 *W,
                       LsArgs ls, const SolveState *st, BatchRef br) {
  constexpr int L = P::L;
  const Slot *sl;
  const int64_t blk = batch_block<JX_ROWS_INV_REVERSE != 0>(br, sl);
  if (sl) {
    spec = sl->spec;
    y = MODE == 3 ? sl->cg_q : sl->eps;
    ls.rhs = sl->rhs;
    ls.partials = sl->partials;
    st = sl->st;
    if constexpr (MODE == 3) {
      ls.cgp = sl->cg_p;
      ls.wd = sl->cg_wd;
      ls.wt = sl->cg_wt;
    }
  }
  spec = global_ptr(spec);
  y = global_ptr(y);
  ls.rhs = global_ptr(ls.rhs);
  ls.partials = global_ptr(ls.partials);
  st = global_ptr(st);
  W = global_ptr(W);
  if constexpr (MODE == 3) ls.cgp = global_ptr(ls.cgp);
  if (solve_done(st)) return;
  constexpr int TPP = reg_tpp<P>(), LS = reg_lds_stride(L), NH = L / 2 + 1;
  constexpr int V = reg_v<P>();
  constexpr int NT = NP * TPP;
  constexpr int RB = 2 * NP / C;
  constexpr bool LS_UPDATE = MODE == 1 || MODE == 2;
  extern __shared__ double2 smem[];
  // one partial sum per wave (384-thread blocks: 512- and 1024-point plans)
  __shared__ double red[(NT + 63) / 64];
  const int tid = threadIdx.x;
  const int64_t row0 = blk * RB;
  const int rb = (int)min((int64_t)RB, nrows - row0);
  const int p = rows_tfast<P>() ? tid / TPP : tid % NP;
  const int t = rows_tfast<P>() ? tid % TPP : tid / NP;
  __shared__ double2 twl[JX_ROWS_TQ ? L / 4 : 1];
  const double2 *Wt = stage_twiddles<L, NT, JX_ROWS_TQ>(W, twl);
  // half spectra into the pair slots: slot k <- Xa[k], slot L-k <- Xb[k]
  // (0 < k < L/2; for C = 6 already combined into Xa + i Xb); DC and
  // Nyquist slots <- (Re Xa, Re Xb), their imaginary parts dropped as
  // FFTW's and pocketfft's c2r do
  if constexpr (C == 6) {
    load_half_spectra6<L, RB, NT, LS>(spec, lay, row0, rb, tid, smem);
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
      const double *sr = spec + spec_row_off(row0 + r, NH * 2 * C, lay) + k * 2 * C;
      {
        const double2 z = *(const double2 *)(sr + 2 * c);
        re = z.x;
        im = z.y;
      }
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
  if (C != 6 && HALF > 0) {
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
  for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<P>(t, i))];
  __syncthreads();
  reg_fft<P, 1, JX_ROWS_TQ>(v, t, seq, Wt);
  double acc = 0.;
  if constexpr (MODE == 3) {
    // all C components of a voxel meet in LDS (the C0 product has a trace
    // term)
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i)
      seq[lp(reg_out_pos<P>(t, i))] =
          make_double2(scale * v[i].x, scale * v[i].y);
    __syncthreads();
    constexpr int KV = (RB * L + NT - 1) / NT;
#pragma unroll
    for (int kv = 0; kv < KV; ++kv) {
      const int w = tid + kv * NT;
      if (w >= rb * L) continue;
      int r = w / L, j = w - r * L;
      double yv[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        int q = r * C + c;
        yv[c] = ((const double *)&smem[(q >> 1) * LS + lp(j)])[q & 1];
      }
      acc += cg_qp_voxel<C>(ls, y, (row0 + r) * L + j, yv);
    }
    double sum = block_sum(acc, red);
    if (tid == 0) ls.partials[blk] = sum;
    return;
  } else if constexpr (C == 6) {
    // pair p = components (c0, c0 + 1) of row r: 16-byte accesses
    const int r = p / 3, c0 = 2 * (p - r * 3);
    if (r < rb) {
      double2 *yr = (double2 *)(y + (row0 + r) * L * 6 + c0);
      const double2 *hr =
          MODE == 2 ? (const double2 *)(ls.rhs + (row0 + r) * L * 6 + c0)
                    : nullptr;
      const double e0 = LS_UPDATE ? st->E[c0] : 0.;
      const double e1 = LS_UPDATE ? st->E[c0 + 1] : 0.;
#pragma unroll
      for (int i = 0; i < V; ++i) {
        int64_t off = (int64_t)reg_out_pos<P>(t, i) * 3;
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
          int64_t off = (int64_t)reg_out_pos<P>(t, i) * C;
          double raw = part ? v[i].y : v[i].x;
          double val = scale * raw;
          if constexpr (LS_UPDATE) {
            double nv = ls_new(st->E[c], raw, scale);
            if constexpr (MODE == 2) nv = ls_new_rhs(st->E[c], hr[off], raw, scale);
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

// ============================================================ launchers

template <class P, int C>
static int launch_rows_inverse_reg_lc(const Grid &g, const double *spec,
                                      double *y, const LsArgs &la, bool upd,
                                      const SolveState *st, const double2 *W,
                                      hipStream_t s, const Batch *b) {
  constexpr int L = P::L;
  constexpr int NP = rows_np<P>();
  constexpr int RB = 2 * NP / C;
  size_t lds = (size_t)NP * reg_lds_stride(L) * sizeof(double2);
  int rc = check_lds(lds, "rows_inverse");
  if (rc) return rc;
  const int64_t nbe = (g.rows + RB - 1) / RB;
  const BatchRef br = batch_ref(b, nbe);
  dim3 grid((unsigned)(nbe * batch_count(b))), block(NP * reg_tpp<P>());
  double scale = 1.0 / (double)g.N;
  if (upd) {
    if constexpr (C == 3 || C == 6) {
      if (la.rhs) {
        set_lds_attr(k_rows_inverse_reg<P, C, NP, 2>, lds);
        hipLaunchKernelGGL((k_rows_inverse_reg<P, C, NP, 2>), grid, block,
                           lds, s, spec, y, g.rows, spec_layout(g, C), scale, W, la, st, br);
      } else {
        set_lds_attr(k_rows_inverse_reg<P, C, NP, 1>, lds);
        hipLaunchKernelGGL((k_rows_inverse_reg<P, C, NP, 1>), grid, block,
                           lds, s, spec, y, g.rows, spec_layout(g, C), scale, W, la, st, br);
      }
    } else {
      return fail(JX_EINVAL, "fixed point needs sym components");
    }
  } else {
    set_lds_attr(k_rows_inverse_reg<P, C, NP, 0>, lds);
    hipLaunchKernelGGL((k_rows_inverse_reg<P, C, NP, 0>), grid, block, lds,
                       s, spec, y, g.rows, spec_layout(g, C), scale, W, la, st, br);
  }
  JX_HIP(hipGetLastError());
  return JX_OK;
}

static int rows_inverse_reg(const Grid &g, int C, const double *spec, double *y,
                            const LsArgs &la, bool upd, const SolveState *st,
                            const double2 *W, hipStream_t s, const Batch *b) {
  JX_WITH_PLAN_C((int)g.L, g.plan[PASS_ROWS_INV], launch_rows_inverse_reg_lc,
                 g, spec, y, la, upd, st, W, s, b);
}

template <class P, int C>
static int launch_rows_inverse_cg_lc(const Grid &g, const double *spec,
                                     double *q, const LsArgs &la,
                                     const SolveState *st, const double2 *W,
                                     hipStream_t s, int *nblocks_out,
                                     const Batch *b) {
  if constexpr (C == 3 || C == 6) {
    constexpr int L = P::L;
    constexpr int NP = rows_np<P>();
    constexpr int RB = 2 * NP / C;
    size_t lds = (size_t)NP * reg_lds_stride(L) * sizeof(double2);
    int rc = check_lds(lds, "rows_inverse_cg");
    if (rc) return rc;
    const int64_t nb = (g.rows + RB - 1) / RB;
    if (nblocks_out) *nblocks_out = (int)nb;
    double scale = 1.0 / (double)g.N;
    set_lds_attr(k_rows_inverse_reg<P, C, NP, 3>, lds);
    hipLaunchKernelGGL((k_rows_inverse_reg<P, C, NP, 3>),
                       dim3((unsigned)(nb * batch_count(b))),
                       dim3(NP * reg_tpp<P>()), lds, s, spec, q, g.rows, spec_layout(g, C), scale,
                       W, la, st, batch_ref(b, nb));
    JX_HIP(hipGetLastError());
    return JX_OK;
  } else {
    return fail(JX_EINVAL, "conjugate gradients need sym components");
  }
}

int launch_rows_inverse_cg(const Grid &g, int ncomp, const double *spec,
                           double *q, const double *p, double *partials,
                           double wd, double wt, const SolveState *st,
                           int *nblocks_out, hipStream_t s, const Batch *b) {
  if (!rows_cg_fusable(g, ncomp))
    return fail(JX_EUNSUPPORTED, "rows_inverse_cg needs a register FFT length");
  FftLen P;
  int rc = get_fft_len((int)g.L, &P);
  if (rc) return rc;
  LsArgs la;
  la.partials = partials;
  la.cgp = p;
  la.wd = wd;
  la.wt = wt;
  const int C = ncomp;
  JX_WITH_PLAN_C((int)g.L, g.plan[PASS_ROWS_INV], launch_rows_inverse_cg_lc,
                 g, spec, q, la, st, P.W, s, nblocks_out, b);
}

int rows_inverse_nblocks(const Grid &g, int ncomp) {
  if (reg_fft_length((int)g.L)) return (int)reg_rows_max_blocks(g, ncomp);
  int RB = rows_per_block((int)g.L, ncomp);
  return (int)((g.rows + RB - 1) / RB);
}

int launch_rows_inverse(const Grid &g, int ncomp, const double *spec, double *y,
                        const RowsInverseLs &ls, const SolveState *st,
                        int *nblocks_out, hipStream_t s, const Batch *b) {
  FftLen P;
  int rc = get_fft_len((int)g.L, &P);
  if (rc) return rc;
  if (int RBr = reg_rows_per_block((int)g.L, ncomp, g.plan[PASS_ROWS_INV])) {
    LsArgs la;
    la.rhs = ls.rhs;
    la.partials = ls.partials;
    if (nblocks_out) *nblocks_out = (int)((g.rows + RBr - 1) / RBr);
    return rows_inverse_reg(g, ncomp, spec, y, la, ls.enabled, st, P.W, s, b);
  }
  if (b) return fail(JX_EUNSUPPORTED, "rows_inverse: batch needs a register FFT length");
  RowsArgs a;
  a.nrows = g.rows;
  a.L = (int)g.L;
  a.nh = (int)g.nh;
  a.Ls = lds_stride(a.L);
  a.RB = rows_per_block(a.L, ncomp);
  a.scale = 1.0 / (double)g.N;
  a.sp = spec_layout(g, ncomp);
  size_t lds = rows_lds_bytes(a.L, ncomp, a.RB);
  if ((rc = check_lds(lds, "rows_inverse"))) return rc;
  int64_t nb = (g.rows + a.RB - 1) / a.RB;
  if (nblocks_out) *nblocks_out = (int)nb;
  LsArgs la;
  la.rhs = ls.rhs;
  la.partials = ls.partials;
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

}  // namespace jx
