// Pass A of the Lippmann-Schwinger iteration: forward real transform of the
// last axis, with the local operator (C - C0) applied on the way in.
//
// Data layouts (FP64, C order):
//   real field  (rows, L, C)        rows = n0 [* n1], L = n_last, C comps
//   spectrum    (rows, nh, C, 2)    nh = L / 2 + 1; complex values, the C
//                                   components of a frequency side by side
//                                   (for C = 1: janus's r2c layout,
//                                   _serial_fft.pyx:57-74, 181-200)
// Real sequences are paired two by two as the real and imaginary parts of
// one complex sequence, which is then transformed once.

#include "jx_device.h"

namespace jx {

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
    double *sr = spec + spec_row_off(row0 + r, (int64_t)nh * 2 * C, a.sp) +
                 (int64_t)k * 2 * C;
    *(double2 *)(sr + 2 * c) = make_double2(re, im);
  }
}


// Conjugate gradients: p <- r + beta p and x <- x + alpha p (the previous
// p) at one voxel, with the operations of the separate vector kernels; the
// new p is returned in e.
template <int C>
__device__ __forceinline__ void cg_update_voxel(const CgDev &cg, int64_t vox,
                                                double alpha, double beta,
                                                double *e) {
#pragma clang fp contract(off)
  double rv[C], pv[C], xv[C];
  if constexpr (C == 6) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double2 a = ((const double2 *)(cg.r + vox * 6))[i];
      double2 b = ((const double2 *)(cg.p + vox * 6))[i];
      double2 c = ((const double2 *)(cg.x + vox * 6))[i];
      rv[2 * i] = a.x; rv[2 * i + 1] = a.y;
      pv[2 * i] = b.x; pv[2 * i + 1] = b.y;
      xv[2 * i] = c.x; xv[2 * i + 1] = c.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      rv[c] = cg.r[vox * C + c];
      pv[c] = cg.p[vox * C + c];
      xv[c] = cg.x[vox * C + c];
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    double ap = alpha * pv[c];
    double bp = beta * pv[c];
    xv[c] = xv[c] + ap;
    e[c] = rv[c] + bp;
  }
  if constexpr (C == 6) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ((double2 *)(cg.x + vox * 6))[i] = make_double2(xv[2 * i], xv[2 * i + 1]);
      ((double2 *)(cg.p + vox * 6))[i] = make_double2(e[2 * i], e[2 * i + 1]);
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      cg.x[vox * C + c] = xv[c];
      cg.p[vox * C + c] = e[c];
    }
  }
}

// Register-resident last-axis transform for the register FFT lengths.  A
// block handles RB rows, i.e. 2 NP real sequences paired into NP complex
// ones (sequence q = r C + c; pair p = q / 2, part q % 2), TPP = L / V
// threads per pair.  Batched: the entry's iterate, spectrum, local operator
// and state come from its slot.
// LOCAL: 0 no local operator, else 1 + its LocalFamily.  CGP: the input is
// the conjugate-gradient direction, updated first (cg_update_voxel).
// Rows in reverse order within each XCD's range (xcd_tile_rev): the forward
// axis-1 pass that follows walks its range forwards, so it starts on the
// slabs this pass wrote last (256^3: Green apply 1.698 -> 1.686 ms,
// five-pass iteration 1.842 -> 1.832 ms, CG iteration 3.110 -> 3.095 ms).
#ifndef JX_ROWS_FWD_REVERSE
#define JX_ROWS_FWD_REVERSE 1
#endif

template <class P, int C, int NP, int LOCAL, bool CGP>
__global__ void __launch_bounds__(NP * reg_tpp<P>())
    k_rows_forward_reg(const double *__restrict__ x, double *__restrict__ spec,
                       int64_t nrows, SpecPad lay, const double2 *W, LocalDev loc,
                       const SolveState *st, BatchRef br, CgDev cg) {
  constexpr int L = P::L;
  const Slot *sl;
  const int64_t blk = batch_block<JX_ROWS_FWD_REVERSE != 0>(br, sl);
  if (sl) {
    x = sl->eps;
    spec = sl->spec;
    loc = sl->loc;
    st = sl->st;
    if constexpr (CGP) {
      cg.r = sl->cg_r;
      cg.p = sl->cg_p;
      cg.x = sl->eps;
      cg.cs = sl->cg_cs;
    }
  }
  x = global_ptr(x);
  spec = global_ptr(spec);
  loc.phase = global_ptr(loc.phase);
  loc.table = global_ptr(loc.table);
  loc.sph = global_ptr(loc.sph);
  loc.dev = global_ptr(loc.dev);
  loc.mat = global_ptr(loc.mat);
  st = global_ptr(st);
  W = global_ptr(W);
  double alpha = 0., beta = 0.;
  if constexpr (CGP) {
    cg.r = global_ptr(cg.r);
    cg.p = global_ptr(cg.p);
    cg.x = global_ptr(cg.x);
  }
  if (solve_done(st)) return;
  if constexpr (CGP) {
    alpha = cg.cs->alpha;
    beta = cg.cs->beta;
  }
  constexpr int TPP = reg_tpp<P>(), LS = reg_lds_stride(L), NH = L / 2 + 1;
  constexpr int V = reg_v<P>();
  constexpr int NT = NP * TPP;
  constexpr int RB = 2 * NP / C;
  static_assert((2 * NP) % C == 0, "whole rows per block");
  extern __shared__ double2 smem[];
  const int tid = threadIdx.x;
  const int64_t row0 = blk * RB;
  const int rb = (int)min((int64_t)RB, nrows - row0);
  __shared__ double2 twl[JX_ROWS_TQ ? L / 4 : 1];
  const double2 *Wt = stage_twiddles<L, NT, JX_ROWS_TQ>(W, twl);
  // stage the rows (local operator applied) into the pair sequences; a
  // compile-time trip count lets the loads of all rounds be issued together
  constexpr int KV = (RB * L + NT - 1) / NT;
#pragma unroll
  for (int kv = 0; kv < KV; ++kv) {
    const int w = tid + kv * NT;
    if (w >= RB * L) continue;
    int r = w / L, j = w - r * L;
    double t[C];
    if (r < rb) {
      int64_t vox = (row0 + r) * L + j;
      double e[C];
      if constexpr (CGP) {
        cg_update_voxel<C>(cg, vox, alpha, beta, e);
      } else if constexpr (C == 6) {
        const double2 *xv = (const double2 *)(x + vox * 6);
        double2 a = xv[0], b = xv[1], cc = xv[2];
        e[0] = a.x; e[1] = a.y; e[2] = b.x; e[3] = b.y; e[4] = cc.x; e[5] = cc.y;
      } else {
#pragma unroll
        for (int c = 0; c < C; ++c) e[c] = x[vox * C + c];
      }
      if constexpr (LOCAL != 0) {
        local_voxel_f<C, LOCAL - 1>(loc, vox, e, t);
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
  const int p = rows_tfast<P>() ? tid / reg_tpp<P>() : tid % NP;
  const int t = rows_tfast<P>() ? tid % reg_tpp<P>() : tid / NP;
  double2 *seq = smem + p * LS;
  double2 v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<P>(t, i))];
  __syncthreads();
  reg_fft<P, -1, JX_ROWS_TQ>(v, t, seq, Wt);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < V; ++i) seq[lp(reg_out_pos<P>(t, i))] = v[i];
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
      double *sp = spec + spec_row_off(row0 + r, NH * 12, lay) + k * 12 + 4 * cp;
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
    double *sr = spec + spec_row_off(row0 + r, NH * 2 * C, lay) + k * 2 * C;
    *(double2 *)(sr + 2 * c) = make_double2(re, im);
  }
}

// ============================================================ launchers

template <class P, int C>
static int launch_rows_forward_reg_lc(const Grid &g, const double *x,
                                      double *spec, const LocalDev &loc,
                                      const SolveState *st, const double2 *W,
                                      hipStream_t s, const Batch *b,
                                      const CgDev *cg) {
  constexpr int L = P::L;
  constexpr int NP = rows_np<P>();
  constexpr int RB = 2 * NP / C;
  size_t lds = (size_t)NP * reg_lds_stride(L) * sizeof(double2);
  int rc = check_lds(lds, "rows_forward");
  if (rc) return rc;
  const int64_t nbe = (g.rows + RB - 1) / RB;
  const BatchRef br = batch_ref(b, nbe);
  dim3 grid((unsigned)(nbe * batch_count(b))), block(NP * reg_tpp<P>());
  const CgDev cgd = cg ? *cg : CgDev();
  auto go = [&](auto kernel) {
    set_lds_attr(kernel, lds);
    hipLaunchKernelGGL(kernel, grid, block, lds, s, x, spec, g.rows,
                       spec_layout(g, C), W, loc, st, br, cgd);
  };
  // a batch always applies the local operator of its slots (all of one
  // family: the first slot's, which the caller passes as loc)
  if (cg) {
    if constexpr (C == 3 || C == 6) {
      if (loc.mode == LOCAL_NONE)  // polarization form: transform p itself
        go(k_rows_forward_reg<P, C, NP, 0, true>);
      else if (local_family(loc) == LF_MATRIX)
        go(k_rows_forward_reg<P, C, NP, 1 + LF_MATRIX, true>);
      else
        go(k_rows_forward_reg<P, C, NP, 1 + LF_ISO, true>);
    } else {
      return fail(JX_EINVAL, "conjugate gradients need sym components");
    }
  } else if (b || loc.mode != LOCAL_NONE) {
    if constexpr (C == 3 || C == 6) {
      if (local_family(loc) == LF_MATRIX)
        go(k_rows_forward_reg<P, C, NP, 1 + LF_MATRIX, false>);
      else
        go(k_rows_forward_reg<P, C, NP, 1 + LF_ISO, false>);
    } else {
      return fail(JX_EINVAL, "local operator needs sym components");
    }
  } else {
    go(k_rows_forward_reg<P, C, NP, 0, false>);
  }
  JX_HIP(hipGetLastError());
  return JX_OK;
}

static int rows_forward_reg(const Grid &g, int C, const double *x, double *spec,
                            const LocalDev &loc, const SolveState *st,
                            const double2 *W, hipStream_t s, const Batch *b,
                            const CgDev *cg = nullptr) {
  JX_WITH_PLAN_C((int)g.L, g.plan[PASS_ROWS_FWD], launch_rows_forward_reg_lc,
                 g, x, spec, loc, st, W, s, b, cg);
}

bool rows_cg_fusable(const Grid &g, int ncomp) {
  return (ncomp == 3 || ncomp == 6) && reg_fft_length((int)g.L);
}

int launch_rows_forward_cg(const Grid &g, int ncomp, const CgDev &cg,
                           double *spec, const LocalDev &loc,
                           const SolveState *st, hipStream_t s,
                           const Batch *b) {
  if (!rows_cg_fusable(g, ncomp))
    return fail(JX_EUNSUPPORTED, "rows_forward_cg needs a register FFT length");
  FftLen P;
  int rc = get_fft_len((int)g.L, &P);
  if (rc) return rc;
  return rows_forward_reg(g, ncomp, nullptr, spec, loc, st, P.W, s, b, &cg);
}

int launch_rows_forward(const Grid &g, int ncomp, const double *x, double *spec,
                        const LocalDev &loc, const SolveState *st,
                        hipStream_t s, const Batch *b) {
  FftLen P;
  int rc = get_fft_len((int)g.L, &P);
  if (rc) return rc;
  if (reg_fft_length((int)g.L))
    return rows_forward_reg(g, ncomp, x, spec, loc, st, P.W, s, b);
  if (b) return fail(JX_EUNSUPPORTED, "rows_forward: batch needs a register FFT length");
  RowsArgs a;
  a.nrows = g.rows;
  a.L = (int)g.L;
  a.nh = (int)g.nh;
  a.Ls = lds_stride(a.L);
  a.RB = rows_per_block(a.L, ncomp);
  a.scale = 1.0;
  a.sp = spec_layout(g, ncomp);
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

}  // namespace jx
