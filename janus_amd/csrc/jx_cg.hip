// Conjugate gradients on the Lippmann-Schwinger operator (no janus
// counterpart in the library; janus drives Krylov solvers from its
// examples, examples/square_inclusion.py:40-47, 71-86 with PETSc CG on the
// polarization form).
//
// Strain form: the fluctuation x = eps - E solves (I + G dC) x = -G[dC : E]
// with G = Gamma0, in the inner product <a, b> = sum_voxels a : C0 : b.
// G C0 is the C0-orthogonal projector onto compatible zero-mean fields for
// the truncated operator (exactly on odd grids), so the operator is
// self-adjoint and positive definite there whenever C = C0 + dC is, for any
// isotropic C0 -- including C0 = C_matrix, where the basic scheme diverges.
// For the filtered operator it is close to self-adjoint and CG converges in
// practice.  The residual after iteration k is ||r_k|| / (sqrt(N) ||E||),
// which for the basic scheme is its own residual ||eps_{k+1} - eps_k||.
//
// One iteration: q = A p through the five passes of the Green operator
// (pass A applies dC to p) and three vector kernels with deterministic
// per-block partial sums; alpha, beta and the residual are computed on the
// device by a one-block kernel, and every kernel returns at once when the
// device state says the solve is done.

#include "jx_device.h"
#include "jx_handles.h"

namespace jx {

int ls_matvec(jx_solver *S, const double *x, double *y, const SolveState *st,
              hipStream_t s, const LocalDev *loc, double *spec) {
  jx_green *G = S->green;
  const Grid &g = G->g;
  const int C = G->sym;
  if (!spec) spec = G->spec;
  int rc = launch_rows_forward(g, C, x, spec, loc ? *loc : S->local->dev,
                               st, s);
  if (rc) return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), -1, spec, spec, nullptr, st, s)))
    return rc;
  if ((rc = launch_axis(g, C, pivot_axis(g), 0, spec, spec, &G->dev, st, s)))
    return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), +1, spec, spec, nullptr, st, s)))
    return rc;
  RowsInverseLs plain;
  return launch_rows_inverse(g, C, spec, y, plain, st, nullptr, s);
}

namespace {

enum CgMode {
  CG_FILL, CG_INIT, CG_QP, CG_XR, CG_P, CG_FINAL, CG_R, CG_FINAL_P,
  CG_INIT_E,  // polarization form: r = p = E, x = 0
  CG_QL,      // polarization form: q = G[p] + L p
  CG_STEP_X   // x += alpha p (the fused iterations' last step)
};
enum CgPhase { PH_INIT, PH_ALPHA, PH_BETA };

constexpr int kCgBlocks = 4096;  // blocks of the vector kernels (partials)
// 1: the vector updates are folded into passes A and E where the last axis
// has a register FFT length (cg_iteration_fused); 0: separate kernels.
#ifndef JX_CG_FUSED
#define JX_CG_FUSED 1
#endif
constexpr int kScalarThreads = 1024;

// Batched (tab != nullptr): block e nb + b works as block b of entry e's
// solve, whose fields, scalars, weights, state and partial sums come from
// its slot, so that every entry gets the single solve's partial sums.
template <int C, int MODE>
__global__ void __launch_bounds__(kThreads)
    k_cg_vec(double *x, double *r, double *p, double *q, int64_t N, double wd,
             double wt, const CgScalars *cs, const SolveState *st,
             double *partials, const double *rhs, LocalDev loc,
             const Slot *tab, int nb) {
#pragma clang fp contract(off)
  int blk = blockIdx.x, nblk = gridDim.x;
  if (tab) {
    const Slot *sl = tab + blk / nb;
    blk -= (blk / nb) * nb;
    nblk = nb;
    x = sl->eps;
    r = sl->cg_r;
    p = sl->cg_p;
    q = sl->cg_q;
    cs = sl->cg_cs;
    st = sl->st;
    partials = sl->cg_partials;
    wd = sl->cg_wd;
    wt = sl->cg_wt;
  }
  // the fill and the final sum run whatever the state
  constexpr bool FINAL =
      MODE == CG_FINAL || MODE == CG_FINAL_P || MODE == CG_STEP_X;
  if (MODE != CG_FILL && !FINAL && solve_done(st)) return;
  __shared__ double red[kThreads / 64];
  double acc0 = 0., acc1 = 0.;
  const double alpha = (MODE == CG_XR || MODE == CG_R || MODE == CG_FINAL_P ||
                        MODE == CG_STEP_X)
                           ? cs->alpha
                           : 0.;
  const double beta = MODE == CG_P ? cs->beta : 0.;
  for (int64_t v = (int64_t)blk * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)nblk * blockDim.x) {
    double e[C];
    if constexpr (MODE == CG_FILL) {
#pragma unroll
      for (int c = 0; c < C; ++c) e[c] = st->E[c];
      store_voxel<C>(x, v, e);
    } else if constexpr (MODE == CG_FINAL) {
      load_voxel<C>(x, v, e);
#pragma unroll
      for (int c = 0; c < C; ++c) e[c] = st->E[c] + e[c];
      store_voxel<C>(x, v, e);
    } else if constexpr (MODE == CG_FINAL_P) {
      // the fused iteration leaves x + alpha p of the last step to here
      double pv[C];
      load_voxel<C>(x, v, e);
      load_voxel<C>(p, v, pv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        double ap = alpha * pv[c];
        e[c] = e[c] + ap;
        e[c] = st->E[c] + e[c];
      }
      store_voxel<C>(x, v, e);
    } else if constexpr (MODE == CG_R) {
      // r -= alpha q (x is updated by the next pass A)
      double rv[C], qv[C];
      load_voxel<C>(r, v, rv);
      load_voxel<C>(q, v, qv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        double aq = alpha * qv[c];
        rv[c] = rv[c] - aq;
      }
      store_voxel<C>(r, v, rv);
      acc0 += c0_dot<C>(rv, rv, wd, wt);
#pragma unroll
      for (int c = 0; c < C; ++c) acc1 += rv[c] * rv[c];
    } else if constexpr (MODE == CG_INIT) {
      // r = rhs - G[dC : E] (G[dC : E] in q, rhs in r or absent), p = r,
      // x = 0
      double rv[C], z[C];
      load_voxel<C>(q, v, rv);
      double hv[C];
      if (rhs) load_voxel<C>(rhs, v, hv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        rv[c] = rhs ? hv[c] - rv[c] : -rv[c];
        z[c] = 0.;
      }
      store_voxel<C>(r, v, rv);
      store_voxel<C>(p, v, rv);
      store_voxel<C>(x, v, z);
      acc0 += c0_dot<C>(rv, rv, wd, wt);
#pragma unroll
      for (int c = 0; c < C; ++c) acc1 += rv[c] * rv[c];
    } else if constexpr (MODE == CG_INIT_E) {
      double z[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        e[c] = st->E[c];
        z[c] = 0.;
      }
      store_voxel<C>(r, v, e);
      store_voxel<C>(p, v, e);
      store_voxel<C>(x, v, z);
      acc0 += c0_dot<C>(e, e, wd, wt);
#pragma unroll
      for (int c = 0; c < C; ++c) acc1 += e[c] * e[c];
    } else if constexpr (MODE == CG_STEP_X) {
      double pv[C];
      load_voxel<C>(x, v, e);
      load_voxel<C>(p, v, pv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        double ap = alpha * pv[c];
        e[c] = e[c] + ap;
      }
      store_voxel<C>(x, v, e);
    } else if constexpr (MODE == CG_QL) {
      // q = G[p] + L p (the local operator applied voxel by voxel)
      double pv[C], qv[C], lv[C];
      load_voxel<C>(p, v, pv);
      load_voxel<C>(q, v, qv);
      local_voxel<C>(loc, v, pv, lv);
#pragma unroll
      for (int c = 0; c < C; ++c) qv[c] = qv[c] + lv[c];
      store_voxel<C>(q, v, qv);
      acc0 += c0_dot<C>(pv, qv, wd, wt);
    } else if constexpr (MODE == CG_QP) {
      // q = p + G[dC : p]
      double pv[C], qv[C];
      load_voxel<C>(p, v, pv);
      load_voxel<C>(q, v, qv);
#pragma unroll
      for (int c = 0; c < C; ++c) qv[c] = pv[c] + qv[c];
      store_voxel<C>(q, v, qv);
      acc0 += c0_dot<C>(pv, qv, wd, wt);
    } else if constexpr (MODE == CG_XR) {
      double xv[C], rv[C], pv[C], qv[C];
      load_voxel<C>(x, v, xv);
      load_voxel<C>(r, v, rv);
      load_voxel<C>(p, v, pv);
      load_voxel<C>(q, v, qv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        double ap = alpha * pv[c];
        double aq = alpha * qv[c];
        xv[c] = xv[c] + ap;
        rv[c] = rv[c] - aq;
      }
      store_voxel<C>(x, v, xv);
      store_voxel<C>(r, v, rv);
      acc0 += c0_dot<C>(rv, rv, wd, wt);
#pragma unroll
      for (int c = 0; c < C; ++c) acc1 += rv[c] * rv[c];
    } else {  // CG_P
      double rv[C], pv[C];
      load_voxel<C>(r, v, rv);
      load_voxel<C>(p, v, pv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        double bp = beta * pv[c];
        pv[c] = rv[c] + bp;
      }
      store_voxel<C>(p, v, pv);
    }
  }
  if constexpr (MODE == CG_INIT || MODE == CG_QP || MODE == CG_XR ||
                MODE == CG_R || MODE == CG_INIT_E || MODE == CG_QL) {
    double s0 = block_sum(acc0, red);
    __syncthreads();
    double s1 = block_sum(acc1, red);
    if (threadIdx.x == 0) {
      partials[2 * blk] = s0;
      partials[2 * blk + 1] = s1;
    }
  }
}

// Fixed-order sum of column `col` of the partials (nb rows of `stride`
// values); valid in thread 0.
__device__ double sum_partials(const double *partials, int nb, int stride,
                               int col, double *red) {
  double acc = 0.;
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    acc += partials[stride * i + col];
  return block_sum(acc, red);
}

// Scalars from the partial sums: PH_INIT and PH_BETA read two columns
// (stride 2), PH_ALPHA the first column (stride 2 from the vector kernel,
// 1 from the fused pass E).
// Batched (tab != nullptr): block e serves entry e, from its slot; the
// partial sums are the pass-E ones (slot partials) or, with vec_sums, the
// vector kernel's.
template <int PHASE>
__global__ void __launch_bounds__(kScalarThreads)
    k_cg_scalar(const double *partials, int nb, int stride, CgScalars *cs,
                SolveState *st, double *res_hist, const Slot *tab,
                int vec_sums) {
  if (tab) {
    const Slot *sl = tab + blockIdx.x;
    partials = vec_sums ? sl->cg_partials : sl->partials;
    cs = sl->cg_cs;
    st = sl->st;
    res_hist = sl->res_hist;
  }
  if (solve_done(st)) return;
  __shared__ double red[kScalarThreads / 64];
  double s0 = sum_partials(partials, nb, stride, 0, red);
  double s1 = 0.;
  if (PHASE != PH_ALPHA) {
    __syncthreads();
    s1 = sum_partials(partials, nb, stride, 1, red);
  }
  if (threadIdx.x != 0) return;
  if constexpr (PHASE == PH_INIT) {
    cs->rr = s0;
    cs->alpha = 0.;
    cs->beta = 0.;
    // homogeneous medium: the initial iterate E is the solution
    if (s0 == 0.) st->done = JX_SOLVE_CONVERGED;
  } else if constexpr (PHASE == PH_ALPHA) {
    cs->alpha = s0 != 0. ? cs->rr / s0 : 0.;
  } else {
    double res = sqrt(s1) * st->res_scale;
    int it = st->iter;
    res_hist[it] = res;
    st->iter = it + 1;
    if (!isfinite(res))
      st->done = JX_SOLVE_NONFINITE;
    else if (res < st->tol)
      st->done = JX_SOLVE_CONVERGED;
    else if (it + 1 >= st->max_iter)
      st->done = JX_SOLVE_MAXITER;
    cs->beta = cs->rr != 0. ? s0 / cs->rr : 0.;
    cs->rr = s0;
  }
}

template <int C, int MODE>
int vec(double *x, double *r, double *p, double *q, const Grid &g, double wd,
        double wt, const CgScalars *cs, const SolveState *st,
        double *partials, hipStream_t s, const double *rhs = nullptr,
        const LocalDev &loc = LocalDev()) {
  hipLaunchKernelGGL((k_cg_vec<C, MODE>), dim3(kCgBlocks), dim3(kThreads), 0,
                     s, x, r, p, q, g.N, wd, wt, cs, st, partials, rhs, loc,
                     nullptr, kCgBlocks);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

// The vector kernel over the entries of a batch (kCgBlocks blocks each).
template <int C, int MODE>
int vec_batched(const Grid &g, const Batch &b, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_vec<C, MODE>), dim3(kCgBlocks * b.n),
                     dim3(kThreads), 0, s, nullptr, nullptr, nullptr, nullptr,
                     g.N, 0., 0., nullptr, nullptr, nullptr, nullptr,
                     LocalDev(), b.tab, kCgBlocks);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

template <int PHASE>
int scalar(const double *partials, CgScalars *cs, SolveState *st,
           double *res_hist, hipStream_t s, int nb = kCgBlocks,
           int stride = 2) {
  hipLaunchKernelGGL((k_cg_scalar<PHASE>), dim3(1), dim3(kScalarThreads), 0, s,
                     partials, nb, stride, cs, st, res_hist, nullptr, 0);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

template <int PHASE>
int scalar_batched(const Batch &b, int nb, int stride, int vec_sums,
                   hipStream_t s) {
  hipLaunchKernelGGL((k_cg_scalar<PHASE>), dim3(b.n), dim3(kScalarThreads), 0,
                     s, nullptr, nb, stride, nullptr, nullptr, nullptr, b.tab,
                     vec_sums);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

template <int C>
int cg_iteration(jx_solver *S, const CgWork &w, double *x, double wd,
                 double wt, hipStream_t s) {
  const Grid &g = S->green->g;
  int rc;
  if ((rc = ls_matvec(S, w.p, w.q, w.st, s, nullptr, w.spec))) return rc;
  if ((rc = vec<C, CG_QP>(x, w.r, w.p, w.q, g, wd, wt, w.cs, w.st,
                          w.vec_partials, s)))
    return rc;
  if ((rc = scalar<PH_ALPHA>(w.vec_partials, w.cs, w.st, w.res_hist, s)))
    return rc;
  if ((rc = vec<C, CG_XR>(x, w.r, w.p, w.q, g, wd, wt, w.cs, w.st,
                          w.vec_partials, s)))
    return rc;
  if ((rc = scalar<PH_BETA>(w.vec_partials, w.cs, w.st, w.res_hist, s)))
    return rc;
  return vec<C, CG_P>(x, w.r, w.p, w.q, g, wd, wt, w.cs, w.st,
                      w.vec_partials, s);
}

// One iteration with the vector updates folded into the Green operator's
// last-axis passes (register lengths): pass A first sets p <- r + beta p
// and x <- x + alpha p, then transforms dC : p; pass E writes q = p +
// G[dC : p] with the block sums of p : C0 : q; then alpha, r -= alpha q
// with its sums, beta: 768 Nh + 481 N bytes instead of 768 Nh + 673 N in 3D.
template <int C>
int cg_iteration_fused(jx_solver *S, const CgWork &w, double *x, double wd,
                       double wt, hipStream_t s) {
  jx_green *G = S->green;
  const Grid &g = G->g;
  CgDev cg;
  cg.r = w.r;
  cg.p = w.p;
  cg.x = x;
  cg.cs = w.cs;
  int rc = launch_rows_forward_cg(g, C, cg, w.spec, S->local->dev, w.st, s);
  if (rc) return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), -1, w.spec, w.spec, nullptr, w.st, s)))
    return rc;
  if ((rc = launch_axis(g, C, pivot_axis(g), 0, w.spec, w.spec, &G->dev, w.st, s)))
    return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), +1, w.spec, w.spec, nullptr, w.st, s)))
    return rc;
  int nb = 0;
  if ((rc = launch_rows_inverse_cg(g, C, w.spec, w.q, w.p, w.rows_partials,
                                   wd, wt, w.st, &nb, s)))
    return rc;
  if ((rc = scalar<PH_ALPHA>(w.rows_partials, w.cs, w.st, w.res_hist, s, nb,
                             1)))
    return rc;
  if ((rc = vec<C, CG_R>(x, w.r, w.p, w.q, g, wd, wt, w.cs, w.st,
                         w.vec_partials, s)))
    return rc;
  return scalar<PH_BETA>(w.vec_partials, w.cs, w.st, w.res_hist, s);
}

// q = G[dC : E], then r = p = rhs - q, x = 0 (x is eps), and the first
// residual.
template <int C>
int cg_start_c(jx_solver *S, const CgWork &w, double *eps, const double *rhs,
               double wd, double wt, hipStream_t s) {
  const Grid &g = S->green->g;
  int rc;
  if ((rc = vec<C, CG_FILL>(eps, w.r, w.p, w.q, g, wd, wt, w.cs, w.st,
                            w.vec_partials, s)))
    return rc;
  if ((rc = ls_matvec(S, eps, w.q, w.st, s, nullptr, w.spec))) return rc;
  if ((rc = vec<C, CG_INIT>(eps, w.r, w.p, w.q, g, wd, wt, w.cs, w.st,
                            w.vec_partials, s, rhs)))
    return rc;
  return scalar<PH_INIT>(w.vec_partials, w.cs, w.st, w.res_hist, s);
}

// eps = E + x (+ alpha p, the fused iterations' last x step)
template <int C>
int cg_finish_c(jx_solver *S, const CgWork &w, double *eps, double wd,
                double wt, bool fused, hipStream_t s) {
  const Grid &g = S->green->g;
  if (fused)
    return vec<C, CG_FINAL_P>(eps, w.r, w.p, w.q, g, wd, wt, w.cs, w.st,
                              w.vec_partials, s);
  return vec<C, CG_FINAL>(eps, w.r, w.p, w.q, g, wd, wt, w.cs, w.st,
                          w.vec_partials, s);
}

bool cg_fused_grid(const Grid &g, int C) {
  return JX_CG_FUSED && rows_cg_fusable(g, C);
}

template <int C>
int cg_solve(jx_solver *S, double *eps, const double *rhs, int max_iter,
             double wd, double wt, hipStream_t s) {
  const Grid &g = S->green->g;
  CgWork w;
  w.r = S->cg_buf;
  w.p = w.r + C * g.N;
  w.q = w.p + C * g.N;
  w.spec = S->green->spec;
  w.rows_partials = S->partials;
  w.vec_partials = S->cg_partials;
  w.cs = S->cg_s;
  w.st = S->st;
  w.res_hist = S->res_dev;
  int rc = cg_start_c<C>(S, w, eps, rhs, wd, wt, s);
  if (rc) return rc;
  const bool fused = cg_fused_grid(g, C);
  int enqueued = 0, chunk = 4;
  while (true) {
    int n = std::min(chunk, max_iter - enqueued);
    for (int i = 0; i < n; ++i) {
      rc = fused ? cg_iteration_fused<C>(S, w, eps, wd, wt, s)
                 : cg_iteration<C>(S, w, eps, wd, wt, s);
      if (rc) return rc;
    }
    enqueued += n;
    JX_HIP(hipMemcpyAsync(S->st_host, S->st, sizeof(SolveState),
                          hipMemcpyDeviceToHost, s));
    JX_HIP(hipStreamSynchronize(s));
    if (S->st_host->done || enqueued >= max_iter) break;
    chunk = std::min(chunk * 2, 64);
  }
  return cg_finish_c<C>(S, w, eps, wd, wt, fused, s);
}

// Polarization form (examples/square_inclusion.py:17-52, 77-85): with the
// solver's local operator L = dC^-1 (the example's tau2eps), tau solves
// (L + G) tau = E by CG in the Euclidean inner product (wd = 1, wt = 0),
// from tau = 0; eps = L tau.  Needs L + G definite: C0 strictly below (or
// above) every phase.
template <int C>
int cg_pol_iteration(jx_solver *S, double *x, hipStream_t s) {
  const Grid &g = S->green->g;
  double *r = S->cg_buf, *p = r + C * g.N, *q = p + C * g.N;
  const LocalDev none;
  int rc;
  if ((rc = ls_matvec(S, p, q, S->st, s, &none))) return rc;
  if ((rc = vec<C, CG_QL>(x, r, p, q, g, 1., 0., S->cg_s, S->st,
                          S->cg_partials, s, nullptr, S->local->dev)))
    return rc;
  if ((rc = scalar<PH_ALPHA>(S->cg_partials, S->cg_s, S->st, S->res_dev, s)))
    return rc;
  if ((rc = vec<C, CG_XR>(x, r, p, q, g, 1., 0., S->cg_s, S->st,
                          S->cg_partials, s)))
    return rc;
  if ((rc = scalar<PH_BETA>(S->cg_partials, S->cg_s, S->st, S->res_dev, s)))
    return rc;
  return vec<C, CG_P>(x, r, p, q, g, 1., 0., S->cg_s, S->st, S->cg_partials,
                      s);
}

// The same iteration with the direction and iterate updates folded into
// pass A (register lengths): pass A sets p <- r + beta p and x <- x +
// alpha p (the previous p) and transforms p; then q = G[p] (plain pass E),
// q += L p with the sums p . q, alpha, r -= alpha q with its sums, beta.
// The last step of x is applied after the loop (CG_STEP_X).
template <int C>
int cg_pol_iteration_fused(jx_solver *S, double *x, hipStream_t s) {
  jx_green *G = S->green;
  const Grid &g = G->g;
  double *r = S->cg_buf, *p = r + C * g.N, *q = p + C * g.N;
  CgDev cg;
  cg.r = r;
  cg.p = p;
  cg.x = x;
  cg.cs = S->cg_s;
  const LocalDev none;
  int rc = launch_rows_forward_cg(g, C, cg, G->spec, none, S->st, s);
  if (rc) return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), -1, G->spec, G->spec, nullptr, S->st, s)))
    return rc;
  if ((rc = launch_axis(g, C, pivot_axis(g), 0, G->spec, G->spec, &G->dev, S->st, s)))
    return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), +1, G->spec, G->spec, nullptr, S->st, s)))
    return rc;
  RowsInverseLs plain;
  if ((rc = launch_rows_inverse(g, C, G->spec, q, plain, S->st, nullptr, s)))
    return rc;
  if ((rc = vec<C, CG_QL>(x, r, p, q, g, 1., 0., S->cg_s, S->st,
                          S->cg_partials, s, nullptr, S->local->dev)))
    return rc;
  if ((rc = scalar<PH_ALPHA>(S->cg_partials, S->cg_s, S->st, S->res_dev, s)))
    return rc;
  if ((rc = vec<C, CG_R>(x, r, p, q, g, 1., 0., S->cg_s, S->st,
                         S->cg_partials, s)))
    return rc;
  return scalar<PH_BETA>(S->cg_partials, S->cg_s, S->st, S->res_dev, s);
}

template <int C>
int cg_pol_solve(jx_solver *S, double *tau, double *eps, int max_iter,
                 hipStream_t s) {
  const Grid &g = S->green->g;
  double *r = S->cg_buf, *p = r + C * g.N, *q = p + C * g.N;
  int rc;
  if ((rc = vec<C, CG_INIT_E>(tau, r, p, q, g, 1., 0., S->cg_s, S->st,
                              S->cg_partials, s)))
    return rc;
  if ((rc = scalar<PH_INIT>(S->cg_partials, S->cg_s, S->st, S->res_dev, s)))
    return rc;
  const bool fused = JX_CG_FUSED && rows_cg_fusable(g, C);
  int enqueued = 0, chunk = 4;
  while (true) {
    int n = std::min(chunk, max_iter - enqueued);
    for (int i = 0; i < n; ++i)
      if ((rc = fused ? cg_pol_iteration_fused<C>(S, tau, s)
                      : cg_pol_iteration<C>(S, tau, s)))
        return rc;
    enqueued += n;
    JX_HIP(hipMemcpyAsync(S->st_host, S->st, sizeof(SolveState),
                          hipMemcpyDeviceToHost, s));
    JX_HIP(hipStreamSynchronize(s));
    if (S->st_host->done || enqueued >= max_iter) break;
    chunk = std::min(chunk * 2, 64);
  }
  // the fused iterations leave the last step of tau to here; eps = L tau
  if (fused && (rc = vec<C, CG_STEP_X>(tau, r, p, q, g, 1., 0., S->cg_s,
                                       S->st, S->cg_partials, s)))
    return rc;
  return launch_local_apply(g, C, S->local->dev, tau, eps, s);
}

}  // namespace

int64_t cg_vec_partials_doubles() { return 2 * (int64_t)kCgBlocks; }

void cg_weights(const jx_green *G, double *wd, double *wt) {
  // C0 inner product: 2 g0 (a . b) + (k0 - 2 g0 / dim) tr a tr b
  const double g0 = G->g0, nu0 = G->nu0;
  const int dim = G->g.dim;
  const double k0 = dim == 2 ? g0 / (1. - 2. * nu0)
                             : 2. / 3. * (1. + nu0) / (1. - 2. * nu0) * g0;
  *wd = 2. * g0;
  *wt = k0 - 2. * g0 / dim;
}

int cg_start(jx_solver *S, const CgWork &w, double *eps, const double *rhs,
             double wd, double wt, hipStream_t s) {
  return S->green->sym == 3 ? cg_start_c<3>(S, w, eps, rhs, wd, wt, s)
                            : cg_start_c<6>(S, w, eps, rhs, wd, wt, s);
}

int cg_finish(jx_solver *S, const CgWork &w, double *eps, double wd,
              double wt, hipStream_t s) {
  const bool fused = cg_fused_grid(S->green->g, S->green->sym);
  return S->green->sym == 3
             ? cg_finish_c<3>(S, w, eps, wd, wt, fused, s)
             : cg_finish_c<6>(S, w, eps, wd, wt, fused, s);
}

bool cg_batchable(const Grid &g, int C) {
  return JX_CG_FUSED && batchable(g, C) && rows_cg_fusable(g, C);
}

int cg_iteration_batched(const jx_green *G, const LocalDev &loc,
                         const Batch &b, hipStream_t s) {
  const Grid &g = G->g;
  const int C = G->sym;
  if (!cg_batchable(g, C))
    return fail(JX_EUNSUPPORTED, "batched conjugate gradients need register "
                                 "FFT lengths on every axis");
  const CgDev none;
  int rc = launch_rows_forward_cg(g, C, none, nullptr, loc, nullptr, s, &b);
  if (rc) return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), -1, nullptr, nullptr, nullptr,
                        nullptr, s, &b)))
    return rc;
  if ((rc = launch_axis(g, C, pivot_axis(g), 0, nullptr, nullptr, &G->dev,
                        nullptr, s, &b)))
    return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), +1, nullptr, nullptr, nullptr,
                        nullptr, s, &b)))
    return rc;
  int nb = 0;
  if ((rc = launch_rows_inverse_cg(g, C, nullptr, nullptr, nullptr, nullptr,
                                   0., 0., nullptr, &nb, s, &b)))
    return rc;
  if ((rc = scalar_batched<PH_ALPHA>(b, nb, 1, 0, s))) return rc;
  if ((rc = C == 3 ? vec_batched<3, CG_R>(g, b, s)
                   : vec_batched<6, CG_R>(g, b, s)))
    return rc;
  return scalar_batched<PH_BETA>(b, kCgBlocks, 2, 1, s);
}

}  // namespace jx

using namespace jx;

extern "C" {

int jx_solver_solve_cg(jx_solver *S, double *eps, const double *E,
                       const double *rhs, double tol, int max_iter,
                       double res_scale, double *res_hist, int *iters,
                       int *status, void *stream) {
  if (!S || !eps || !E) return fail(JX_EINVAL, "jx_solver_solve_cg: NULL argument");
  JX_ALIGNED(eps, "jx_solver_solve_cg");
  JX_ALIGNED(rhs, "jx_solver_solve_cg");
  if (max_iter < 1) return fail(JX_EINVAL, "max_iter must be >= 1");
  if (!S->local->symmetric)
    return fail(JX_EINVAL, "conjugate gradients need a symmetric local operator");
  hipStream_t s = (hipStream_t)stream;
  jx_green *G = S->green;
  const Grid &g = G->g;
  const int C = G->sym;
  const int dim = g.dim;
  if (res_scale <= 0.) {
    double nE = 0.;
    for (int c = 0; c < C; ++c) nE += E[c] * E[c];
    nE = std::sqrt(nE);
    if (nE == 0.) return fail(JX_EINVAL, "res_scale must be given when E = 0");
    res_scale = 1.0 / std::sqrt((double)g.N) / nE;  // square_basic.py:86
  }
  // buffers: r, p, q fields, partial sums, scalars, residual history
  if (!S->cg_buf) {
    JX_HIP(hipMalloc((void **)&S->cg_buf, sizeof(double) * 3 * C * g.N));
    JX_HIP(hipMalloc((void **)&S->cg_partials, sizeof(double) * 2 * kCgBlocks));
    JX_HIP(hipMalloc((void **)&S->cg_s, sizeof(CgScalars)));
  }
  if (S->res_cap < max_iter) {
    if (S->res_dev) (void)hipFree(S->res_dev);
    S->res_dev = nullptr;
    S->res_cap = 0;
    JX_HIP(hipMalloc((void **)&S->res_dev, sizeof(double) * max_iter));
    S->res_cap = max_iter;
  }
  SolveState init = {};
  init.tol = tol;
  init.res_scale = res_scale;
  init.max_iter = max_iter;
  for (int c = 0; c < 6; ++c) init.E[c] = c < C ? E[c] : 0.;
  *S->st_host = init;
  JX_HIP(hipMemcpyAsync(S->st, S->st_host, sizeof(SolveState),
                        hipMemcpyHostToDevice, s));
  double wd = 0., wt = 0.;
  cg_weights(G, &wd, &wt);
  int rc = C == 3 ? cg_solve<3>(S, eps, rhs, max_iter, wd, wt, s)
                  : cg_solve<6>(S, eps, rhs, max_iter, wd, wt, s);
  if (rc) return rc;
  JX_HIP(hipStreamSynchronize(s));
  int it = S->st_host->iter;
  if (iters) *iters = it;
  if (status) *status = S->st_host->done ? S->st_host->done : JX_SOLVE_MAXITER;
  if (res_hist && it > 0)
    JX_HIP(hipMemcpy(res_hist, S->res_dev, sizeof(double) * it,
                     hipMemcpyDeviceToHost));
  return JX_OK;
}

int jx_solver_solve_cg_polarization(jx_solver *S, double *tau, double *eps,
                                    const double *E, double tol, int max_iter,
                                    double res_scale, double *res_hist,
                                    int *iters, int *status, void *stream) {
  if (!S || !tau || !eps || !E)
    return fail(JX_EINVAL, "jx_solver_solve_cg_polarization: NULL argument");
  JX_ALIGNED(tau, "jx_solver_solve_cg_polarization");
  JX_ALIGNED(eps, "jx_solver_solve_cg_polarization");
  if (tau == eps) return fail(JX_EINVAL, "tau and eps must not overlap");
  if (max_iter < 1) return fail(JX_EINVAL, "max_iter must be >= 1");
  if (!S->local->symmetric)
    return fail(JX_EINVAL, "conjugate gradients need a symmetric local operator");
  hipStream_t s = (hipStream_t)stream;
  const Grid &g = S->green->g;
  const int C = S->green->sym;
  if (res_scale <= 0.) {
    double nE = 0.;
    for (int c = 0; c < C; ++c) nE += E[c] * E[c];
    nE = std::sqrt(nE);
    if (nE == 0.) return fail(JX_EINVAL, "res_scale must be given when E = 0");
    res_scale = 1.0 / std::sqrt((double)g.N) / nE;
  }
  if (!S->cg_buf) {
    JX_HIP(hipMalloc((void **)&S->cg_buf, sizeof(double) * 3 * C * g.N));
    JX_HIP(hipMalloc((void **)&S->cg_partials, sizeof(double) * 2 * kCgBlocks));
    JX_HIP(hipMalloc((void **)&S->cg_s, sizeof(CgScalars)));
  }
  if (S->res_cap < max_iter) {
    if (S->res_dev) (void)hipFree(S->res_dev);
    S->res_dev = nullptr;
    S->res_cap = 0;
    JX_HIP(hipMalloc((void **)&S->res_dev, sizeof(double) * max_iter));
    S->res_cap = max_iter;
  }
  SolveState init = {};
  init.tol = tol;
  init.res_scale = res_scale;
  init.max_iter = max_iter;
  for (int c = 0; c < 6; ++c) init.E[c] = c < C ? E[c] : 0.;
  *S->st_host = init;
  JX_HIP(hipMemcpyAsync(S->st, S->st_host, sizeof(SolveState),
                        hipMemcpyHostToDevice, s));
  int rc = C == 3 ? cg_pol_solve<3>(S, tau, eps, max_iter, s)
                  : cg_pol_solve<6>(S, tau, eps, max_iter, s);
  if (rc) return rc;
  JX_HIP(hipStreamSynchronize(s));
  int it = S->st_host->iter;
  if (iters) *iters = it;
  if (status) *status = S->st_host->done ? S->st_host->done : JX_SOLVE_MAXITER;
  if (res_hist && it > 0)
    JX_HIP(hipMemcpy(res_hist, S->res_dev, sizeof(double) * it,
                     hipMemcpyDeviceToHost));
  return JX_OK;
}

}  // extern "C"
