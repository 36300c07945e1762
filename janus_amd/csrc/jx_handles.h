// Handle structures behind the opaque pointers of include/janus_amd.h.
#pragma once

#include <vector>

#include "jx_internal.h"

struct jx_fft {
  jx::Grid g;
  double *scratch = nullptr;  // spectrum copy for c2r (input is preserved)
};

struct jx_green {
  jx::Grid g;
  int sym = 0;
  double h = 1.0;
  double g0 = 0, nu0 = 0;
  jx::GreenDev dev;
  double *tables = nullptr;   // device per-axis tables
  int device = -1;            // device holding tables (and spec)
  std::vector<double> host_tables;
  jx::GreenDev host;          // same tables, host pointers
  double *spec = nullptr;     // spectrum scratch (Nh * 2 sym)
};

struct jx_local {
  jx::Grid g;
  int sym = 0;
  jx::LocalDev dev;
  double *table = nullptr;     // device (tr, dev) per phase, or matrices
  double *partials = nullptr;  // device, 1024 blocks x 2 sym columns + 2 sym
  bool symmetric = true;       // every block is a symmetric matrix
};

struct jx_solver {
  jx_green *green = nullptr;
  jx_local *local = nullptr;
  bool fused = true;  // pass E of one iteration fused with pass A of the next
  jx::SolveState *st = nullptr;       // device
  jx::SolveState *st_host = nullptr;  // pinned host mirror
  double *partials = nullptr;
  int nblocks = 0;
  double *res_dev = nullptr;
  int res_cap = 0;
  // conjugate gradients (allocated on first use): r, p, q fields, block
  // partial sums, scalars
  double *cg_buf = nullptr;
  double *cg_partials = nullptr;
  jx::CgScalars *cg_s = nullptr;
};

namespace jx {
// Device tables, spectrum scratch and FFT plans of a Green operator, set up
// on first use.
int green_ensure_device(jx_green *G);
// One fixed-point iteration of solver S on iterate eps, spectrum buffer
// spec and solve state st (jx_api.hip).
int solver_iteration(jx_solver *S, double *eps, double *spec, SolveState *st,
                     const RowsInverseLs &ls, bool first, int *nb,
                     hipStream_t s);
// y = Gamma0[dC : x] with the operators of solver S (five passes, the last
// one a plain scaled c2r); x and y are fields of the grid (jx_cg.hip).  loc
// replaces the solver's local operator (a default LocalDev: y = Gamma0[x]).
// spec replaces the Green operator's spectrum scratch.
int ls_matvec(jx_solver *S, const double *x, double *y, const SolveState *st,
              hipStream_t s, const LocalDev *loc = nullptr,
              double *spec = nullptr);

// Work buffers of one strain-form conjugate-gradient solve (jx_cg.hip).
struct CgWork {
  double *r = nullptr, *p = nullptr, *q = nullptr;  // fields of the grid
  double *spec = nullptr;           // spectrum scratch
  double *rows_partials = nullptr;  // block sums of the fused pass E
  double *vec_partials = nullptr;   // block sums of the vector kernels
  CgScalars *cs = nullptr;
  SolveState *st = nullptr;
  double *res_hist = nullptr;
};
// Doubles of CgWork::vec_partials.
int64_t cg_vec_partials_doubles();
// C0 inner-product weights 2 g0 and k0 - 2 g0 / dim of the solver's
// reference medium.
void cg_weights(const jx_green *G, double *wd, double *wt);
// x = 0 in eps, r = p = rhs - Gamma0[dC : E], the first residual.
int cg_start(jx_solver *S, const CgWork &w, double *eps, const double *rhs,
             double wd, double wt, hipStream_t s);
// eps = E + x + alpha p: the last step the fused iterations leave.
int cg_finish(jx_solver *S, const CgWork &w, double *eps, double wd,
              double wt, hipStream_t s);
// True when batched conjugate gradients run on the grid (the fused
// iteration, register FFT lengths on every axis).
bool cg_batchable(const Grid &g, int C);
// One fused iteration of the batched conjugate-gradient entries b (their
// slots hold the fields; register FFT lengths on every axis), with the
// kernels and block decomposition of the single solve.
int cg_iteration_batched(const jx_green *G, const LocalDev &loc,
                         const Batch &b, hipStream_t s);
}  // namespace jx
