/*
 * janus_amd -- C ABI of the MI355X (gfx950) Lippmann-Schwinger hot path.
 *
 * Plain C: opaque handles, device pointers, sizes and int status codes.  No
 * torch types cross this boundary.  Every compute call is stream ordered on
 * the hipStream_t passed as `void *stream` (NULL = the null stream).  All
 * device arrays are FP64, C order, in janus's layouts:
 *
 *   real field      (n0, n1[, n2], sym)               sym strain components
 *   scalar field    (n0, n1[, n2])                    FFT input / output
 *   spectrum        (n0, n1[, n2'])                   n_last' = 2 (n_last/2 + 1),
 *                   Re and Im interleaved along the last axis
 *                   (janus/fft/serial/_serial_fft.pyx:57-74, 181-200).
 * The Green operator's spectrum scratch is internal (per frequency, the sym
 * complex components side by side) and never crosses this boundary.
 *
 * sym = 3 (2D, Mandel order xx, yy, sqrt2 xy) or 6 (3D, xx, yy, zz, sqrt2 yz,
 * sqrt2 zx, sqrt2 xy) -- janus/mandelvoigt.py:26-40.
 *
 * Device arrays passed to the compute calls must be 16-byte aligned (any
 * hipMalloc or torch allocation is); a misaligned one gets JX_EINVAL.
 *
 * Status codes: JX_OK (0) or a negative JX_E* value; jx_last_error() returns
 * a thread-local message for the last failure of the calling thread.
 * Handles are not thread safe (they own scratch memory), as in janus.
 */
#ifndef JANUS_AMD_H
#define JANUS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JX_OK 0
#define JX_EINVAL (-1)       /* invalid argument (maps to ValueError) */
#define JX_EHIP (-2)         /* HIP runtime error */
#define JX_ENOMEM (-3)       /* device allocation failed */
#define JX_EUNSUPPORTED (-4) /* size or mode this build does not handle */

#define JX_GREEN_TRUNCATED 0
#define JX_GREEN_FILTERED 1
#define JX_GREEN_WILLOT 2 /* janus/green.pyx:976-1036 willot2015 */

#define JX_SOLVE_CONVERGED 1 /* residual < tol */
#define JX_SOLVE_MAXITER 2   /* max_iter iterations without reaching tol */
#define JX_SOLVE_NONFINITE 3 /* residual became NaN or Inf */

typedef struct jx_fft jx_fft;
typedef struct jx_green jx_green;
typedef struct jx_local jx_local;
typedef struct jx_solver jx_solver;
typedef struct jx_batch jx_batch;

/* Message describing the last failure on this thread ("" if none). */
const char *jx_last_error(void);
/* ABI version of this library. */
int jx_version(void);

/* ---------------------------------------------------------------------
 * Real FFT.  Replaces janus.fft.serial.create_real and _RealFFT2D/3D
 * (janus/fft/serial/_serial_fft.pyx:20-53 create_real/create_real_2D/3D,
 * :57-74 and :181-200 shapes, :152-163 r2c, :165-177 c2r).
 * r2c is unnormalized, c2r scales by 1/(n0 n1 [n2]).  Inputs are preserved.
 * ------------------------------------------------------------------- */
int jx_fft_create(int dim, const int64_t *shape, jx_fft **out);
int jx_fft_destroy(jx_fft *fft);
/* r: (n0, n1[, n2]) real; c: (n0, n1[, ...], 2 (n_last/2 + 1)). */
int jx_fft_r2c(jx_fft *fft, const double *r, double *c, void *stream);
int jx_fft_c2r(jx_fft *fft, const double *c, double *r, void *stream);

/* ---------------------------------------------------------------------
 * Discrete Green operator for strains of an isotropic reference material
 * (g0, nu0).  Replaces janus.green.truncated / filtered
 * (janus/green.pyx:22-36) with DiscreteGreenOperator2D/3D.apply
 * (:182-216, :353-390, truncated :453-495, :531-580, filtered
 * :607-694, :730-973) and the continuous operator of
 * janus/material/elastic/linear/isotropic.pyx:133-156, 173-200, 242-308.
 * kind = JX_GREEN_TRUNCATED or JX_GREEN_FILTERED; h > 0 is the cell size.
 * ------------------------------------------------------------------- */
int jx_green_create(int dim, const int64_t *shape, double h, int kind,
                    double g0, double nu0, jx_green **out);
int jx_green_destroy(jx_green *green);
/* y = Gamma0[x]; x, y: (n0, n1[, n2], sym) device arrays; x == y allowed. */
int jx_green_apply(jx_green *green, const double *x, double *y, void *stream);
/* Allocation-free variant of jx_green_apply (no counterpart in janus: it is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results; the spectrum scratch is a file-scope static device
 * array of 64 MiB that each GPU holds until the process ends, instead of
 * the handle's own buffer.  The call makes no hipMalloc, hipFree or other
 * allocating call: the grid's spectrum, 16 sym n0 n1 [n2] / n_last (n_last / 2
 * + 1) bytes (in 3D plus 8448 bytes after each n0 slab whose size is a
 * multiple of 4 KB), must fit the 64 MiB (3D up to about 100^3, 2D up to
 * 1024^2), else JX_EINVAL; the handle must have been created with a GPU
 * present and this device current (its tables and FFT plans are made by
 * jx_green_create then), else JX_EINVAL.  Every call on one GPU shares the
 * array: calls that overlap on one GPU (two streams, two threads) race and
 * must be serialized by the caller. */
int jx_green_apply_static(jx_green *green, const double *x, double *y,
                          void *stream);
/* Host-side per-frequency matrix Gamma-hat(b), b in [0, n_i), written to
 * out[sym * sym] (row major).  Replaces set_frequency + to_memoryview
 * (janus/green.pyx:146-171, 312-346). */
int jx_green_matrix_at(jx_green *green, const int32_t *b, double *out);
/* Continuous Gamma0(k) of the material (g, nu) in dimension dim, written to
 * out[sym * sym]; replaces _GreenOperatorForStrains*.set_frequency +
 * to_memoryview (isotropic.pyx:173-200, 242-308, 213-224, 333-371). */
int jx_gamma0_at(int dim, double g, double nu, const double *k, double *out);

/* ---------------------------------------------------------------------
 * Block-diagonal isotropic local operator x -> (sph J + dev K) : x, voxel
 * by voxel.  Replaces janus.operators.BlockDiagonalOperator2D/3D whose loc
 * entries are isotropic_4(sph, dev, dim) (janus/operators.pyx:53-67,
 * 292-299, 315-324, 352-365, 617-716).  The handle keeps the device
 * pointers it is given (a shallow reference, as operators.pyx:76-77 keeps
 * loc); the caller keeps them alive.
 * ------------------------------------------------------------------- */
/* phase: (n0, n1[, n2]) uint8 device array; table: host array of nphase
 * (sph, dev) pairs, nphase in [1, 256]. */
int jx_local_create_phases(int dim, const int64_t *shape, const uint8_t *phase,
                           const double *table, int nphase, jx_local **out);
/* sph, dev: (n0, n1[, n2]) device arrays of per-voxel coefficients. */
int jx_local_create_fields(int dim, const int64_t *shape, const double *sph,
                           const double *dev, jx_local **out);
int jx_local_destroy(jx_local *local);
/* y = local : x; x, y: (n0, n1[, n2], sym); x == y allowed. */
int jx_local_apply(jx_local *local, const double *x, double *y, void *stream);
/* mean over the grid of local : x (out_tau[sym]) and of x (out_eps[sym]),
 * written to host memory (synchronizes the stream).  Either may be NULL. */
int jx_local_means(jx_local *local, const double *x, double *out_tau,
                   double *out_eps, void *stream);

/* ---------------------------------------------------------------------
 * Anisotropic local operators: a general sym x sym matrix per voxel, as a
 * janus BlockDiagonalOperator2D/3D whose loc entries are any sym x sym
 * linear operators (e.g. FourthRankCubicTensor2D, operators.pyx:411-450)
 * or a square BlockDiagonalLinearOperator2D/3D (operators.pyx:719-890).
 * y_i = sum_j a_ij x_j (operators.pyx:739-756); the handle works wherever
 * a jx_local does: jx_local_apply, jx_local_means, jx_solver_*, jx_batch_*.
 * Conjugate gradients (jx_solver_solve_cg) need symmetric matrices and
 * fail with JX_EINVAL otherwise.
 * ------------------------------------------------------------------- */
/* phase: (n0, n1[, n2]) uint8 device array; mats: host array of nphase
 * row-major sym x sym matrices, nphase in [1, 256] (copied). */
int jx_local_create_phase_matrices(int dim, const int64_t *shape,
                                   const uint8_t *phase, const double *mats,
                                   int nphase, jx_local **out);
/* mats: (n0, n1[, n2], sym, sym) device array, kept by reference. */
int jx_local_create_matrix_field(int dim, const int64_t *shape,
                                 const double *mats, jx_local **out);
/* *out = 1 when every block of the operator is a symmetric matrix. */
int jx_local_is_symmetric(jx_local *local, int *out);
/* y = local^T : x, y_i = sum_j a_ji x_j (operators.pyx:758-775, 843-861);
 * the isotropic modes are symmetric.  x == y allowed. */
int jx_local_apply_transpose(jx_local *local, const double *x, double *y,
                             void *stream);
/* Block-diagonal linear operator with blocks of any size m x n
 * (BlockDiagonalLinearOperator2D/3D, operators.pyx:719-890): a is the
 * (n0, n1[, n2], m, n) device array; transpose = 0: x (.., n) -> y (.., m),
 * y_i = sum_j a_ij x_j; transpose = 1: x (.., m) -> y (.., n),
 * y_i = sum_j a_ji x_j.  x and y must not overlap. */
int jx_block_linear_apply(int dim, const int64_t *shape, int m, int n,
                          const double *a, const double *x, double *y,
                          int transpose, void *stream);

/* ---------------------------------------------------------------------
 * Basic (Moulinec-Suquet) fixed-point solver, the loop of
 * sphinx/tutorials/square_basic/square_basic.py:83-92:
 *     eps <- E + rhs - Gamma0[local : eps]
 * with residual ||eps_{k+1} - eps_k|| * res_scale; the default res_scale
 * (pass 0) is 1 / (sqrt(N) ||E||) as in square_basic.py:86, 91.  Stops
 * after the first iteration whose residual is < tol, or after max_iter.
 * The fused passes keep Gamma0 and the local operator out of memory.
 * ------------------------------------------------------------------- */
int jx_solver_create(jx_green *green, jx_local *local, jx_solver **out);
int jx_solver_destroy(jx_solver *solver);
/* eps: device field (n0, n1[, n2], sym), holds the initial iterate on entry
 * (e.g. E everywhere) and the last iterate on return.  E: host sym vector.
 * rhs: optional device field added to E (NULL for none).  res_hist: host
 * array of max_iter doubles (may be NULL).  iters, status: host outputs.
 * Synchronizes the stream. */
int jx_solver_solve(jx_solver *solver, double *eps, const double *E,
                    const double *rhs, double tol, int max_iter,
                    double res_scale, double *res_hist, int *iters,
                    int *status, void *stream);
/* Conjugate gradients on the same operators (no janus counterpart in the
 * library; janus's examples/square_inclusion.py:40-47, 71-86 drives PETSc CG
 * on the polarization form).  Solves (I + Gamma0 dC) x = -Gamma0[dC : E]
 * for the fluctuation x = eps - E in the inner product sum a : C0 : b, in
 * which the operator is self-adjoint (truncated operator; close to it for
 * the filtered one) and positive definite for any isotropic C0 when C is:
 * it converges where the basic scheme diverges (C0 = the matrix).  The
 * residual after iteration k is ||b - A x_k|| / (sqrt(N) ||E||) (res_scale
 * overrides the 1 / (sqrt(N) ||E||)), the basic scheme's own residual
 * definition (square_basic.py:86, 91); the solve stops after the first
 * iteration below tol.  eps: device field, output only (the iteration
 * starts from E).  rhs: optional device field added to E, as in
 * jx_solver_solve (eps solves (I + Gamma0 dC) eps = E + rhs); it must be
 * compatible and zero-mean (a Gamma0 output, as in the adjoint solve).  It
 * enters the initial residual only.  Other arguments as jx_solver_solve.
 * Allocates three fields of the grid on first use.  Synchronizes the
 * stream. */
int jx_solver_solve_cg(jx_solver *solver, double *eps, const double *E,
                       const double *rhs, double tol, int max_iter,
                       double res_scale, double *res_hist, int *iters,
                       int *status, void *stream);
/* Conjugate gradients on the polarization form, as janus's
 * examples/square_inclusion.py:17-52, 77-85 drives PETSc CG: the solver's
 * local operator L is the inverse of C - C0 (the example's tau2eps,
 * isotropic_4(1 / (dim dk), 1 / (2 dg))), and tau solves (L + Gamma0) tau =
 * E in the Euclidean inner product, from tau = 0; eps = L tau is written
 * too.  L + Gamma0 is definite when C0 lies strictly below (or above) every
 * phase.  The residual after iteration k is ||E - (L + Gamma0) tau_k|| /
 * (sqrt(N) ||E||), PETSc's relative residual with PCNONE.  tau, eps:
 * device fields, output only, not overlapping.  Other arguments as
 * jx_solver_solve_cg; shares its buffers. */
int jx_solver_solve_cg_polarization(jx_solver *solver, double *tau,
                                    double *eps, const double *E, double tol,
                                    int max_iter, double res_scale,
                                    double *res_hist, int *iters, int *status,
                                    void *stream);
/* Restarted GMRES on the same operators (no janus counterpart in the
 * library; janus's examples/square_inclusion_scipy.py:52-59, 88 hands the
 * strain form (I + Gamma0 dC) x = -Gamma0[dC : E] to scipy's GMRES).  Same
 * equation, arguments, rhs and residual definition as jx_solver_solve_cg,
 * but in the Euclidean inner product and with no symmetry needed: any
 * sym x sym local operator, including non-symmetric matrix fields.
 * restart in [1, 256]: Arnoldi steps per cycle; the iterate is updated and
 * the residual recomputed after each cycle.  The residual after iteration k
 * is the Arnoldi estimate of ||b - A x_k|| / (sqrt(N) ||E||).  Allocates
 * restart + 3 fields of the grid for the call and frees them before it
 * returns; reads one Hessenberg column back per iteration.  Synchronizes
 * the stream. */
int jx_solver_solve_gmres(jx_solver *solver, double *eps, const double *E,
                          const double *rhs, double tol, int max_iter,
                          int restart, double res_scale, double *res_hist,
                          int *iters, int *status, void *stream);
/* Schedule choice (no janus counterpart).  fused = 1 (default): the last
 * pass of an iteration also applies the local operator to the new iterate
 * and transforms it for the next iteration, so the iterate is written once
 * and not read back; fused = 0: five separate passes per iteration.  Both
 * give the same iterates up to rounding.  get returns the schedule actually
 * used (the fused pass needs a last-axis length with a register FFT:
 * 16, 32, 64, 128, 256, 512, 1024, 80, 96 or 160). */
int jx_solver_set_fused(jx_solver *solver, int fused);
int jx_solver_get_fused(jx_solver *solver);

/* Register FFT plans (no janus counterpart).  Lengths 64, 128, 256 and 512
 * have two plans: variant 0 holds 16 values per thread, variant 1 holds 8
 * (fewer registers and more waves per SIMD for one more LDS exchange).
 * Each pass of an operator picks its own: pass 0 last-axis r2c, 1 axis-1
 * transforms, 2 axis-0 transforms with Gamma-hat, 3 last-axis c2r (with the
 * update), 4 the fused update pass.  Variants change rounding, not the
 * operator.  Defaults come from timings at 256^3; jx_solver_autotune
 * measures them for the solver's grid. */
int jx_green_set_plan(jx_green *green, int pass, int variant);
int jx_green_get_plan(jx_green *green, int pass);
/* Times every pass of the solver's grid with each plan variant (`iters`
 * launches each, on scratch fields) and keeps the fastest in the solver's
 * Green operator, then picks the fused schedule when its pass is faster
 * than the separate last-axis passes.  ms_out: NULL or 10 doubles, the
 * average time of pass p with variant k at [2 p + k] (0 when not run).
 * Synchronizes the stream. */
int jx_solver_autotune(jx_solver *solver, int iters, double *ms_out,
                       void *stream);

/* ---------------------------------------------------------------------
 * Batched solves (no janus counterpart: janus runs one fixed point at a
 * time from user scripts, sphinx/tutorials/square_basic/square_basic.py:
 * 83-97).  A batch holds many right-hand sides -- the load cases of one
 * solver, solvers on different grids -- and runs them all to their stop in
 * one sequence of launches.  Entries on the same grid, Green operator kind
 * and h, schedule and rhs use form a group; `width` entries of a group
 * share each kernel launch when every axis has a register FFT length (else
 * one launch per entry).  Every replay runs iters_per_launch iterations of
 * each chunk that still has a running entry; after the first (eager)
 * replay the replays are captured in a hipGraph, captured again only when
 * a chunk has finished.  Each entry stops exactly as jx_solver_solve would
 * and gets the same iterates and iteration count.
 * ------------------------------------------------------------------- */
int jx_batch_create(jx_batch **out);
int jx_batch_destroy(jx_batch *batch);
/* One right-hand side of `solver`; the arguments are those of
 * jx_solver_solve (eps: the entry's own device field, initial iterate in,
 * last iterate out; it must stay valid while the batch runs).  Allocates the
 * entry's spectrum, partial sums and residual history on the device.
 * index: host output, the entry's index (0, 1, ... in order of addition). */
int jx_batch_add(jx_batch *batch, jx_solver *solver, double *eps,
                 const double *E, const double *rhs, double tol, int max_iter,
                 double res_scale, int *index);
/* The same entry solved by conjugate gradients: the arguments, the result
 * (eps = E + x; the content of eps on entry is not used, so a second run
 * solves again from E), the iterates, residual history and iteration count
 * of jx_solver_solve_cg (janus drives CG from its examples,
 * examples/square_inclusion.py:40-47, 71-86).  Needs a symmetric local
 * operator (else JX_EINVAL) and register FFT lengths on every axis (else
 * JX_EUNSUPPORTED).  Also allocates the entry's r, p, q fields, partial
 * sums and scalars.  Entries of both methods can share a batch; they never
 * share a launch. */
int jx_batch_add_cg(jx_batch *batch, jx_solver *solver, double *eps,
                    const double *E, const double *rhs, double tol,
                    int max_iter, double res_scale, int *index);
/* Groups in order of first appearance: grid (dim, shape[3], 1 for unused
 * axes), entry count, entries per launch (after jx_batch_run: the width
 * used) and whether the group's launches are batched (1) or per entry. */
int jx_batch_ngroups(jx_batch *batch, int *ngroups);
int jx_batch_group(jx_batch *batch, int group, int *dim, int64_t *shape,
                   int *size, int *width, int *batched);
/* Autotuning of the last run for a group: *n candidate widths and their
 * measured ms per entry-iteration (1e30 when no entry ran), at most cap of
 * them copied. */
int jx_batch_tuning(jx_batch *batch, int group, int cap, int *n, int *widths,
                    double *ms);
/* Entries per launch of a group (0 = the whole group). */
int jx_batch_set_width(jx_batch *batch, int group, int width);
/* iters_per_launch >= 1 (default 8); graph = 1 (default) replays a
 * captured hipGraph; autotune = 1: after the first replay, one timed replay
 * per candidate width (all, 12, 6, 3, 2, 1 entries per launch) of every
 * group whose width was not set, then the width with the least time per
 * entry-iteration run is kept (default 0: whole groups). */
int jx_batch_set_schedule(jx_batch *batch, int iters_per_launch, int graph,
                          int autotune);
/* Concurrent groups (no janus counterpart): nstreams in [1, 8] (default 1).
 * Above 1, each enqueued replay (graph = 0) after the autotuning ones forks
 * the groups onto up to nstreams streams (groups balanced by running
 * entries times voxels) and joins them back, so that the launches of small
 * grids run beside those of large ones.  Captured graphs (graph = 1) keep
 * one stream.  Groups share no fields: every entry's iterates are
 * unchanged. */
int jx_batch_set_streams(jx_batch *batch, int nstreams);
/* Runs every entry to its stop on an internal stream after synchronizing
 * `stream`; returns when all have stopped.  A batch may be run again: each
 * entry restarts from the content of its eps. */
int jx_batch_run(jx_batch *batch, void *stream);
/* Result of entry `index` of the last run: iteration count, status and the
 * residual history (host array of iters doubles, may be NULL). */
int jx_batch_result(jx_batch *batch, int index, double *res_hist, int *iters,
                    int *status);
/* Wall time of the last run in ms, number of replays and of graph
 * captures. */
int jx_batch_stats(jx_batch *batch, double *ms, int *replays, int *captures);

/* Adjoint contraction (no janus counterpart; implicit differentiation of
 * the fixed point of jx_solver_solve).  With eps the converged strain and
 * mu the solution of the adjoint fixed point mu <- G g - G[local : mu]
 * (g = dL/deps, G = Gamma0), the gradient with respect to the per-voxel
 * coefficients of local = sph J + dev K is
 *     dsph = -(J eps) . mu = -tr(eps) tr(mu) / dim,
 *     ddev = -(K eps) . mu = -(eps . mu) + tr(eps) tr(mu) / dim.
 * eps, mu: (n0, n1[, n2], sym) device arrays; dsph, ddev: (n0, n1[, n2]). */
int jx_adjoint_contract(int dim, const int64_t *shape, const double *eps,
                        const double *mu, double *dsph, double *ddev,
                        void *stream);
/* Adjoint contraction for a per-voxel matrix local operator a (no janus
 * counterpart): with eps and mu as in jx_adjoint_contract, except that mu
 * solves the adjoint fixed point with the transposed matrices, mu <- G g -
 * G[a^T : mu], the gradient is da[v][i][j] = -mu[v][i] eps[v][j].  eps, mu:
 * (n0, n1[, n2], sym); da: (n0, n1[, n2], sym, sym) device arrays. */
int jx_adjoint_outer(int dim, const int64_t *shape, const double *eps,
                     const double *mu, double *da, void *stream);

/* Measurement helper (no janus counterpart): device-to-device copy of
 * `bytes` (a multiple of 16; 16-byte aligned pointers) with one 16-byte load
 * and store per lane -- the reference for the practical HBM bandwidth and
 * for the calibration of traffic counters. */
int jx_stream_copy(const void *src, void *dst, int64_t bytes, void *stream);

/* Measurement helper (no janus counterpart): the memory traffic of the
 * strided transform passes without their arithmetic.  Copies the spectrum of
 * the 3D grid `shape` (6 components, the solvers' padded layout) from src to
 * dst along `axis` (0 or 1; that axis has 256 or 512 points) in tiles of
 * `points` points (1, 2 or 4), in the passes' tile and lane order.  src and
 * dst are 16-byte aligned and hold `bytes` each, at least the layout's size
 * from jx_tile_probe_bytes. */
int jx_tile_probe_bytes(const int64_t *shape, int64_t *bytes);
int jx_tile_probe(const int64_t *shape, int axis, int points, const void *src,
                  void *dst, int64_t bytes, void *stream);

/* Measurement helper (no janus counterpart): runs `iters` fixed-point
 * iterations exactly as jx_solver_solve does (tol ignored) after one untimed
 * iteration, and records a hipEvent before and after every pass on
 * `stream`.  ms_out[6] receives the average duration in milliseconds of the
 * passes: 0 last-axis r2c with the local operator (0 in the fused
 * schedule), 1 axis-1 forward (3D), 2 axis-0 forward + Gamma-hat + inverse,
 * 3 axis-1 inverse (3D), 4 last-axis c2r with the fixed-point update (fused:
 * plus the local operator and r2c of the next iteration), 5 residual
 * reduction.  Synchronizes the stream. */
int jx_solver_profile(jx_solver *solver, double *eps, const double *E,
                      int iters, double *ms_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* JANUS_AMD_H */
