// Internal declarations shared by the janus_amd HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/janus_amd.h"

namespace jx {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

// Device arrays crossing the C ABI must be 16-byte aligned: the kernels
// move complex values and component pairs with 16-byte loads and stores.
#define JX_ALIGNED(p, what)                                             \
  do {                                                                  \
    if ((p) != nullptr && (reinterpret_cast<uintptr_t>(p) & 15))        \
      return jx::fail(JX_EINVAL, std::string(what) +                    \
                                     ": device arrays must be 16-byte " \
                                     "aligned");                        \
  } while (0)

#define JX_HIP(call)                                  \
  do {                                                \
    hipError_t jx_e_ = (call);                        \
    if (jx_e_ != hipSuccess) return jx::hip_fail(jx_e_, #call); \
  } while (0)

// ---------------------------------------------------------------- limits
constexpr int kMaxStages = 24;   // radix stages per FFT length
constexpr int kMaxDim = 3;
constexpr int kThreads = 256;    // threads per workgroup of the FFT kernels

// One FFT length: radix sequence and twiddle table W[m] = exp(-2 pi i m / L).
struct FftLen {
  int L = 0;
  int nstages = 0;
  int radix[kMaxStages] = {};
  const double2 *W = nullptr;  // device, L entries
};

// Per-axis tables of the discrete Green operators (device pointers).
//   kt: truncated wave number  s (b - n) if 2b > n else s b   (green.pyx:432-445)
//   km, kp: filtered aliases    s (b - n), s b                 (green.pyx:740-800)
//   wm, wp: filtered weights    cos^2(h k / 4) of km, kp
//   sw, cw: Willot's sin, cos of phi = s h b / 2              (green.pyx:984-1023)
struct AxisTab {
  const double *kt = nullptr, *km = nullptr, *kp = nullptr, *wm = nullptr,
               *wp = nullptr, *sw = nullptr, *cw = nullptr;
};

struct GreenDev {
  int dim = 0;
  int kind = 0;  // JX_GREEN_TRUNCATED / JX_GREEN_FILTERED / JX_GREEN_WILLOT
  double daux1 = 0, daux2 = 0, daux3 = 0, daux4 = 0;
  AxisTab ax[kMaxDim];
};

// Local operator modes: isotropic tensors sph J + dev K from a phase map and
// a (tr, dev) table or from per-voxel fields; general C x C matrices from a
// phase map and a matrix table or from a per-voxel matrix field.
enum LocalMode {
  LOCAL_NONE = 0,
  LOCAL_PHASES = 1,
  LOCAL_FIELDS = 2,
  LOCAL_PHASE_MATRIX = 3,
  LOCAL_MATRIX_FIELD = 4
};

struct LocalDev {
  int mode = LOCAL_NONE;
  const uint8_t *phase = nullptr;
  const double *table = nullptr;  // device, 2 * nphase (tr, dev) pairs
  const double *sph = nullptr;
  const double *dev = nullptr;
  const double *mat = nullptr;    // device, C * C per phase or per voxel
};

// Families of local operators that the register-heavy kernels are compiled
// for separately, so that the isotropic kernels keep their register budget:
// LF_ISO covers LOCAL_PHASES and LOCAL_FIELDS, LF_MATRIX the matrix modes.
enum LocalFamily { LF_ISO = 0, LF_MATRIX = 1 };

__host__ __device__ inline int local_family(const LocalDev &loc) {
  return loc.mode >= LOCAL_PHASE_MATRIX ? LF_MATRIX : LF_ISO;
}

// Device-side state of a running solve.
struct SolveState {
  int iter;      // completed iterations
  int done;      // 0 running, else JX_SOLVE_* status
  double tol;
  double res_scale;
  int max_iter;
  int pad;
  double E[6];   // macroscopic strain (Mandel), read by the update pass
};

// One right-hand side of a batched solve.  A batched launch covers n
// consecutive entries of a device table of slots, all on the same grid;
// each block finds its entry from its index and takes every per-entry
// pointer from the slot.
struct CgScalars;
struct Slot {
  double *eps = nullptr;        // iterate, real field (rows, L, C)
  double *spec = nullptr;       // spectrum (rows, nh, C, 2)
  const double *rhs = nullptr;  // optional field added to E
  double *partials = nullptr;   // residual partial sums, one per block
  double *res_hist = nullptr;   // residual history (max_iter entries)
  SolveState *st = nullptr;
  LocalDev loc;
  double daux[4] = {};          // Green operator constants of C0
  // conjugate-gradient entries: residual, direction and A p fields, their
  // scalars, the vector kernel's partial sums and the C0 inner product
  // weights (eps holds the fluctuation x while the solve runs)
  double *cg_r = nullptr, *cg_p = nullptr, *cg_q = nullptr;
  CgScalars *cg_cs = nullptr;
  double *cg_partials = nullptr;
  double cg_wd = 0., cg_wt = 0.;
};

struct Batch {
  const Slot *tab = nullptr;  // device
  int n = 0;
};

// Passes that choose a register FFT plan variant (jx_fft_reg.h).
enum PlanPass {
  PASS_ROWS_FWD = 0,  // A: last-axis r2c (+ local operator)
  PASS_AXIS = 1,      // B, D: axis-1 transforms
  PASS_GAMMA = 2,     // C: axis-0 transforms with Gamma-hat
  PASS_ROWS_INV = 3,  // E: last-axis c2r (+ update)
  PASS_FUSED = 4,     // E + A fused
  kPlanPasses = 5
};

// Geometry of a grid with ncomp components per point, and the register FFT
// plan variant each pass uses on it.
struct Grid {
  int dim = 0;
  int64_t n[kMaxDim] = {1, 1, 1};
  int64_t N = 0;      // voxels
  int64_t rows = 0;   // N / n_last
  int64_t L = 0;      // n_last
  int64_t nh = 0;     // n_last / 2 + 1
  int64_t Nh = 0;     // rows * nh spectral points
  int plan[kPlanPasses] = {0, 0, 0, 0, 0};
};

Grid make_grid(int dim, const int64_t *shape);

// The strided axis whose pass applies Gamma-hat between its forward and
// inverse transforms (the pivot, transformed last and inverted first), and
// in 3D the other strided axis, whose passes only transform.  Either order
// gives the same operator; the pivot's pass is the heaviest of the
// iteration, and on axis 1 its tiles stay within one axis-0 slab.
#ifndef JX_PIVOT_AXIS
#define JX_PIVOT_AXIS 0
#endif
inline int pivot_axis(const Grid &g) { return g.dim == 3 ? JX_PIVOT_AXIS : 0; }
inline int plain_axis(const Grid &g) { return 1 - pivot_axis(g); }

// Layout of the solver spectra (ncomp complex values per frequency): rows
// (i0[, i1]) of nh frequencies, C order.  In 3D each i0 slab of n1 rows is
// followed by `pad` unused doubles when its size is a multiple of 4 KB: the
// axis-0 transforms stride by whole slabs, and slab strides aligned to 4 KB
// (256^3: 387 x 8 KB; 512^3: 771 x 16 KB) concentrate their row accesses on
// few HBM channels -- a probe of the axis-0 pass's access pattern at 256^3
// moved 3.76 TB/s unpadded and 4.29 TB/s with 8448 bytes of padding
// (two-point tiles; four-point tiles 4.53 and 4.98).  ncomp = 1 (the FFT
// API's janus layout) is never padded.
struct SpecPad {
  int64_t slab_rows = 1;  // rows per padded slab (n1 in 3D)
  int64_t pad = 0;        // doubles after each slab
};
inline SpecPad spec_layout(const Grid &g, int ncomp) {
  SpecPad sp;
  if (g.dim != 3 || ncomp == 1) return sp;
  sp.slab_rows = g.n[1];
  const int64_t slab_bytes = g.n[1] * g.nh * ncomp * 16;
  sp.pad = (slab_bytes % 4096 == 0) ? 8448 / 8 : 0;
  return sp;
}
// Doubles of a spectrum of the grid (padding included).
inline int64_t spec_doubles(const Grid &g, int ncomp) {
  const SpecPad sp = spec_layout(g, ncomp);
  return 2 * ncomp * g.Nh + (sp.pad ? g.n[0] * sp.pad : 0);
}

// Twiddle / plan cache (per device, shared by all handles).
int get_fft_len(int L, FftLen *out);

// LDS limit per workgroup on the current device.
int lds_limit_bytes();

// ------------------------------------------------------------- launchers
// Every solver pass takes an optional batch: when b != nullptr the launch
// covers b->n right-hand sides and the pointer arguments (x, spec, y, loc,
// st, ls pointers) are taken from the slots instead.  Batched launches need
// register FFT lengths on every axis (batchable()).
bool batchable(const Grid &g, int ncomp);
// Forward transform of the last axis of a real field with optional local
// operator, writing the half spectrum.
int launch_rows_forward(const Grid &g, int ncomp, const double *x, double *spec,
                        const LocalDev &loc, const SolveState *st,
                        hipStream_t s, const Batch *b = nullptr);
// Complex transform along axis `axis` (< dim - 1) of a spectrum; dir = -1
// forward, +1 inverse, 0 = forward + Gamma-hat + inverse (axis must be 0).
int launch_axis(const Grid &g, int ncomp, int axis, int dir, const double *in,
                double *out, const GreenDev *green, const SolveState *st,
                hipStream_t s, const Batch *b = nullptr);
// Copy of a 3D 6-component padded spectrum along axis 0 or 1 in the strided
// passes' tile and lane order, without the transforms (jx_tile_probe).
int tile_probe(const Grid &g, int axis, int points, const double *src,
               double *dst, hipStream_t s);
// Scalars of a conjugate-gradient solve, kept on the device.
struct CgScalars {
  double rr;     // <r, r> in the C0 inner product
  double alpha;  // step of the current iteration
  double beta;   // direction update of the current iteration
};

// Fields of a conjugate-gradient solve that pass A updates before it
// transforms: p <- r + beta p and x <- x + alpha p (the previous p).
struct CgDev {
  const double *r = nullptr;
  double *p = nullptr;
  double *x = nullptr;
  const CgScalars *cs = nullptr;
};

// Inverse transform of the last axis (scaled by 1/N).  When enabled, the
// fixed-point update eps <- E + rhs - y (E from the solve state) is applied
// in place on y and per-block partial sums of (eps_new - eps_old)^2 go to
// partials.
struct RowsInverseLs {
  const double *rhs = nullptr;  // optional field
  double *partials = nullptr;
  bool enabled = false;
};
int launch_rows_inverse(const Grid &g, int ncomp, const double *spec, double *y,
                        const RowsInverseLs &ls, const SolveState *st,
                        int *nblocks_out, hipStream_t s,
                        const Batch *b = nullptr);
// Residual partial sums written by the update pass: the most blocks any
// plan variant of the last axis uses (the size of a partials buffer).
int rows_inverse_nblocks(const Grid &g, int ncomp);
// Conjugate-gradient variants of passes A and E for the register lengths
// (C = 3 or 6): pass A updates p and x (CgDev) and transforms dC : p; pass
// E writes q = p + c2r(spec) / N and the block sums of p : C0 : q
// (wd p . q + wt tr p tr q per voxel) into partials[block].  Batched: the
// fields, scalars, weights and state of each entry come from its slot (the
// CgDev, spectrum, q, p, partials and weights arguments are then unused).
bool rows_cg_fusable(const Grid &g, int ncomp);
int launch_rows_forward_cg(const Grid &g, int ncomp, const CgDev &cg,
                           double *spec, const LocalDev &loc,
                           const SolveState *st, hipStream_t s,
                           const Batch *b = nullptr);
int launch_rows_inverse_cg(const Grid &g, int ncomp, const double *spec,
                           double *q, const double *p, double *partials,
                           double wd, double wt, const SolveState *st,
                           int *nblocks_out, hipStream_t s,
                           const Batch *b = nullptr);
// Fused pass E of iteration k + pass A of iteration k + 1 (register path).
bool rows_fusable(const Grid &g, int ncomp);
int launch_rows_update_forward(const Grid &g, int ncomp, double *spec,
                               double *y, const RowsInverseLs &ls,
                               const LocalDev &loc, const SolveState *st,
                               int *nblocks_out, hipStream_t s,
                               const Batch *b = nullptr);
// Residual of the iteration from the partial sums; batched: one block per
// entry.
int launch_finalize(const double *partials, int nblocks, SolveState *st,
                    double *res_hist, hipStream_t s, const Batch *b = nullptr);
int launch_local_apply(const Grid &g, int ncomp, const LocalDev &loc,
                       const double *x, double *y, hipStream_t s,
                       bool transpose = false);
// *out = every one of the count C x C device matrices a is symmetric
// (synchronous; flag_dev: one device int of scratch).
int matrices_symmetric(int ncomp, const double *a, int64_t count, int *flag_dev,
                       bool *out);
// y = a x (or a^T x) voxel by voxel for a per-voxel (m, n) matrix field of
// any size (x, y must not overlap).
int launch_block_linear(int64_t N, int m, int n, const double *a,
                        const double *x, double *y, bool transpose,
                        hipStream_t s);
int launch_adjoint_contract(const Grid &g, int ncomp, const double *eps,
                            const double *mu, double *dsph, double *ddev,
                            hipStream_t s);
int launch_adjoint_outer(const Grid &g, int ncomp, const double *eps,
                         const double *mu, double *da, hipStream_t s);
int launch_local_sums(const Grid &g, int ncomp, const LocalDev &loc,
                      const double *x, double *partials, int *nblocks,
                      hipStream_t s);
int launch_stream_copy(const void *src, void *dst, int64_t n16,
                       hipStream_t s);
int launch_reduce_columns(const double *partials, int nblocks, int ncol,
                          double *out, hipStream_t s);

}  // namespace jx
