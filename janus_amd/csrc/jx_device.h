// Device helpers and launch plumbing shared by the kernel sources.
#pragma once

#include <algorithm>
#include <cmath>

#include "jx_fft_lds.h"
#include "jx_fft_reg.h"
#include "jx_gamma.h"
#include "jx_internal.h"

namespace jx {

// LDS row stride of a sequence: one complex pad for even lengths breaks the
// power-of-two strides between sequences.
static inline int lds_stride(int L) { return (L % 2 == 0) ? L + 1 : L; }

// ------------------------------------------------------------ pointers
// A pointer into device memory as the compiler should see it: global.  The
// kernels that serve batches take their pointers from the slot table or the
// kernel arguments, which leaves them generic; every access then becomes a
// flat instruction, which needs 64-bit VGPR addresses and also counts
// against the LDS wait counter (LDS steps then wait for outstanding HBM
// loads and stores).  A plain pair of address-space casts is folded away,
// so the global pointer passes through an empty asm; the result is also
// opaque to common-subexpression elimination.  p must be uniform across the
// block.
template <class T>
__device__ __forceinline__ T *global_ptr(T *p) {
  typedef __attribute__((address_space(1))) T GT;
  GT *g = (GT *)p;
  asm volatile("" : "+s"(g));
  return (T *)g;
}

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

// t = a e (TRANSPOSE: t = a^T e) for one C x C row-major matrix; each
// output is accumulated over j = 0, 1, ... in the order of janus's c_apply
// and c_apply_transpose loops (operators.pyx:739-756, 758-775).
template <int C, bool TRANSPOSE>
__device__ __forceinline__ void mat_voxel(const double *__restrict__ a,
                                          const double *e, double *t) {
#pragma unroll
  for (int i = 0; i < C; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < C; ++j)
      acc += (TRANSPOSE ? a[j * C + i] : a[i * C + j]) * e[j];
    t[i] = acc;
  }
}

// The C x C matrix of voxel vox in the matrix modes.
template <int C>
__device__ __forceinline__ const double *local_matrix(const LocalDev &loc,
                                                      int64_t vox) {
  int64_t k = loc.mode == LOCAL_PHASE_MATRIX ? (int64_t)loc.phase[vox] : vox;
  return loc.mat + k * (C * C);
}

// Isotropic modes: phase tables hold (tr, dev) per phase; field mode
// computes tr per voxel.
template <int C>
__device__ __forceinline__ void local_voxel_iso(const LocalDev &loc,
                                                int64_t vox, const double *e,
                                                double *t) {
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

// t = local : e at voxel vox, for a local operator of family LF.
template <int C, int LF>
__device__ __forceinline__ void local_voxel_f(const LocalDev &loc, int64_t vox,
                                              const double *e, double *t) {
  if constexpr (LF == LF_MATRIX)
    mat_voxel<C, false>(local_matrix<C>(loc, vox), e, t);
  else
    local_voxel_iso<C>(loc, vox, e, t);
}

// Any mode, chosen at run time (kernels that are not register bound).
template <int C>
__device__ __forceinline__ void local_voxel(const LocalDev &loc, int64_t vox,
                                            const double *e, double *t) {
  if (loc.mode >= LOCAL_PHASE_MATRIX)
    local_voxel_f<C, LF_MATRIX>(loc, vox, e, t);
  else
    local_voxel_f<C, LF_ISO>(loc, vox, e, t);
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
static __device__ double block_sum(double v, double *red) {
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
  SpecPad sp;  // spectrum layout (padded 3D slabs)
};

// Offset in doubles of spectrum row `row` (rowlen doubles per row) in the
// padded layout of spec_layout().
__device__ __forceinline__ int64_t spec_row_off(int64_t row, int64_t rowlen,
                                                const SpecPad &sp) {
  return row * rowlen + (sp.pad ? (row / sp.slab_rows) * sp.pad : 0);
}

__device__ __forceinline__ bool solve_done(const SolveState *st) {
  return st != nullptr && *((volatile const int *)&st->done) != 0;
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

// xcd_tile with each XCD walking its range from the end: a pass that reads
// what the previous pass wrote last (still in the Infinity Cache) first.
__device__ __forceinline__ int64_t xcd_tile_rev(int64_t b, int64_t nb) {
  const int64_t nx = 8;
  int64_t q = nb / nx, r = nb - q * nx;
  int64_t x = b - (b / nx) * nx, slot = b / nx;
  int64_t count = q + (x < r ? 1 : 0);
  return x * q + (x < r ? x : r) + (count - 1 - slot);
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

// ------------------------------------------------------------ batches
// Kernel-side view of a batch: tab == nullptr for a single launch (pointer
// arguments by value), else nbe blocks per entry of the slot table.
struct BatchRef {
  const Slot *tab;
  int64_t nbe;
};

// Index of the block within its entry (XCD-aware order over the whole
// launch); *sl is the entry's slot in a batched launch, else nullptr.
template <bool REV = false>
__device__ __forceinline__ int64_t batch_block(const BatchRef &br,
                                               const Slot *&sl) {
  int64_t blk = REV ? xcd_tile_rev(blockIdx.x, gridDim.x)
                    : xcd_tile(blockIdx.x, gridDim.x);
  sl = nullptr;
  if (br.tab) {
    const int64_t e = blk / br.nbe;
    blk -= e * br.nbe;
    sl = br.tab + e;
  }
  return blk;
}

inline BatchRef batch_ref(const Batch *b, int64_t nbe) {
  BatchRef r;
  r.tab = b ? b->tab : nullptr;
  r.nbe = nbe;
  return r;
}

inline int64_t batch_count(const Batch *b) { return b ? b->n : 1; }

// Fixed-point update arguments (the macroscopic strain is read from the
// solve state).
struct LsArgs {
  const double *rhs = nullptr;
  double *partials = nullptr;
  // conjugate gradients (pass E): the direction p and the C0 weights
  const double *cgp = nullptr;
  double wd = 0., wt = 0.;
};

// The C components of voxel v: 16-byte accesses for C = 6 (48-byte voxels).
template <int C>
__device__ __forceinline__ void load_voxel(const double *a, int64_t v,
                                           double *x) {
  if constexpr (C == 6) {
    const double2 *a2 = (const double2 *)(a + v * 6);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double2 t = a2[i];
      x[2 * i] = t.x;
      x[2 * i + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) x[c] = a[v * C + c];
  }
}

template <int C>
__device__ __forceinline__ void store_voxel(double *a, int64_t v,
                                            const double *x) {
  if constexpr (C == 6) {
    double2 *a2 = (double2 *)(a + v * 6);
#pragma unroll
    for (int i = 0; i < 3; ++i) a2[i] = make_double2(x[2 * i], x[2 * i + 1]);
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) a[v * C + c] = x[c];
  }
}

// a : C0 : b of one voxel, C0 = dim k0 J + 2 g0 K in Mandel notation:
// wd (a . b) + wt tr(a) tr(b) with wd = 2 g0, wt = k0 - 2 g0 / dim.
template <int C>
__device__ __forceinline__ double c0_dot(const double *a, const double *b,
                                         double wd, double wt) {
  constexpr int D = C == 3 ? 2 : 3;
  double dot = 0., ta = 0., tb = 0.;
#pragma unroll
  for (int c = 0; c < C; ++c) dot += a[c] * b[c];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    ta += a[c];
    tb += b[c];
  }
  return wd * dot + wt * ta * tb;
}

// Last-axis c2r passes, 6 components: the half spectra of RB rows into the
// NP = 3 RB pair sequences of LDS (stride LS), pair (A, B) as Z = A + i B:
// Z[k] = A[k] + i B[k] and Z[L - k] = conj(A[k]) + i conj(B[k]); the DC and
// Nyquist slots hold (Re A, Re B), their imaginary parts dropped as FFTW's
// and pocketfft's c2r do.  Rows past rb are zero.  Every global load of the
// block is issued before the first LDS store (a loop with a run-time trip
// count would wait for each round's loads in turn).
template <int L, int RB, int NT, int LS>
__device__ __forceinline__ void load_half_spectra6(const double *spec,
                                                   const SpecPad &lay,
                                                   int64_t row0, int rb,
                                                   int tid, double2 *smem) {
  constexpr int NH = L / 2 + 1;
  constexpr int ITEMS = RB * NH * 3;
  constexpr int K = (ITEMS + NT - 1) / NT;
  double2 xa[K], xb[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int w = tid + k * NT;
    xa[k] = make_double2(0., 0.);
    xb[k] = make_double2(0., 0.);
    const int rest = w / 3, r = rest / NH;
    if (w < ITEMS && r < rb) {
      const int cp = w - rest * 3, kk = rest - r * NH;
      const double *sp =
          spec + spec_row_off(row0 + r, NH * 12, lay) + kk * 12 + 4 * cp;
      xa[k] = *(const double2 *)sp;
      xb[k] = *(const double2 *)(sp + 2);
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int w = tid + k * NT;
    if (w < ITEMS) {
      const int rest = w / 3, r = rest / NH;
      const int cp = w - rest * 3, kk = rest - r * NH;
      double2 *Z = smem + (r * 3 + cp) * LS;
      const double2 a = xa[k], b = xb[k];
      if (kk == 0 || 2 * kk == L) {
        Z[lp(kk)] = make_double2(a.x, b.x);
      } else {
        Z[lp(kk)] = make_double2(a.x - b.y, a.y + b.x);
        Z[lp(L - kk)] = make_double2(a.x + b.y, b.x - a.y);
      }
    }
  }
}

// Lane order of the register FFT kernels of the last axis: sequence fastest
// (lane = t * sequences + p) or thread-in-sequence fastest (lane = p * TPP +
// t: the stage exchanges through LDS touch consecutive 16-byte slots, free
// of bank conflicts).  JX_ROWS_TFAST 0: always the first, 1: always the
// second, 2 (default) -- thread-fast for the mixed-radix plans of 80 and
// 160 points and the 12-value one of 96 (where it measured faster: the
// 160-point fused pass 0.224 against 0.284 ms at 160^3 with 20 values,
// 0.199 against 0.242 ms with 10; the 10-value 80-point one 0.027 against
// 0.051 ms at 80^3), sequence-fast for the power-of-two plans (the 8-value
// 256- and 512-point passes measured slower thread-fast) and the 6-value
// 96-point plan (96^3 fused pass 0.0380 against 0.0387 ms, c2r 0.0211
// against 0.0226 ms).
#ifndef JX_ROWS_TFAST
#define JX_ROWS_TFAST 2
#endif
template <class P>
constexpr bool rows_tfast() {
  return JX_ROWS_TFAST == 1 ||
         (JX_ROWS_TFAST == 2 && (P::L & (P::L - 1)) != 0 &&
          !(P::L == 96 && P::V == 6));
}

// Twiddle source of the last-axis kernels: quarter table in LDS (1) or the
// global table (0, fewer registers; measured as fast or faster).
#ifndef JX_ROWS_TQ
#define JX_ROWS_TQ 0
#endif

// ------------------------------------------------------------ launch helpers
static inline int check_lds(size_t bytes, const char *what) {
  if ((long long)bytes > lds_limit_bytes())
    return fail(JX_EUNSUPPORTED,
                std::string(what) + ": transform length too large for LDS (" +
                    std::to_string(bytes) + " bytes needed)");
  return JX_OK;
}

template <typename K>
static inline void set_lds_attr(K kernel, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute((const void *)kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

// Threads per block of the register kernels: 192, or six transforms when
// those need more (a block of the axis kernels holds whole points of up to
// six components).  Pairs per block of the last-axis kernels and sequences
// per block of the axis kernels follow.
template <class P>
constexpr int reg_block_threads() {
  return 6 * reg_tpp<P>() > 192 ? 6 * reg_tpp<P>() : 192;
}
template <class P>
constexpr int reg_np() {
  return reg_block_threads<P>() / reg_tpp<P>();
}
template <class P>
constexpr int reg_nseq() {
  return reg_block_threads<P>() / reg_tpp<P>();
}

// Pair sequences per block of the last-axis (rows) kernels.  Plans with 64
// threads per transform (512 points with 8 values, 1024 points) take three
// pairs, 192 threads and one 6-component row per block: at 512^3 the fused
// update pass went from 6.06 to 5.39 ms against six pairs (384 threads, 52
// KB of LDS, three blocks per CU).  The others take reg_np.
#ifndef JX_ROWS_NP64
#define JX_ROWS_NP64 3
#endif
template <class P>
constexpr int rows_np() {
  return (reg_tpp<P>() >= 64 && JX_ROWS_NP64 > 0) ? JX_ROWS_NP64
                                                  : reg_np<P>();
}

// Rows per block of the register last-axis kernels with plan variant k
// (0 = no register FFT for L).
static inline int reg_rows_per_block(int L, int C, int k) {
  if (!reg_fft_length(L) || (C != 1 && C != 3 && C != 6)) return 0;
  int np = 0;
  with_plan(L, k, [&](auto plan) -> int {
    np = rows_np<decltype(plan)>();
    return 0;
  });
  return 2 * np / C;
}

// Most rows-kernel blocks over the plan variants of the last axis.
static inline int64_t reg_rows_max_blocks(const Grid &g, int C) {
  int64_t nb = 0;
  for (int k = 0; k < reg_variants((int)g.L); ++k) {
    int rb = reg_rows_per_block((int)g.L, C, k);
    if (rb) nb = std::max<int64_t>(nb, (g.rows + rb - 1) / rb);
  }
  return nb;
}

// Rows per block of the generic last-axis kernels and their LDS bytes.
static inline int rows_per_block(int L, int C) {
  int Ls = lds_stride(L);
  // aim at about 1536 complex entries per ping-pong buffer
  int nseq = std::max(1, 1536 / Ls);
  int rb = std::max(1, (2 * nseq) / C);
  return rb;
}

static inline size_t rows_lds_bytes(int L, int C, int RB) {
  int nseq_max = (RB * C + 1) / 2;
  return 2 * (size_t)nseq_max * lds_stride(L) * sizeof(double2);
}

// Runs FN<P, C>(args) for the plan P of variant K of length LEN and the
// component count C in scope.
#define JX_WITH_PLAN_C(LEN, K, FN, ...)                                    \
  return with_plan((LEN), (K), [&](auto plan_) -> int {                    \
    using P_ = decltype(plan_);                                            \
    switch (C) {                                                           \
      case 1: return FN<P_, 1>(__VA_ARGS__);                               \
      case 3: return FN<P_, 3>(__VA_ARGS__);                               \
      case 6: return FN<P_, 6>(__VA_ARGS__);                               \
      default: return fail(JX_EINVAL, "ncomp must be 1, 3 or 6");          \
    }                                                                      \
  })

}  // namespace jx
