// Fused pass E + pass A of the Lippmann-Schwinger iteration (register FFT
// lengths): the inverse last-axis transform and fixed-point update of
// iteration k followed by the local operator and forward last-axis
// transform of iteration k + 1.  Layouts as in pass A (jx_rows_fwd.hip).

#include "jx_device.h"

namespace jx {

// Fused pass E of iteration k and pass A of iteration k + 1: inverse
// last-axis transform, fixed-point update eps <- E (+ rhs) - y with the
// residual partial sums, then (C - C0) : eps and the forward last-axis
// transform of the result, written over the same spectrum rows.  The new
// iterate never travels back from memory.  MODE 1: E only; 2: E + rhs.
// Minimum waves per SIMD requested for the fused update pass: 3 (about 70
// registers spilled to scratch) measured faster than 2 at 256^3.
#ifndef JX_FUSED_PREFETCH
#define JX_FUSED_PREFETCH 1
#endif
// Rows in reverse order within each XCD's range (xcd_tile_rev): the fused
// pass then starts on the slabs the inverse axis-1 pass wrote last, which
// are still in the 256 MB Infinity Cache, and ends on the slabs the next
// forward axis-1 pass reads first (256^3: EA 0.651 -> 0.631 ms and B 0.312
// -> 0.297 ms, 595 -> 607 it/s; 512^3 unchanged, its spectrum being 25
// times the cache).
#ifndef JX_FUSED_REVERSE
#define JX_FUSED_REVERSE 1
#endif
#ifndef JX_FUSED_MIN_WAVES
#define JX_FUSED_MIN_WAVES 3
#endif
// Same for the 8-value plans (1: the compiler's choice), and for the
// 8-value 3D isotropic-family kernels without a field right-hand side,
// which fit 4 waves per SIMD with the prefetched iterate (5 without it, and
// the prefetch at 5 spills).
#ifndef JX_FUSED8_MIN_WAVES
#define JX_FUSED8_MIN_WAVES 1
#endif
#ifndef JX_FUSED8_ISO_MIN_WAVES
#define JX_FUSED8_ISO_MIN_WAVES (JX_FUSED_PREFETCH ? 4 : 5)
#endif
// The 10-value 160-point plan loads the previous iterate early too (below).
#ifndef JX_FUSED10_PREFETCH
#define JX_FUSED10_PREFETCH 1
#endif
// The other plans (6, 10, 12 and 20 values): the compiler's choice.
#ifndef JX_FUSEDX_MIN_WAVES
#define JX_FUSEDX_MIN_WAVES 1
#endif

template <class P, int C, int NP, int MODE, int LF>
__global__ void __launch_bounds__(NP * reg_tpp<P>(),
                                  reg_v<P>() == 16 ? JX_FUSED_MIN_WAVES
                                  : reg_v<P>() == 8
                                      ? (C == 6 && LF == LF_ISO && MODE == 1
                                             ? JX_FUSED8_ISO_MIN_WAVES
                                             : JX_FUSED8_MIN_WAVES)
                                      : JX_FUSEDX_MIN_WAVES)
    k_rows_update_forward_reg(double *spec, double *y, int64_t nrows,
                              SpecPad lay, double scale, const double2 *W, LsArgs ls,
                              LocalDev loc, const SolveState *st,
                              BatchRef br) {
  constexpr int L = P::L;
  const Slot *sl;
  // each XCD walks its rows from the end of its range, where the inverse
  // axis-1 pass (which walks the same slab range forwards) wrote last
  const int64_t blk = batch_block<JX_FUSED_REVERSE != 0>(br, sl);
  if (sl) {
    spec = sl->spec;
    y = sl->eps;
    ls.rhs = sl->rhs;
    ls.partials = sl->partials;
    loc = sl->loc;
    st = sl->st;
  }
  spec = global_ptr(spec);
  y = global_ptr(y);
  ls.rhs = global_ptr(ls.rhs);
  ls.partials = global_ptr(ls.partials);
  loc.phase = global_ptr(loc.phase);
  loc.table = global_ptr(loc.table);
  loc.sph = global_ptr(loc.sph);
  loc.dev = global_ptr(loc.dev);
  loc.mat = global_ptr(loc.mat);
  st = global_ptr(st);
  W = global_ptr(W);
  if (solve_done(st)) return;
  constexpr int TPP = reg_tpp<P>(), LS = reg_lds_stride(L), NH = L / 2 + 1;
  constexpr int V = reg_v<P>();
  constexpr int NT = NP * TPP;
  constexpr int RB = 2 * NP / C;
  extern __shared__ double2 smem[];
  // one partial sum per wave (384-thread blocks: 512- and 1024-point plans)
  __shared__ double red[(NT + 63) / 64];
  __shared__ double2 twl[JX_ROWS_TQ ? L / 4 : 1];
  const int tid = threadIdx.x;
  const int64_t row0 = blk * RB;
  const int rb = (int)min((int64_t)RB, nrows - row0);
  const int p = rows_tfast<P>() ? tid / TPP : tid % NP;
  const int t = rows_tfast<P>() ? tid % TPP : tid / NP;
  const double2 *Wt = stage_twiddles<L, NT, JX_ROWS_TQ>(W, twl);
  // 8-value plans and the 10-value 160-point plan, 6 components: the
  // previous iterate is loaded before the inverse transform, so that its
  // latency overlaps the spectrum loads and the transform (256^3: 0.654
  // against 0.670 ms, at 4 waves per SIMD instead of 5; 160^3: 0.198
  // against 0.232 ms, which makes the 10-value plan faster than the
  // 20-value one's 0.218; with it the 10-value 80-point plan measured
  // 0.0286 against 0.0264 ms and the 6-value 96-point one 0.0411 against
  // 0.0387 ms)
  constexpr bool PREFETCH =
      C == 6 && ((JX_FUSED_PREFETCH && V == 8) ||
                 (JX_FUSED10_PREFETCH && V == 10 && L == 160));
  double2 prev[PREFETCH ? V : 1];
  if constexpr (PREFETCH) {
    const int r = p / 3, c0 = 2 * (p - r * 3);
    if (r < rb) {
      const double2 *yr = (const double2 *)(y + (row0 + r) * L * 6 + c0);
#pragma unroll
      for (int i = 0; i < V; ++i) prev[i] = yr[(int64_t)reg_out_pos<P>(t, i) * 3];
    }
  }

  // ---- inverse: half spectra into the pair slots (see k_rows_inverse_reg)
  if constexpr (C == 6) {
    load_half_spectra6<L, RB, NT, LS>(spec, lay, row0, rb, tid, smem);
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

  // ---- update; v becomes the new iterate (part 0 in .x, part 1 in .y)
  double acc = 0.;
  if constexpr (C == 6) {
    const int r = p / 3, c0 = 2 * (p - r * 3);
    if (r < rb) {
      double2 *yr = (double2 *)(y + (row0 + r) * L * 6 + c0);
      const double2 *hr =
          MODE == 2 ? (const double2 *)(ls.rhs + (row0 + r) * L * 6 + c0)
                    : nullptr;
      const double e0 = st->E[c0], e1 = st->E[c0 + 1];
#pragma unroll
      for (int i = 0; i < V; ++i) {
        int64_t off = (int64_t)reg_out_pos<P>(t, i) * 3;
        double2 nv = make_double2(ls_new(e0, v[i].x, scale),
                                  ls_new(e1, v[i].y, scale));
        if constexpr (MODE == 2) {
          double2 h = hr[off];
          nv = make_double2(ls_new_rhs(e0, h.x, v[i].x, scale),
                            ls_new_rhs(e1, h.y, v[i].y, scale));
        }
        double2 o;
        if constexpr (PREFETCH) o = prev[i];
        else o = yr[off];
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
          int64_t off = (int64_t)reg_out_pos<P>(t, i) * C;
          double raw = part ? v[i].y : v[i].x;
          nv = ls_new(st->E[c], raw, scale);
          if constexpr (MODE == 2) nv = ls_new_rhs(st->E[c], hr[off], raw, scale);
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
  // (block_sum's barrier has retired the inverse transform's LDS reads)
#pragma unroll
  for (int i = 0; i < V; ++i) seq[lp(reg_out_pos<P>(t, i))] = v[i];
  __syncthreads();
  // a compile-time trip count lets the phase loads of all rounds be issued
  // together
  constexpr int KV = (RB * L + NT - 1) / NT;
#pragma unroll
  for (int kv = 0; kv < KV; ++kv) {
    const int w = tid + kv * NT;
    if (w >= rb * L) continue;
    int r = w / L, j = w - r * L;
    int64_t vox = (row0 + r) * L + j;
    double e[C], tt[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      int q = r * C + c;
      e[c] = ((const double *)&smem[(q >> 1) * LS + lp(j)])[q & 1];
    }
    local_voxel_f<C, LF>(loc, vox, e, tt);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      int q = r * C + c;
      ((double *)&smem[(q >> 1) * LS + lp(j)])[q & 1] = tt[c];
    }
  }
  __syncthreads();

  // ---- forward transform of the polarization rows (see k_rows_forward_reg)
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = seq[lp(reg_in_pos<P>(t, i))];
  __syncthreads();
  // same twiddles as the inverse transform: an opaque copy of the table
  // pointer keeps the compiler from holding them in registers in between
  const double2 *Wf = Wt;
  if constexpr (JX_ROWS_TQ) asm volatile("" : "+s"(Wf));
  else Wf = global_ptr(Wt);
  reg_fft<P, -1, JX_ROWS_TQ>(v, t, seq, Wf);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < V; ++i) seq[lp(reg_out_pos<P>(t, i))] = v[i];
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
      double *sp = spec + spec_row_off(row0 + r, NH * 12, lay) + k * 12 + 4 * cp;
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
      double *sr = spec + spec_row_off(row0 + r, NH * 2 * C, lay) + k * 2 * C;
      *(double2 *)(sr + 2 * c) = make_double2(re, im);
    }
  }
}
// ============================================================ launchers

template <class P, int C>
static int launch_rows_update_forward_reg_lc(const Grid &g, double *spec,
                                             double *y, const LsArgs &la,
                                             const LocalDev &loc,
                                             const SolveState *st,
                                             const double2 *W, hipStream_t s,
                                             const Batch *b) {
  constexpr int L = P::L;
  if constexpr (C == 3 || C == 6) {
    constexpr int NP = rows_np<P>();
    constexpr int RB = 2 * NP / C;
    size_t lds = (size_t)NP * reg_lds_stride(L) * sizeof(double2);
    int rc = check_lds(lds, "rows_update_forward");
    if (rc) return rc;
    const int64_t nbe = (g.rows + RB - 1) / RB;
    const BatchRef br = batch_ref(b, nbe);
    dim3 grid((unsigned)(nbe * batch_count(b))), block(NP * reg_tpp<P>());
    double scale = 1.0 / (double)g.N;
    auto go = [&](auto kernel) {
      set_lds_attr(kernel, lds);
      hipLaunchKernelGGL(kernel, grid, block, lds, s, spec, y, g.rows,
                         spec_layout(g, C), scale, W, la, loc, st, br);
    };
    const bool mat = local_family(loc) == LF_MATRIX;
    if (la.rhs && mat)
      go(k_rows_update_forward_reg<P, C, NP, 2, LF_MATRIX>);
    else if (la.rhs)
      go(k_rows_update_forward_reg<P, C, NP, 2, LF_ISO>);
    else if (mat)
      go(k_rows_update_forward_reg<P, C, NP, 1, LF_MATRIX>);
    else
      go(k_rows_update_forward_reg<P, C, NP, 1, LF_ISO>);
    JX_HIP(hipGetLastError());
    return JX_OK;
  } else {
    return fail(JX_EINVAL, "fixed point needs sym components");
  }
}

static int rows_update_forward_reg(const Grid &g, int C, double *spec,
                                   double *y, const LsArgs &la,
                                   const LocalDev &loc, const SolveState *st,
                                   const double2 *W, hipStream_t s,
                                   const Batch *b) {
  JX_WITH_PLAN_C((int)g.L, g.plan[PASS_FUSED],
                 launch_rows_update_forward_reg_lc, g, spec, y, la, loc, st, W,
                 s, b);
}

bool rows_fusable(const Grid &g, int ncomp) {
  return (ncomp == 3 || ncomp == 6) && reg_fft_length((int)g.L);
}

int launch_rows_update_forward(const Grid &g, int ncomp, double *spec,
                               double *y, const RowsInverseLs &ls,
                               const LocalDev &loc, const SolveState *st,
                               int *nblocks_out, hipStream_t s,
                               const Batch *b) {
  if (!rows_fusable(g, ncomp))
    return fail(JX_EUNSUPPORTED, "fused update pass needs a register FFT length");
  FftLen P;
  int rc = get_fft_len((int)g.L, &P);
  if (rc) return rc;
  LsArgs la;
  la.rhs = ls.rhs;
  la.partials = ls.partials;
  int RB = reg_rows_per_block((int)g.L, ncomp, g.plan[PASS_FUSED]);
  if (nblocks_out) *nblocks_out = (int)((g.rows + RB - 1) / RB);
  return rows_update_forward_reg(g, ncomp, spec, y, la, loc, st, P.W, s, b);
}

}  // namespace jx
