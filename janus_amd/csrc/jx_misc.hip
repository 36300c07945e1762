// Residual reduction, local-operator, adjoint-contraction and stream-copy
// kernels of janus_amd, with their launchers.

#include "jx_device.h"

namespace jx {

// ------------------------------------------------------------ residual
constexpr int kFinalizeThreads = 1024;

__global__ void __launch_bounds__(kFinalizeThreads)
    k_finalize(const double *partials, int nb, SolveState *st,
               double *res_hist, const Slot *tab) {
  if (tab) {
    // batched: one block per entry
    const Slot &sl = tab[blockIdx.x];
    partials = sl.partials;
    st = sl.st;
    res_hist = sl.res_hist;
  }
  partials = global_ptr(partials);
  st = global_ptr(st);
  res_hist = global_ptr(res_hist);
  if (solve_done(st)) return;
  __shared__ double red[kFinalizeThreads / 64];
  // eight independent accumulators keep eight loads in flight per thread;
  // the summation order is fixed, so the result is deterministic
  double acc[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  const int stride = kFinalizeThreads * 8;
  // whole strides without bounds checks (four per unrolled step: 32 loads
  // in flight per thread), then the checked tail; accumulator u sums the
  // same partials in the same order either way
  int i0 = 0;
#pragma unroll 4
  for (; i0 + stride <= nb; i0 += stride) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      x[u] = partials[i0 + u * kFinalizeThreads + threadIdx.x];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += x[u];
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    int i = i0 + u * kFinalizeThreads + threadIdx.x;
    if (i < nb) acc[u] += partials[i];
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

// Gradient with respect to a per-voxel matrix field: da[v][i][j] =
// -mu[v][i] eps[v][j] (row-major sym x sym blocks).
template <int C>
__global__ void __launch_bounds__(kThreads)
    k_adjoint_outer(const double *eps, const double *mu, int64_t N,
                    double *da) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double e[C], m[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      e[c] = eps[v * C + c];
      m[c] = mu[v * C + c];
    }
    double *dv = da + v * (C * C);
#pragma unroll
    for (int i = 0; i < C; ++i)
#pragma unroll
      for (int j = 0; j < C; ++j) dv[i * C + j] = -(m[i] * e[j]);
  }
}

// ------------------------------------------------------------ local apply
// TRANSPOSE applies the transposed local operator; the isotropic modes are
// symmetric, so only the matrix modes differ.
template <int C, bool TRANSPOSE>
__global__ void __launch_bounds__(kThreads)
    k_local_apply(const double *x, double *y, int64_t N, LocalDev loc) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double e[C], t[C];
#pragma unroll
    for (int c = 0; c < C; ++c) e[c] = x[v * C + c];
    if (TRANSPOSE && loc.mode >= LOCAL_PHASE_MATRIX)
      mat_voxel<C, true>(local_matrix<C>(loc, v), e, t);
    else
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

int launch_finalize(const double *partials, int nblocks, SolveState *st,
                    double *res_hist, hipStream_t s, const Batch *b) {
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)batch_count(b)),
                     dim3(kFinalizeThreads), 0, s, partials, nblocks, st,
                     res_hist, b ? b->tab : nullptr);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

int launch_local_apply(const Grid &g, int ncomp, const LocalDev &loc,
                       const double *x, double *y, hipStream_t s,
                       bool transpose) {
  int64_t nb = std::min<int64_t>((g.N + kThreads - 1) / kThreads, 65536);
  dim3 grid((unsigned)nb), block(kThreads);
  if (ncomp == 3 && !transpose)
    hipLaunchKernelGGL((k_local_apply<3, false>), grid, block, 0, s, x, y, g.N, loc);
  else if (ncomp == 3)
    hipLaunchKernelGGL((k_local_apply<3, true>), grid, block, 0, s, x, y, g.N, loc);
  else if (ncomp == 6 && !transpose)
    hipLaunchKernelGGL((k_local_apply<6, false>), grid, block, 0, s, x, y, g.N, loc);
  else if (ncomp == 6)
    hipLaunchKernelGGL((k_local_apply<6, true>), grid, block, 0, s, x, y, g.N, loc);
  else
    return fail(JX_EINVAL, "local_apply: ncomp must be 3 or 6");
  JX_HIP(hipGetLastError());
  return JX_OK;
}

// Sets *flag to 1 when the matrix of some voxel (or phase) is not
// symmetric; exact comparison.
template <int C>
__global__ void __launch_bounds__(kThreads)
    k_asymmetric(const double *__restrict__ a, int64_t count, int *flag) {
  int bad = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < count;
       v += (int64_t)gridDim.x * blockDim.x) {
    const double *av = a + v * (C * C);
#pragma unroll
    for (int i = 0; i < C; ++i)
#pragma unroll
      for (int j = i + 1; j < C; ++j) bad |= av[i * C + j] != av[j * C + i];
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

int matrices_symmetric(int ncomp, const double *a, int64_t count, int *flag_dev,
                       bool *out) {
  // the matrices may have been written on any stream of the caller (the
  // creation call takes none): wait for all of them before reading
  JX_HIP(hipDeviceSynchronize());
  JX_HIP(hipMemset(flag_dev, 0, sizeof(int)));
  int64_t nb = std::max<int64_t>(1, std::min<int64_t>((count + kThreads - 1) / kThreads, 4096));
  if (ncomp == 3)
    hipLaunchKernelGGL(k_asymmetric<3>, dim3((unsigned)nb), dim3(kThreads), 0, 0,
                       a, count, flag_dev);
  else if (ncomp == 6)
    hipLaunchKernelGGL(k_asymmetric<6>, dim3((unsigned)nb), dim3(kThreads), 0, 0,
                       a, count, flag_dev);
  else
    return fail(JX_EINVAL, "matrices_symmetric: ncomp must be 3 or 6");
  JX_HIP(hipGetLastError());
  int flag = 0;
  JX_HIP(hipMemcpy(&flag, flag_dev, sizeof(int), hipMemcpyDeviceToHost));
  *out = flag == 0;
  return JX_OK;
}

// Block-diagonal linear operator of any block size (operators.pyx:719-890):
// one thread per voxel, the voxel's (m, n) matrix read row by row.
__global__ void __launch_bounds__(kThreads)
    k_block_linear(const double *__restrict__ a, const double *__restrict__ x,
                   double *__restrict__ y, int64_t N, int m, int n,
                   int transpose) {
  int64_t mn = (int64_t)m * n;
  int xs = transpose ? m : n, ys = transpose ? n : m;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    const double *av = a + v * mn;
    const double *xv = x + v * xs;
    double *yv = y + v * ys;
    for (int i = 0; i < ys; ++i) {
      double acc = 0.0;
      if (transpose)
        for (int j = 0; j < xs; ++j) acc += av[(int64_t)j * n + i] * xv[j];
      else
        for (int j = 0; j < xs; ++j) acc += av[(int64_t)i * n + j] * xv[j];
      yv[i] = acc;
    }
  }
}

int launch_block_linear(int64_t N, int m, int n, const double *a,
                        const double *x, double *y, bool transpose,
                        hipStream_t s) {
  if (N <= 0) return JX_OK;
  int64_t nb = std::min<int64_t>((N + kThreads - 1) / kThreads, 65536);
  hipLaunchKernelGGL(k_block_linear, dim3((unsigned)nb), dim3(kThreads), 0, s,
                     a, x, y, N, m, n, (int)transpose);
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

int launch_adjoint_outer(const Grid &g, int ncomp, const double *eps,
                         const double *mu, double *da, hipStream_t s) {
  int64_t nb = std::min<int64_t>((g.N + kThreads - 1) / kThreads, 65536);
  if (ncomp == 3)
    hipLaunchKernelGGL(k_adjoint_outer<3>, dim3((unsigned)nb), dim3(kThreads),
                       0, s, eps, mu, g.N, da);
  else if (ncomp == 6)
    hipLaunchKernelGGL(k_adjoint_outer<6>, dim3((unsigned)nb), dim3(kThreads),
                       0, s, eps, mu, g.N, da);
  else
    return fail(JX_EINVAL, "adjoint_outer: ncomp must be 3 or 6");
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

bool batchable(const Grid &g, int ncomp) {
  if (ncomp != 3 && ncomp != 6) return false;
  for (int d = 0; d < g.dim; ++d)
    if (!reg_fft_length((int)g.n[d])) return false;
  return true;
}

}  // namespace jx
