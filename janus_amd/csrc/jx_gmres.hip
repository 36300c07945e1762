// Restarted GMRES on the Lippmann-Schwinger operator (no janus counterpart
// in the library; janus's examples/square_inclusion_scipy.py:52-59, 88 hands
// the same strain form to scipy's GMRES).
//
// The fluctuation x = eps - E solves (I + G dC) x = b with G = Gamma0 and
// b = rhs - G[dC : E], in the Euclidean inner product; dC need not be
// symmetric.  Arnoldi step j: w = v_j + G[dC : v_j] (the Green operator's
// five passes), one classical Gram-Schmidt projection h_i = <w, v_i>,
// w -= sum_i h_i v_i, h_{j+1} = ||w||, v_{j+1} = w / h_{j+1}.  The field
// work runs in tiles of eight basis vectors per pass with deterministic
// per-block partial sums; the (restart + 1) x restart Hessenberg matrix,
// its Givens rotations and the triangular solve stay on the host, which
// reads the new column once per step.  The residual after iteration k is
// the Arnoldi estimate |g_{k+1}| = ||b - A x_k|| times 1 / (sqrt(N) ||E||).

#include <algorithm>
#include <cmath>
#include <vector>

#include "jx_device.h"
#include "jx_handles.h"

namespace jx {

namespace {

constexpr int kGmBlocks = 2048;  // blocks of the field kernels (partials)
constexpr int kGmTile = 8;       // basis vectors per field pass

// Block sums of <w, V_{i0 + k}> for k < cnt into partials[blk * kGmTile + k].
// form: w = vj + y is formed and stored first.
template <int C>
__global__ void __launch_bounds__(kThreads)
    k_gm_dots(double *w, const double *y, const double *vj, const double *V,
              int64_t vs, int64_t N, int i0, int cnt, int form,
              double *partials) {
  __shared__ double red[kThreads / 64];
  double acc[kGmTile];
#pragma unroll
  for (int k = 0; k < kGmTile; ++k) acc[k] = 0.;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double wv[C];
    if (form) {
      double a[C], b[C];
      load_voxel<C>(vj, v, a);
      load_voxel<C>(y, v, b);
#pragma unroll
      for (int c = 0; c < C; ++c) wv[c] = a[c] + b[c];
      store_voxel<C>(w, v, wv);
    } else {
      load_voxel<C>(w, v, wv);
    }
#pragma unroll
    for (int k = 0; k < kGmTile; ++k) {
      if (k < cnt) {
        double a[C];
        load_voxel<C>(V + (int64_t)(i0 + k) * vs, v, a);
        double d = 0.;
#pragma unroll
        for (int c = 0; c < C; ++c) d += wv[c] * a[c];
        acc[k] += d;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kGmTile; ++k) {
    double s = block_sum(acc[k], red);
    if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * kGmTile + k] = s;
    __syncthreads();
  }
}

// w += sign sum_{k < cnt} coef[i0 + k] V_{i0 + k}; norm: block sums of
// w . w after the update into partials[blk].
template <int C>
__global__ void __launch_bounds__(kThreads)
    k_gm_axpy(double *w, const double *V, int64_t vs, int64_t N, int i0,
              int cnt, const double *coef, double sign, int norm,
              double *partials) {
  __shared__ double red[kThreads / 64];
  double cf[kGmTile];
#pragma unroll
  for (int k = 0; k < kGmTile; ++k) cf[k] = k < cnt ? sign * coef[i0 + k] : 0.;
  double acc = 0.;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double wv[C];
    load_voxel<C>(w, v, wv);
#pragma unroll
    for (int k = 0; k < kGmTile; ++k) {
      if (k < cnt) {
        double a[C];
        load_voxel<C>(V + (int64_t)(i0 + k) * vs, v, a);
#pragma unroll
        for (int c = 0; c < C; ++c) wv[c] += cf[k] * a[c];
      }
    }
    store_voxel<C>(w, v, wv);
    if (norm) {
#pragma unroll
      for (int c = 0; c < C; ++c) acc += wv[c] * wv[c];
    }
  }
  if (norm) {
    double s = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
}

struct Evec {
  double e[6];
};

enum GmMode {
  GM_FILL,   // x = E
  GM_INIT,   // b = rhs - y, r = b, x = 0; norm of r
  GM_RESID,  // r = b - x - y; norm of r
  GM_SCALE,  // r *= f
  GM_FINAL   // x = E + x
};

template <int C, int MODE>
__global__ void __launch_bounds__(kThreads)
    k_gm_vec(double *x, double *r, double *b, const double *y,
             const double *rhs, int64_t N, Evec E, double f,
             double *partials) {
  __shared__ double red[kThreads / 64];
  double acc = 0.;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double a[C];
    if constexpr (MODE == GM_FILL) {
#pragma unroll
      for (int c = 0; c < C; ++c) a[c] = E.e[c];
      store_voxel<C>(x, v, a);
    } else if constexpr (MODE == GM_INIT) {
      double yv[C], hv[C], z[C];
      load_voxel<C>(y, v, yv);
      if (rhs) load_voxel<C>(rhs, v, hv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        a[c] = rhs ? hv[c] - yv[c] : -yv[c];
        z[c] = 0.;
        acc += a[c] * a[c];
      }
      store_voxel<C>(b, v, a);
      store_voxel<C>(r, v, a);
      store_voxel<C>(x, v, z);
    } else if constexpr (MODE == GM_RESID) {
      double bv[C], xv[C], yv[C];
      load_voxel<C>(b, v, bv);
      load_voxel<C>(x, v, xv);
      load_voxel<C>(y, v, yv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        a[c] = bv[c] - (xv[c] + yv[c]);
        acc += a[c] * a[c];
      }
      store_voxel<C>(r, v, a);
    } else if constexpr (MODE == GM_SCALE) {
      load_voxel<C>(r, v, a);
#pragma unroll
      for (int c = 0; c < C; ++c) a[c] = f * a[c];
      store_voxel<C>(r, v, a);
    } else {
      load_voxel<C>(x, v, a);
#pragma unroll
      for (int c = 0; c < C; ++c) a[c] = E.e[c] + a[c];
      store_voxel<C>(x, v, a);
    }
  }
  if constexpr (MODE == GM_INIT || MODE == GM_RESID) {
    double s = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
}

template <int C, int MODE>
int gm_vec(double *x, double *r, double *b, const double *y,
           const double *rhs, int64_t N, const Evec &E, double f,
           double *partials, hipStream_t s) {
  hipLaunchKernelGGL((k_gm_vec<C, MODE>), dim3(kGmBlocks), dim3(kThreads), 0,
                     s, x, r, b, y, rhs, N, E, f, partials);
  JX_HIP(hipGetLastError());
  return JX_OK;
}

// Device buffers of a GMRES solve and the host copies it reads back.
struct GmWork {
  double *V;         // restart + 1 basis fields
  double *y, *b;     // operator output, right-hand side
  double *partials;  // kGmBlocks x kGmTile
  double *small;     // device: h (restart + 1 + kGmTile), norm, coefficients
  double *host;      // pinned: small_len() doubles
  int m;
  // h takes restart + 1 entries plus a tile of padding (each reduction
  // writes a whole tile of columns); the norm and restart coefficients follow
  int hlen() const { return m + 1 + kGmTile; }
  int small_len() const { return hlen() + 1 + m; }
};

// Sum of the partials of `ncol` columns into out[0 .. ncol) (device), in a
// fixed order.
int reduce(const GmWork &W, int ncol, double *out, hipStream_t s) {
  return launch_reduce_columns(W.partials, kGmBlocks, ncol, out, s);
}

template <int C>
int gmres_solve(jx_solver *S, double *x, const double *rhs, const Evec &E,
                int max_iter, double tol, double res_scale, const GmWork &W,
                std::vector<double> &hist, int *status, hipStream_t s) {
  const Grid &g = S->green->g;
  const int64_t N = g.N, vs = (int64_t)C * N;
  const int m = W.m;
  const int HL = W.hlen();
  double *h_dev = W.small, *nrm_dev = W.small + HL,
         *coef_dev = W.small + HL + 1;
  int rc;
  // y = G[dC : E]; b = rhs - y; r = V_0 = b; x = 0
  if ((rc = gm_vec<C, GM_FILL>(x, nullptr, nullptr, nullptr, nullptr, N, E,
                               0., nullptr, s)))
    return rc;
  if ((rc = ls_matvec(S, x, W.y, S->st, s))) return rc;
  if ((rc = gm_vec<C, GM_INIT>(x, W.V, W.b, W.y, rhs, N, E, 0., W.partials,
                               s)))
    return rc;
  if ((rc = reduce(W, 1, nrm_dev, s))) return rc;
  JX_HIP(hipMemcpyAsync(W.host, nrm_dev, sizeof(double),
                        hipMemcpyDeviceToHost, s));
  JX_HIP(hipStreamSynchronize(s));
  double beta = std::sqrt(W.host[0]);
  *status = JX_SOLVE_MAXITER;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1),
      yv(m);
  auto Hij = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
  bool done = false;
  if (beta == 0.) {
    *status = JX_SOLVE_CONVERGED;
    done = true;
  } else if (!std::isfinite(beta)) {
    *status = JX_SOLVE_NONFINITE;
    done = true;
  }
  while (!done) {
    if ((rc = gm_vec<C, GM_SCALE>(nullptr, W.V, nullptr, nullptr, nullptr, N,
                                  E, 1.0 / beta, nullptr, s)))
      return rc;
    std::fill(H.begin(), H.end(), 0.);
    std::fill(gv.begin(), gv.end(), 0.);
    gv[0] = beta;
    int k = 0;
    bool stop = false;
    for (int j = 0; j < m; ++j) {
      double *vj = W.V + (int64_t)j * vs, *w = W.V + (int64_t)(j + 1) * vs;
      if ((rc = ls_matvec(S, vj, W.y, S->st, s))) return rc;
      // h_i = <v_j + y, v_i>, i <= j
      for (int i0 = 0; i0 <= j; i0 += kGmTile) {
        const int cnt = std::min(kGmTile, j + 1 - i0);
        hipLaunchKernelGGL((k_gm_dots<C>), dim3(kGmBlocks), dim3(kThreads), 0,
                           s, w, W.y, vj, W.V, vs, N, i0, cnt, i0 == 0 ? 1 : 0,
                           W.partials);
        JX_HIP(hipGetLastError());
        if ((rc = reduce(W, kGmTile, h_dev + i0, s))) return rc;
      }
      // w -= sum_i h_i v_i, then ||w||^2
      for (int i0 = 0; i0 <= j; i0 += kGmTile) {
        const int cnt = std::min(kGmTile, j + 1 - i0);
        const int last = i0 + kGmTile > j ? 1 : 0;
        hipLaunchKernelGGL((k_gm_axpy<C>), dim3(kGmBlocks), dim3(kThreads), 0,
                           s, w, W.V, vs, N, i0, cnt, h_dev, -1.0, last,
                           W.partials);
        JX_HIP(hipGetLastError());
      }
      if ((rc = reduce(W, 1, nrm_dev, s))) return rc;
      JX_HIP(hipMemcpyAsync(W.host, h_dev, sizeof(double) * (HL + 1),
                            hipMemcpyDeviceToHost, s));
      JX_HIP(hipStreamSynchronize(s));
      const double hn = std::sqrt(W.host[HL]);
      for (int i = 0; i <= j; ++i) Hij(i, j) = W.host[i];
      Hij(j + 1, j) = hn;
      for (int i = 0; i < j; ++i) {
        const double t = cs[i] * Hij(i, j) + sn[i] * Hij(i + 1, j);
        Hij(i + 1, j) = -sn[i] * Hij(i, j) + cs[i] * Hij(i + 1, j);
        Hij(i, j) = t;
      }
      const double d = std::hypot(Hij(j, j), Hij(j + 1, j));
      cs[j] = Hij(j, j) / d;
      sn[j] = Hij(j + 1, j) / d;
      Hij(j, j) = d;
      Hij(j + 1, j) = 0.;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      k = j + 1;
      const double res = std::fabs(gv[j + 1]) * res_scale;
      hist.push_back(res);
      if (!std::isfinite(res)) {
        *status = JX_SOLVE_NONFINITE;
        stop = true;
      } else if (res < tol) {
        *status = JX_SOLVE_CONVERGED;
        stop = true;
      } else if ((int)hist.size() >= max_iter) {
        stop = true;
      }
      if (stop || hn == 0.) break;
      if ((rc = gm_vec<C, GM_SCALE>(nullptr, w, nullptr, nullptr, nullptr, N,
                                    E, 1.0 / hn, nullptr, s)))
        return rc;
    }
    if (*status == JX_SOLVE_NONFINITE) break;
    // x += sum_i y_i v_i with H y = g (upper triangular, k x k)
    for (int i = k - 1; i >= 0; --i) {
      double t = gv[i];
      for (int l = i + 1; l < k; ++l) t -= Hij(i, l) * yv[l];
      yv[i] = t / Hij(i, i);
    }
    for (int i = 0; i < k; ++i) W.host[i] = yv[i];
    JX_HIP(hipMemcpyAsync(coef_dev, W.host, sizeof(double) * k,
                          hipMemcpyHostToDevice, s));
    for (int i0 = 0; i0 < k; i0 += kGmTile) {
      const int cnt = std::min(kGmTile, k - i0);
      hipLaunchKernelGGL((k_gm_axpy<C>), dim3(kGmBlocks), dim3(kThreads), 0, s,
                         x, W.V, vs, N, i0, cnt, coef_dev, 1.0, 0, W.partials);
      JX_HIP(hipGetLastError());
    }
    if (stop) break;
    // restart: r = V_0 = b - (x + G[dC : x])
    if ((rc = ls_matvec(S, x, W.y, S->st, s))) return rc;
    if ((rc = gm_vec<C, GM_RESID>(x, W.V, W.b, W.y, nullptr, N, E, 0.,
                                  W.partials, s)))
      return rc;
    if ((rc = reduce(W, 1, nrm_dev, s))) return rc;
    JX_HIP(hipMemcpyAsync(W.host, nrm_dev, sizeof(double),
                          hipMemcpyDeviceToHost, s));
    JX_HIP(hipStreamSynchronize(s));
    beta = std::sqrt(W.host[0]);
    if (beta == 0.) {
      *status = JX_SOLVE_CONVERGED;
      break;
    }
  }
  // eps = E + x
  return gm_vec<C, GM_FINAL>(x, nullptr, nullptr, nullptr, nullptr, N, E, 0.,
                             nullptr, s);
}

}  // namespace

}  // namespace jx

using namespace jx;

extern "C" {

int jx_solver_solve_gmres(jx_solver *S, double *eps, const double *E,
                          const double *rhs, double tol, int max_iter,
                          int restart, double res_scale, double *res_hist,
                          int *iters, int *status, void *stream) {
  if (!S || !eps || !E)
    return fail(JX_EINVAL, "jx_solver_solve_gmres: NULL argument");
  JX_ALIGNED(eps, "jx_solver_solve_gmres");
  JX_ALIGNED(rhs, "jx_solver_solve_gmres");
  if (max_iter < 1) return fail(JX_EINVAL, "max_iter must be >= 1");
  if (restart < 1 || restart > 256)
    return fail(JX_EINVAL, "restart must be in [1, 256]");
  hipStream_t s = (hipStream_t)stream;
  jx_green *G = S->green;
  const Grid &g = G->g;
  const int C = G->sym;
  if (res_scale <= 0.) {
    double nE = 0.;
    for (int c = 0; c < C; ++c) nE += E[c] * E[c];
    nE = std::sqrt(nE);
    if (nE == 0.) return fail(JX_EINVAL, "res_scale must be given when E = 0");
    res_scale = 1.0 / std::sqrt((double)g.N) / nE;  // square_basic.py:86
  }
  // the Green passes return at once when the solve state says done: keep
  // it running (GMRES stops on the host)
  SolveState init = {};
  init.max_iter = max_iter;
  for (int c = 0; c < 6; ++c) init.E[c] = c < C ? E[c] : 0.;
  *S->st_host = init;
  JX_HIP(hipMemcpyAsync(S->st, S->st_host, sizeof(SolveState),
                        hipMemcpyHostToDevice, s));
  const int64_t field = (int64_t)C * g.N;
  GmWork W = {};
  W.m = restart;
  int rc = JX_OK;
  hipError_t e = hipMalloc((void **)&W.V, sizeof(double) * field * (restart + 3));
  if (e != hipSuccess) return hip_fail(e, "jx_solver_solve_gmres: basis");
  W.y = W.V + field * (restart + 1);
  W.b = W.y + field;
  e = hipMalloc((void **)&W.partials, sizeof(double) * kGmBlocks * kGmTile);
  if (e == hipSuccess)
    e = hipMalloc((void **)&W.small, sizeof(double) * W.small_len());
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&W.host, sizeof(double) * W.small_len());
  std::vector<double> hist;
  int st = JX_SOLVE_MAXITER;
  if (e != hipSuccess) {
    rc = hip_fail(e, "jx_solver_solve_gmres: buffers");
  } else {
    Evec Ev = {};
    for (int c = 0; c < C; ++c) Ev.e[c] = E[c];
    rc = C == 3 ? gmres_solve<3>(S, eps, rhs, Ev, max_iter, tol, res_scale, W,
                                 hist, &st, s)
                : gmres_solve<6>(S, eps, rhs, Ev, max_iter, tol, res_scale, W,
                                 hist, &st, s);
    if (!rc) {
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) rc = hip_fail(e, "jx_solver_solve_gmres");
    }
  }
  if (W.host) (void)hipHostFree(W.host);
  if (W.small) (void)hipFree(W.small);
  if (W.partials) (void)hipFree(W.partials);
  (void)hipFree(W.V);
  if (rc) return rc;
  const int it = (int)hist.size();
  if (iters) *iters = it;
  if (status) *status = st;
  if (res_hist)
    for (int i = 0; i < it; ++i) res_hist[i] = hist[i];
  return JX_OK;
}

}  // extern "C"
