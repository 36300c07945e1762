// C ABI of janus_amd: handles, plan/twiddle cache and the fused fixed-point
// solver.  See include/janus_amd.h for the contract of each entry point.

#include <cmath>
#include <map>
#include <mutex>
#include <string>

#include "jx_fft_reg.h"
#include "jx_gamma.h"
#include "jx_handles.h"
#include "jx_internal.h"

namespace jx {

// ------------------------------------------------------------ errors
static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? JX_ENOMEM : JX_EHIP;
}

// ------------------------------------------------------------ grids
Grid make_grid(int dim, const int64_t *shape) {
  Grid g;
  g.dim = dim;
  g.N = 1;
  for (int d = 0; d < dim; ++d) {
    g.n[d] = shape[d];
    g.N *= shape[d];
  }
  g.L = shape[dim - 1];
  g.rows = g.N / g.L;
  g.nh = g.L / 2 + 1;
  g.Nh = g.rows * g.nh;
  // default register FFT plans, from timings at 256^3 (6 components): the
  // fused update pass is faster with the 8-value plan (no spilled
  // registers, 4 waves per SIMD), every other pass with the 16-value one
  if (g.L == 256) g.plan[PASS_FUSED] = 1;
  return g;
}

static int check_shape(int dim, const int64_t *shape) {
  if (dim != 2 && dim != 3)
    return fail(JX_EINVAL, "dim must be 2 or 3 (was " + std::to_string(dim) + ")");
  if (!shape) return fail(JX_EINVAL, "shape is NULL");
  for (int d = 0; d < dim; ++d)
    if (shape[d] <= 0)
      return fail(JX_EINVAL, "shape" + std::to_string(d) + " should be > 0 (was " +
                                 std::to_string(shape[d]) + ")");
  for (int d = 0; d < dim; ++d)
    if (shape[d] > (1 << 20))
      return fail(JX_EUNSUPPORTED, "grid dimension too large");
  return JX_OK;
}

// ------------------------------------------------------------ device limits
int lds_limit_bytes() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (dev >= 0 && dev < 64 && cache[dev] > 0) return cache[dev];
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess ||
      v <= 0)
    v = 65536;
  if (dev >= 0 && dev < 64) cache[dev] = v;
  return v;
}

// ------------------------------------------------------------ FFT plans
// Radix sequence: 8s, then one 4 or 2 for powers of two, then the odd prime
// factors in increasing order (handled by the direct-DFT stage).
static void factorize(int L, std::vector<int> &rad) {
  rad.clear();
  int m = L;
  while (m % 8 == 0) { rad.push_back(8); m /= 8; }
  if (m % 4 == 0) { rad.push_back(4); m /= 4; }
  if (m % 2 == 0) { rad.push_back(2); m /= 2; }
  for (int p = 3; (long long)p * p <= m; p += 2)
    while (m % p == 0) { rad.push_back(p); m /= p; }
  if (m > 1) rad.push_back(m);
}

// exp(-2 pi i m / L), reduced to the first octant and evaluated in long
// double so that each entry is the correctly rounded double in practice.
static void twiddle_host(int m, int L, double *re, double *im) {
  // angle = 2 pi m / L; use symmetries on the integer ratio
  long long a = (long long)m * 8;  // in units of L/8 turns: angle = a/L * (pi/4)
  long long q = a / L;             // octant index 0..7
  long long r = a - q * L;         // remainder, angle within octant = r/L * pi/4
  long double th = (long double)r / (long double)L * 0.78539816339744830961566084581987572L;
  long double c = cosl(th), s = sinl(th);
  // (cos, sin) of 2 pi m / L from the octant decomposition
  long double C, S;
  const long double h = 0.70710678118654752440084436210484903L;
  switch (q & 7) {
    case 0: C = c; S = s; break;
    case 1: C = h * (c - s); S = h * (c + s); break;
    case 2: C = -s; S = c; break;
    case 3: C = -h * (c + s); S = h * (c - s); break;
    case 4: C = -c; S = -s; break;
    case 5: C = -h * (c - s); S = -h * (c + s); break;
    case 6: C = s; S = -c; break;
    default: C = h * (c + s); S = -h * (c - s); break;
  }
  *re = (double)C;
  *im = (double)(-S);
}

struct LenEntry {
  FftLen plan;
  double2 *W = nullptr;
};

static std::mutex g_plan_mu;
static std::map<std::pair<int, int>, LenEntry> g_plans;  // (device, L)

// True when the plan of length L exists on the current device (a lookup
// only: nothing is allocated).
static bool fft_len_cached(int L) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  std::lock_guard<std::mutex> lock(g_plan_mu);
  return g_plans.find({dev, L}) != g_plans.end();
}

int get_fft_len(int L, FftLen *out) {
  if (L <= 0) return fail(JX_EINVAL, "FFT length must be > 0");
  int dev = 0;
  JX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_plan_mu);
  auto it = g_plans.find({dev, L});
  if (it != g_plans.end()) {
    *out = it->second.plan;
    return JX_OK;
  }
  std::vector<int> rad;
  factorize(L, rad);
  if ((int)rad.size() > kMaxStages) return fail(JX_EUNSUPPORTED, "too many FFT stages");
  std::vector<double2> W(L);
  for (int m = 0; m < L; ++m) twiddle_host(m, L, &W[m].x, &W[m].y);
  LenEntry e;
  JX_HIP(hipMalloc((void **)&e.W, sizeof(double2) * L));
  JX_HIP(hipMemcpy(e.W, W.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
  e.plan.L = L;
  e.plan.nstages = (int)rad.size();
  for (size_t i = 0; i < rad.size(); ++i) e.plan.radix[i] = rad[i];
  e.plan.W = e.W;
  g_plans[{dev, L}] = e;
  *out = e.plan;
  return JX_OK;
}

}  // namespace jx

using namespace jx;

static int local_alloc_partials(jx_local *l) {
  JX_HIP(hipMalloc((void **)&l->partials, sizeof(double) * (1024 + 1) * 2 * l->sym));
  return JX_OK;
}

extern "C" {

const char *jx_last_error(void) { return g_last_error.c_str(); }

int jx_version(void) { return 1; }

// ------------------------------------------------------------ FFT
int jx_fft_create(int dim, const int64_t *shape, jx_fft **out) {
  if (!out) return fail(JX_EINVAL, "out is NULL");
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  jx_fft *f = new jx_fft();
  f->g = make_grid(dim, shape);
  // one-component transforms: the 8-value plans of the last-axis passes
  // (twice the threads per row) measured faster at 128^3, 256^3 and 512^3
  // (r2c 0.205 -> 0.188 ms at 256^3, 1.92 -> 1.66 ms at 512^3); the
  // strided passes keep the 16-value plans (the 8-value c2r was slower)
  f->g.plan[PASS_ROWS_FWD] = f->g.plan[PASS_ROWS_INV] = 1;
  FftLen P;
  for (int d = 0; d < dim; ++d)
    if ((rc = get_fft_len((int)shape[d], &P))) {
      delete f;
      return rc;
    }
  hipError_t e = hipMalloc((void **)&f->scratch, sizeof(double) * 2 * f->g.Nh);
  if (e != hipSuccess) {
    delete f;
    return hip_fail(e, "hipMalloc(fft scratch)");
  }
  *out = f;
  return JX_OK;
}

int jx_fft_destroy(jx_fft *f) {
  if (!f) return JX_OK;
  if (f->scratch) (void)hipFree(f->scratch);
  delete f;
  return JX_OK;
}

int jx_fft_r2c(jx_fft *f, const double *r, double *c, void *stream) {
  if (!f || !r || !c) return fail(JX_EINVAL, "jx_fft_r2c: NULL argument");
  JX_ALIGNED(r, "jx_fft_r2c");
  JX_ALIGNED(c, "jx_fft_r2c");
  hipStream_t s = (hipStream_t)stream;
  LocalDev none;
  int rc = launch_rows_forward(f->g, 1, r, c, none, nullptr, s);
  if (rc) return rc;
  for (int ax = f->g.dim - 2; ax >= 0; --ax)
    if ((rc = launch_axis(f->g, 1, ax, -1, c, c, nullptr, nullptr, s))) return rc;
  return JX_OK;
}

int jx_fft_c2r(jx_fft *f, const double *c, double *r, void *stream) {
  if (!f || !r || !c) return fail(JX_EINVAL, "jx_fft_c2r: NULL argument");
  JX_ALIGNED(r, "jx_fft_c2r");
  JX_ALIGNED(c, "jx_fft_c2r");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  const double *src = c;
  for (int ax = 0; ax <= f->g.dim - 2; ++ax) {
    if ((rc = launch_axis(f->g, 1, ax, +1, src, f->scratch, nullptr, nullptr, s)))
      return rc;
    src = f->scratch;
  }
  RowsInverseLs none;
  return launch_rows_inverse(f->g, 1, src, r, none, nullptr, nullptr, s);
}

// ------------------------------------------------------------ Green
static void build_axis_tables(jx_green *G) {
  // per axis: kt, km, kp, wm, wp, sw, cw (7 * n_i doubles)
  const Grid &g = G->g;
  size_t total = 0;
  for (int d = 0; d < g.dim; ++d) total += 7 * (size_t)g.n[d];
  G->host_tables.assign(total, 0.);
  size_t off = 0;
  for (int d = 0; d < g.dim; ++d) {
    int64_t n = g.n[d];
    // s_i = 2 pi / (h n_i), green.pyx:140-141, 305-307
    double s = 2. * M_PI / (G->h * (double)n);
    double *kt = &G->host_tables[off];
    double *km = kt + n, *kp = km + n, *wm = kp + n, *wp = wm + n;
    double *sw = wp + n, *cw = sw + n;
    for (int64_t b = 0; b < n; ++b) {
      // truncated alias rule, green.pyx:432-445
      kt[b] = (2 * b > n) ? s * (double)(b - n) : s * (double)b;
      // filtered aliases and weights, green.pyx:740-800
      km[b] = s * (double)(b - n);
      double w = cos(0.25 * G->h * km[b]);
      wm[b] = w * w;
      kp[b] = s * (double)b;
      w = cos(0.25 * G->h * kp[b]);
      wp[b] = w * w;
      // Willot's modified wave vector, green.pyx:984-1023
      double phi = 0.5 * s * G->h * (double)b;
      sw[b] = sin(phi);
      cw[b] = cos(phi);
    }
    off += 7 * (size_t)n;
  }
}

static void point_tables(GreenDev &D, const Grid &g, const double *base) {
  size_t off = 0;
  for (int d = 0; d < g.dim; ++d) {
    int64_t n = g.n[d];
    D.ax[d].kt = base + off;
    D.ax[d].km = base + off + n;
    D.ax[d].kp = base + off + 2 * n;
    D.ax[d].wm = base + off + 3 * n;
    D.ax[d].wp = base + off + 4 * n;
    D.ax[d].sw = base + off + 5 * n;
    D.ax[d].cw = base + off + 6 * n;
    off += 7 * (size_t)n;
  }
}

// ------------------------------------------------------------ static scratch
// Spectrum scratch of jx_green_apply_static: a file-scope device array that
// every GPU the library runs on holds until the process ends, so that the
// call allocates nothing.  In host code its name is a host-side stand-in,
// not a device address: the kernels get the per-device address that
// hipGetSymbolAddress returns.
#define JX_STATIC_SCRATCH_BYTES (64ll << 20)
static __device__ __align__(256) double
    scratch[JX_STATIC_SCRATCH_BYTES / sizeof(double)];

static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;  // device -> address of scratch

// Address of `scratch` on the current device, looked up once per device.
static int static_scratch(double **out) {
  int dev = 0;
  JX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (ptr == nullptr) return fail(JX_EHIP, "hipGetSymbolAddress(scratch): NULL");
    it = g_scratch_ptr.emplace(dev, (double *)ptr).first;
  }
  *out = it->second;
  return JX_OK;
}

}  // extern "C"

// Device tables and FFT plans of a Green operator on the current device
// (small; no spectrum), and the static scratch's address there.
static int green_prepare_device(jx_green *G) {
  if (G->tables) return JX_OK;
  int dev = 0;
  JX_HIP(hipGetDevice(&dev));
  JX_HIP(hipMalloc((void **)&G->tables, sizeof(double) * G->host_tables.size()));
  JX_HIP(hipMemcpy(G->tables, G->host_tables.data(),
                   sizeof(double) * G->host_tables.size(), hipMemcpyHostToDevice));
  G->device = dev;
  point_tables(G->dev, G->g, G->tables);
  FftLen P;
  for (int d = 0; d < G->g.dim; ++d) {
    int rc = get_fft_len((int)G->g.n[d], &P);
    if (rc) return rc;
  }
  double *unused = nullptr;
  return static_scratch(&unused);
}

extern "C" {

int jx_green_create(int dim, const int64_t *shape, double h, int kind, double g0,
                    double nu0, jx_green **out) {
  if (!out) return fail(JX_EINVAL, "out is NULL");
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  if (!(h > 0.)) return fail(JX_EINVAL, "h must be > 0");
  if (kind != JX_GREEN_TRUNCATED && kind != JX_GREEN_FILTERED &&
      kind != JX_GREEN_WILLOT)
    return fail(JX_EINVAL, "unknown Green operator kind");
  if (!(g0 > 0.)) return fail(JX_EINVAL, "g0 must be > 0");
  jx_green *G = new jx_green();
  G->g = make_grid(dim, shape);
  G->sym = dim == 2 ? 3 : 6;
  G->h = h;
  G->g0 = g0;
  G->nu0 = nu0;
  GreenDev D;
  D.dim = dim;
  D.kind = kind;
  // isotropic.pyx:145-156
  D.daux1 = 1.0 / g0;
  D.daux2 = 0.5 / (g0 * (1.0 - nu0));
  D.daux3 = 0.25 / g0;
  D.daux4 = 0.5 / g0;
  build_axis_tables(G);
  G->host = D;
  point_tables(G->host, G->g, G->host_tables.data());
  G->dev = D;
  // with a GPU present the device tables and FFT plans are made now (the
  // spectrum waits for the first apply); without one the host-side
  // per-frequency API still works
  int count = 0;
  if (hipGetDeviceCount(&count) == hipSuccess && count > 0) {
    rc = green_prepare_device(G);
    if (rc) {
      jx_green_destroy(G);
      return rc;
    }
  } else {
    (void)hipGetLastError();
  }
  *out = G;
  return JX_OK;
}

}  // extern "C"

// Device tables, spectrum scratch and FFT plans, on first use when the
// handle was created without a GPU.
int jx::green_ensure_device(jx_green *G) {
  if (G->spec) return JX_OK;
  int rc = green_prepare_device(G);
  if (rc) return rc;
  JX_HIP(hipMalloc((void **)&G->spec, sizeof(double) * spec_doubles(G->g, G->sym)));
  return JX_OK;
}

extern "C" {

int jx_green_destroy(jx_green *G) {
  if (!G) return JX_OK;
  if (G->tables) (void)hipFree(G->tables);
  if (G->spec) (void)hipFree(G->spec);
  delete G;
  return JX_OK;
}

static int green_passes(jx_green *G, double *spec, const double *x, double *y,
                        const LocalDev &loc, const RowsInverseLs &ls,
                        const SolveState *st, int *nb, hipStream_t s) {
  const Grid &g = G->g;
  int C = G->sym;
  int rc = launch_rows_forward(g, C, x, spec, loc, st, s);
  if (rc) return rc;
  if (g.dim == 3 && (rc = launch_axis(g, C, plain_axis(g), -1, spec, spec, nullptr, st, s)))
    return rc;
  if ((rc = launch_axis(g, C, pivot_axis(g), 0, spec, spec, &G->dev, st, s))) return rc;
  if (g.dim == 3 && (rc = launch_axis(g, C, plain_axis(g), +1, spec, spec, nullptr, st, s)))
    return rc;
  return launch_rows_inverse(g, C, spec, y, ls, st, nb, s);
}

int jx_green_apply(jx_green *G, const double *x, double *y, void *stream) {
  if (!G || !x || !y) return fail(JX_EINVAL, "jx_green_apply: NULL argument");
  JX_ALIGNED(x, "jx_green_apply");
  JX_ALIGNED(y, "jx_green_apply");
  int rc = green_ensure_device(G);
  if (rc) return rc;
  LocalDev none;
  RowsInverseLs plain;
  return green_passes(G, G->spec, x, y, none, plain, nullptr, nullptr,
                      (hipStream_t)stream);
}

int jx_green_apply_static(jx_green *G, const double *x, double *y,
                          void *stream) {
  if (!G || !x || !y) return fail(JX_EINVAL, "jx_green_apply_static: NULL argument");
  JX_ALIGNED(x, "jx_green_apply_static");
  JX_ALIGNED(y, "jx_green_apply_static");
  const long long need = (long long)sizeof(double) * spec_doubles(G->g, G->sym);
  if (need > JX_STATIC_SCRATCH_BYTES)
    return fail(JX_EINVAL, "jx_green_apply_static: the grid needs " +
                               std::to_string(need) + " bytes of spectrum "
                               "scratch, more than the static " +
                               std::to_string(JX_STATIC_SCRATCH_BYTES) +
                               " (use jx_green_apply)");
  int dev = 0;
  JX_HIP(hipGetDevice(&dev));
  if (!G->tables || G->device != dev)
    return fail(JX_EINVAL, "jx_green_apply_static: the operator's device tables "
                           "are not on the current device (create it with "
                           "that device current)");
  for (int d = 0; d < G->g.dim; ++d)
    if (!fft_len_cached((int)G->g.n[d]))
      return fail(JX_EINVAL, "jx_green_apply_static: FFT plans missing on the "
                             "current device");
  double *spec = nullptr;
  int rc = static_scratch(&spec);
  if (rc) return rc;
  LocalDev none;
  RowsInverseLs plain;
  return green_passes(G, spec, x, y, none, plain, nullptr, nullptr,
                      (hipStream_t)stream);
}

int jx_green_matrix_at(jx_green *G, const int32_t *b, double *out) {
  if (!G || !b || !out) return fail(JX_EINVAL, "jx_green_matrix_at: NULL argument");
  const Grid &g = G->g;
  int bb[3];
  for (int d = 0; d < g.dim; ++d) {
    if (b[d] < 0 || b[d] >= g.n[d])
      return fail(JX_EINVAL, "index must be >= 0 and < " + std::to_string(g.n[d]) +
                                 " (was " + std::to_string(b[d]) + ")");
    bb[d] = b[d];
  }
  int C = G->sym;
  double t[21];
  if (g.dim == 2)
    gamma_hat<2>(G->host, bb, t);
  else
    gamma_hat<3>(G->host, bb, t);
  for (int i = 0; i < C; ++i)
    for (int j = 0; j < C; ++j)
      out[i * C + j] = t[tri_index(C, i < j ? i : j, i < j ? j : i)];
  return JX_OK;
}

int jx_gamma0_at(int dim, double g, double nu, const double *k, double *out) {
  if (!k || !out) return fail(JX_EINVAL, "jx_gamma0_at: NULL argument");
  if (dim != 2 && dim != 3) return fail(JX_EINVAL, "dim must be 2 or 3");
  double d1 = 1.0 / g, d2 = 0.5 / (g * (1.0 - nu)), d3 = 0.25 / g, d4 = 0.5 / g;
  double t[21];
  int C = dim == 2 ? 3 : 6;
  if (dim == 2)
    gamma0_2d(k[0], k[1], d1, d2, d3, d4, t);
  else
    gamma0_3d(k[0], k[1], k[2], d1, d2, d3, d4, t);
  for (int i = 0; i < C; ++i)
    for (int j = 0; j < C; ++j)
      out[i * C + j] = t[tri_index(C, i < j ? i : j, i < j ? j : i)];
  return JX_OK;
}

// ------------------------------------------------------------ local
int jx_local_create_phases(int dim, const int64_t *shape, const uint8_t *phase,
                           const double *table, int nphase, jx_local **out) {
  if (!out) return fail(JX_EINVAL, "out is NULL");
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  if (!phase || !table) return fail(JX_EINVAL, "phase/table is NULL");
  if (nphase < 1 || nphase > 256) return fail(JX_EINVAL, "nphase must be in [1, 256]");
  jx_local *l = new jx_local();
  l->g = make_grid(dim, shape);
  l->sym = dim == 2 ? 3 : 6;
  std::vector<double> tab(2 * nphase);
  for (int p = 0; p < nphase; ++p) {
    double sph = table[2 * p], dv = table[2 * p + 1];
    tab[2 * p] = (sph - dv) / (double)dim;  // tr, operators.pyx:299
    tab[2 * p + 1] = dv;
  }
  hipError_t e = hipMalloc((void **)&l->table, sizeof(double) * tab.size());
  if (e == hipSuccess)
    e = hipMemcpy(l->table, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    jx_local_destroy(l);
    return hip_fail(e, "jx_local_create_phases");
  }
  l->dev.mode = LOCAL_PHASES;
  l->dev.phase = phase;
  l->dev.table = l->table;
  if ((rc = local_alloc_partials(l))) {
    jx_local_destroy(l);
    return rc;
  }
  *out = l;
  return JX_OK;
}

int jx_local_create_fields(int dim, const int64_t *shape, const double *sph,
                           const double *dev, jx_local **out) {
  if (!out) return fail(JX_EINVAL, "out is NULL");
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  if (!sph || !dev) return fail(JX_EINVAL, "sph/dev is NULL");
  jx_local *l = new jx_local();
  l->g = make_grid(dim, shape);
  l->sym = dim == 2 ? 3 : 6;
  l->dev.mode = LOCAL_FIELDS;
  l->dev.sph = sph;
  l->dev.dev = dev;
  if ((rc = local_alloc_partials(l))) {
    jx_local_destroy(l);
    return rc;
  }
  *out = l;
  return JX_OK;
}

int jx_local_create_phase_matrices(int dim, const int64_t *shape,
                                   const uint8_t *phase, const double *mats,
                                   int nphase, jx_local **out) {
  if (!out) return fail(JX_EINVAL, "out is NULL");
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  if (!phase || !mats) return fail(JX_EINVAL, "phase/mats is NULL");
  if (nphase < 1 || nphase > 256) return fail(JX_EINVAL, "nphase must be in [1, 256]");
  jx_local *l = new jx_local();
  l->g = make_grid(dim, shape);
  int C = dim == 2 ? 3 : 6;
  l->sym = C;
  size_t count = (size_t)nphase * C * C;
  for (int p = 0; p < nphase; ++p)
    for (int i = 0; i < C; ++i)
      for (int j = i + 1; j < C; ++j)
        if (mats[(size_t)p * C * C + i * C + j] != mats[(size_t)p * C * C + j * C + i])
          l->symmetric = false;
  hipError_t e = hipMalloc((void **)&l->table, sizeof(double) * count);
  if (e == hipSuccess)
    e = hipMemcpy(l->table, mats, sizeof(double) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    jx_local_destroy(l);
    return hip_fail(e, "jx_local_create_phase_matrices");
  }
  l->dev.mode = LOCAL_PHASE_MATRIX;
  l->dev.phase = phase;
  l->dev.mat = l->table;
  if ((rc = local_alloc_partials(l))) {
    jx_local_destroy(l);
    return rc;
  }
  *out = l;
  return JX_OK;
}

int jx_local_create_matrix_field(int dim, const int64_t *shape, const double *mats,
                                 jx_local **out) {
  if (!out) return fail(JX_EINVAL, "out is NULL");
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  if (!mats) return fail(JX_EINVAL, "mats is NULL");
  jx_local *l = new jx_local();
  l->g = make_grid(dim, shape);
  l->sym = dim == 2 ? 3 : 6;
  l->dev.mode = LOCAL_MATRIX_FIELD;
  l->dev.mat = mats;
  if ((rc = local_alloc_partials(l)) ||
      (rc = matrices_symmetric(l->sym, mats, l->g.N, (int *)l->partials,
                               &l->symmetric))) {
    jx_local_destroy(l);
    return rc;
  }
  *out = l;
  return JX_OK;
}

int jx_local_is_symmetric(jx_local *l, int *out) {
  if (!l || !out) return fail(JX_EINVAL, "jx_local_is_symmetric: NULL argument");
  *out = l->symmetric ? 1 : 0;
  return JX_OK;
}

int jx_local_destroy(jx_local *l) {
  if (!l) return JX_OK;
  if (l->table) (void)hipFree(l->table);
  if (l->partials) (void)hipFree(l->partials);
  delete l;
  return JX_OK;
}

int jx_local_apply(jx_local *l, const double *x, double *y, void *stream) {
  if (!l || !x || !y) return fail(JX_EINVAL, "jx_local_apply: NULL argument");
  JX_ALIGNED(x, "jx_local_apply");
  JX_ALIGNED(y, "jx_local_apply");
  return launch_local_apply(l->g, l->sym, l->dev, x, y, (hipStream_t)stream);
}

int jx_local_apply_transpose(jx_local *l, const double *x, double *y,
                             void *stream) {
  if (!l || !x || !y) return fail(JX_EINVAL, "jx_local_apply_transpose: NULL argument");
  JX_ALIGNED(x, "jx_local_apply_transpose");
  JX_ALIGNED(y, "jx_local_apply_transpose");
  return launch_local_apply(l->g, l->sym, l->dev, x, y, (hipStream_t)stream, true);
}

int jx_block_linear_apply(int dim, const int64_t *shape, int m, int n,
                          const double *a, const double *x, double *y,
                          int transpose, void *stream) {
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  if (!a || !x || !y) return fail(JX_EINVAL, "jx_block_linear_apply: NULL argument");
  if (m < 1 || n < 1) return fail(JX_EINVAL, "block sizes must be >= 1");
  if ((const double *)y == x)
    return fail(JX_EINVAL, "jx_block_linear_apply: x and y must not overlap");
  int64_t N = 1;
  for (int d = 0; d < dim; ++d) N *= shape[d];
  return launch_block_linear(N, m, n, a, x, y, transpose != 0, (hipStream_t)stream);
}

int jx_local_means(jx_local *l, const double *x, double *out_tau, double *out_eps,
                   void *stream) {
  if (!l || !x) return fail(JX_EINVAL, "jx_local_means: NULL argument");
  JX_ALIGNED(x, "jx_local_means");
  hipStream_t s = (hipStream_t)stream;
  int C = l->sym;
  double *partials = l->partials;
  double *cols = l->partials + 1024 * 2 * C;
  int nb = 0;
  int rc = launch_local_sums(l->g, C, l->dev, x, partials, &nb, s);
  if (!rc) rc = launch_reduce_columns(partials, nb, 2 * C, cols, s);
  double host[12];
  if (!rc) {
    hipError_t e = hipMemcpyAsync(host, cols, sizeof(double) * 2 * C,
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "jx_local_means");
  }
  if (rc) return rc;
  double invN = 1.0 / (double)l->g.N;
  for (int c = 0; c < C; ++c) {
    if (out_tau) out_tau[c] = host[c] * invN;
    if (out_eps) out_eps[c] = host[C + c] * invN;
  }
  return JX_OK;
}

// ------------------------------------------------------------ solver
int jx_solver_create(jx_green *G, jx_local *l, jx_solver **out) {
  if (!G || !l || !out) return fail(JX_EINVAL, "jx_solver_create: NULL argument");
  if (G->g.dim != l->g.dim)
    return fail(JX_EINVAL, "Green operator and local operator dimensions differ");
  for (int d = 0; d < G->g.dim; ++d)
    if (G->g.n[d] != l->g.n[d])
      return fail(JX_EINVAL, "Green operator and local operator grids differ");
  int rc = green_ensure_device(G);
  if (rc) return rc;
  jx_solver *S = new jx_solver();
  S->green = G;
  S->local = l;
  S->nblocks = rows_inverse_nblocks(G->g, G->sym);
  hipError_t e = hipMalloc((void **)&S->st, sizeof(SolveState));
  if (e == hipSuccess) e = hipHostMalloc((void **)&S->st_host, sizeof(SolveState), 0);
  if (e == hipSuccess) e = hipMalloc((void **)&S->partials, sizeof(double) * S->nblocks);
  if (e != hipSuccess) {
    jx_solver_destroy(S);
    return hip_fail(e, "jx_solver_create");
  }
  *out = S;
  return JX_OK;
}

int jx_solver_destroy(jx_solver *S) {
  if (!S) return JX_OK;
  if (S->st) (void)hipFree(S->st);
  if (S->st_host) (void)hipHostFree(S->st_host);
  if (S->partials) (void)hipFree(S->partials);
  if (S->res_dev) (void)hipFree(S->res_dev);
  if (S->cg_buf) (void)hipFree(S->cg_buf);
  if (S->cg_partials) (void)hipFree(S->cg_partials);
  if (S->cg_s) (void)hipFree(S->cg_s);
  delete S;
  return JX_OK;
}

}  // extern "C"

// One fixed-point iteration on iterate eps, spectrum buffer spec and solve
// state st.  Fused schedule: pass A runs only on the first iteration; every
// iteration then runs B, C, D and the fused E+A pass, which leaves the
// spectrum of the next iteration's polarization in place.
int jx::solver_iteration(jx_solver *S, double *eps, double *spec,
                         SolveState *st, const RowsInverseLs &ls, bool first,
                         int *nb, hipStream_t s) {
  jx_green *G = S->green;
  const Grid &g = G->g;
  const int C = G->sym;
  if (!(S->fused && rows_fusable(g, C)))
    return green_passes(G, spec, eps, eps, S->local->dev, ls, st, nb, s);
  int rc;
  if (first &&
      (rc = launch_rows_forward(g, C, eps, spec, S->local->dev, st, s)))
    return rc;
  if (g.dim == 3 && (rc = launch_axis(g, C, plain_axis(g), -1, spec, spec, nullptr, st, s)))
    return rc;
  if ((rc = launch_axis(g, C, pivot_axis(g), 0, spec, spec, &G->dev, st, s))) return rc;
  if (g.dim == 3 && (rc = launch_axis(g, C, plain_axis(g), +1, spec, spec, nullptr, st, s)))
    return rc;
  return launch_rows_update_forward(g, C, spec, eps, ls, S->local->dev, st, nb,
                                    s);
}

extern "C" {

int jx_solver_set_fused(jx_solver *S, int fused) {
  if (!S) return fail(JX_EINVAL, "jx_solver_set_fused: NULL solver");
  S->fused = fused != 0;
  return JX_OK;
}

int jx_solver_get_fused(jx_solver *S) {
  if (!S) return -1;
  return (S->fused && rows_fusable(S->green->g, S->green->sym)) ? 1 : 0;
}

int jx_green_set_plan(jx_green *G, int pass, int variant) {
  if (!G) return fail(JX_EINVAL, "jx_green_set_plan: NULL operator");
  if (pass < 0 || pass >= kPlanPasses)
    return fail(JX_EINVAL, "pass must be in [0, 5)");
  if (variant < 0 || variant > 1)
    return fail(JX_EINVAL, "variant must be 0 or 1");
  G->g.plan[pass] = variant;
  return JX_OK;
}

int jx_green_get_plan(jx_green *G, int pass) {
  if (!G || pass < 0 || pass >= kPlanPasses) return -1;
  return G->g.plan[pass];
}

}  // extern "C"

namespace {

// Average duration (ms) of `iters` launches of one pass of solver S on
// scratch fields, after one untimed launch.
int time_pass(jx_solver *S, int pass, double *eps, int iters, hipStream_t s,
              hipEvent_t e0, hipEvent_t e1, double *ms) {
  jx_green *G = S->green;
  const Grid &g = G->g;
  const int C = G->sym;
  RowsInverseLs ls;
  ls.enabled = true;
  ls.partials = S->partials;
  int rc = JX_OK, nb = 0;
  for (int it = 0; it <= iters && !rc; ++it) {
    if (it == 1) (void)hipEventRecord(e0, s);
    switch (pass) {
      case PASS_ROWS_FWD:
        rc = launch_rows_forward(g, C, eps, G->spec, S->local->dev, S->st, s);
        break;
      case PASS_AXIS:
        rc = launch_axis(g, C, g.dim == 3 ? plain_axis(g) : 0, -1, G->spec, G->spec,
                         nullptr, S->st, s);
        break;
      case PASS_GAMMA:
        rc = launch_axis(g, C, pivot_axis(g), 0, G->spec, G->spec, &G->dev, S->st, s);
        break;
      case PASS_ROWS_INV:
        rc = launch_rows_inverse(g, C, G->spec, eps, ls, S->st, &nb, s);
        break;
      default:
        rc = launch_rows_update_forward(g, C, G->spec, eps, ls, S->local->dev,
                                        S->st, &nb, s);
    }
  }
  (void)hipEventRecord(e1, s);
  hipError_t e = hipEventSynchronize(e1);
  if (!rc && e != hipSuccess) rc = hip_fail(e, "jx_solver_autotune");
  float t = 0.f;
  if (!rc) (void)hipEventElapsedTime(&t, e0, e1);
  *ms = t / iters;
  return rc;
}

}  // namespace

extern "C" {

int jx_solver_autotune(jx_solver *S, int iters, double *ms_out, void *stream) {
  if (!S || iters < 1) return fail(JX_EINVAL, "jx_solver_autotune: invalid argument");
  hipStream_t s = (hipStream_t)stream;
  jx_green *G = S->green;
  Grid &g = G->g;
  const int C = G->sym;
  if (ms_out)
    for (int i = 0; i < 2 * kPlanPasses; ++i) ms_out[i] = 0.;
  // scratch iterate (zero: every pass stays finite) and a running state
  double *eps = nullptr;
  JX_HIP(hipMalloc((void **)&eps, sizeof(double) * C * g.N));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = JX_OK;
  hipError_t he = hipMemsetAsync(eps, 0, sizeof(double) * C * g.N, s);
  if (he == hipSuccess) he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  if (he != hipSuccess) rc = hip_fail(he, "jx_solver_autotune");
  SolveState init = {};
  init.tol = 0.;
  init.res_scale = 1.;
  init.max_iter = 1 << 30;
  *S->st_host = init;
  if (!rc) {
    he = hipMemcpyAsync(S->st, S->st_host, sizeof(SolveState),
                        hipMemcpyHostToDevice, s);
    if (he != hipSuccess) rc = hip_fail(he, "jx_solver_autotune");
  }
  const bool fusable = rows_fusable(g, C);
  double best[kPlanPasses];
  for (int p = 0; p < kPlanPasses && !rc; ++p) {
    best[p] = 0.;
    if (p == PASS_AXIS && g.dim != 3) continue;
    if (p == PASS_FUSED && !fusable) continue;
    const int len = (p == PASS_AXIS) ? (int)g.n[1]
                    : (p == PASS_GAMMA) ? (int)g.n[0] : (int)g.L;
    const int nv = reg_fft_length(len) ? reg_variants(len) : 1;
    int keep = g.plan[p];
    for (int k = 0; k < nv && !rc; ++k) {
      g.plan[p] = k;
      double ms = 0.;
      rc = time_pass(S, p, eps, iters, s, e0, e1, &ms);
      if (ms_out) ms_out[2 * p + k] = ms;
      if (!rc && (k == 0 || ms < best[p])) {
        best[p] = ms;
        keep = k;
      }
    }
    g.plan[p] = keep;
  }
  // schedule: fused pass against the separate passes A and E
  if (!rc && fusable)
    S->fused = best[PASS_FUSED] <= best[PASS_ROWS_FWD] + best[PASS_ROWS_INV];
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipFree(eps);
  return rc;
}

int jx_solver_solve(jx_solver *S, double *eps, const double *E, const double *rhs,
                    double tol, int max_iter, double res_scale, double *res_hist,
                    int *iters, int *status, void *stream) {
  if (!S || !eps || !E) return fail(JX_EINVAL, "jx_solver_solve: NULL argument");
  JX_ALIGNED(eps, "jx_solver_solve");
  JX_ALIGNED(rhs, "jx_solver_solve");
  if (max_iter < 1) return fail(JX_EINVAL, "max_iter must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  jx_green *G = S->green;
  const Grid &g = G->g;
  int C = G->sym;
  if (res_scale <= 0.) {
    double nE = 0.;
    for (int c = 0; c < C; ++c) nE += E[c] * E[c];
    nE = std::sqrt(nE);
    if (nE == 0.)
      return fail(JX_EINVAL, "res_scale must be given when E = 0");
    res_scale = 1.0 / std::sqrt((double)g.N) / nE;  // square_basic.py:86
  }
  if (S->res_cap < max_iter) {
    if (S->res_dev) (void)hipFree(S->res_dev);
    S->res_dev = nullptr;
    S->res_cap = 0;
    JX_HIP(hipMalloc((void **)&S->res_dev, sizeof(double) * max_iter));
    S->res_cap = max_iter;
  }
  SolveState init;
  init.iter = 0;
  init.done = 0;
  init.tol = tol;
  init.res_scale = res_scale;
  init.max_iter = max_iter;
  init.pad = 0;
  for (int c = 0; c < 6; ++c) init.E[c] = c < C ? E[c] : 0.;
  *S->st_host = init;
  JX_HIP(hipMemcpyAsync(S->st, S->st_host, sizeof(SolveState), hipMemcpyHostToDevice, s));
  JX_HIP(hipStreamSynchronize(s));

  RowsInverseLs ls;
  ls.enabled = true;
  ls.rhs = rhs;
  ls.partials = S->partials;
  // Enqueue iterations in chunks; kernels of iterations past convergence
  // return at once (the device state says done), so the iteration count is
  // exact while the host only synchronizes once per chunk.
  int enqueued = 0;
  int chunk = 4;
  while (true) {
    int n = std::min(chunk, max_iter - enqueued);
    for (int i = 0; i < n; ++i) {
      int nb = 0;
      int rc = solver_iteration(S, eps, G->spec, S->st, ls, enqueued + i == 0,
                                &nb, s);
      if (rc) return rc;
      if ((rc = launch_finalize(S->partials, nb, S->st, S->res_dev, s))) return rc;
    }
    enqueued += n;
    JX_HIP(hipMemcpyAsync(S->st_host, S->st, sizeof(SolveState), hipMemcpyDeviceToHost, s));
    JX_HIP(hipStreamSynchronize(s));
    if (S->st_host->done || enqueued >= max_iter) break;
    chunk = std::min(chunk * 2, 64);
  }
  int it = S->st_host->iter;
  if (iters) *iters = it;
  if (status) *status = S->st_host->done ? S->st_host->done : JX_SOLVE_MAXITER;
  if (res_hist && it > 0)
    JX_HIP(hipMemcpy(res_hist, S->res_dev, sizeof(double) * it, hipMemcpyDeviceToHost));
  return JX_OK;
}

int jx_stream_copy(const void *src, void *dst, int64_t bytes, void *stream) {
  if (!src || !dst || bytes < 0 || bytes % 16)
    return fail(JX_EINVAL, "jx_stream_copy: bytes must be a multiple of 16");
  if (((uintptr_t)src | (uintptr_t)dst) % 16)
    return fail(JX_EINVAL, "jx_stream_copy: pointers must be 16-byte aligned");
  return launch_stream_copy(src, dst, bytes / 16, (hipStream_t)stream);
}

int jx_tile_probe_bytes(const int64_t *shape, int64_t *bytes) {
  int rc = check_shape(3, shape);
  if (rc) return rc;
  if (!bytes) return fail(JX_EINVAL, "jx_tile_probe_bytes: bytes is NULL");
  *bytes = (int64_t)sizeof(double) * spec_doubles(make_grid(3, shape), 6);
  return JX_OK;
}

int jx_tile_probe(const int64_t *shape, int axis, int points, const void *src,
                  void *dst, int64_t bytes, void *stream) {
  int rc = check_shape(3, shape);
  if (rc) return rc;
  if (!src || !dst) return fail(JX_EINVAL, "jx_tile_probe: NULL argument");
  JX_ALIGNED(src, "jx_tile_probe");
  JX_ALIGNED(dst, "jx_tile_probe");
  const Grid g = make_grid(3, shape);
  if (bytes < (int64_t)sizeof(double) * spec_doubles(g, 6))
    return fail(JX_EINVAL, "jx_tile_probe: buffers smaller than the spectrum layout");
  return tile_probe(g, axis, points, (const double *)src, (double *)dst,
                    (hipStream_t)stream);
}

int jx_adjoint_contract(int dim, const int64_t *shape, const double *eps,
                        const double *mu, double *dsph, double *ddev,
                        void *stream) {
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  if (!eps || !mu || !dsph || !ddev)
    return fail(JX_EINVAL, "jx_adjoint_contract: NULL argument");
  JX_ALIGNED(eps, "jx_adjoint_contract");
  JX_ALIGNED(mu, "jx_adjoint_contract");
  Grid g = make_grid(dim, shape);
  return launch_adjoint_contract(g, dim == 2 ? 3 : 6, eps, mu, dsph, ddev,
                                 (hipStream_t)stream);
}

int jx_adjoint_outer(int dim, const int64_t *shape, const double *eps,
                     const double *mu, double *da, void *stream) {
  int rc = check_shape(dim, shape);
  if (rc) return rc;
  if (!eps || !mu || !da)
    return fail(JX_EINVAL, "jx_adjoint_outer: NULL argument");
  Grid g = make_grid(dim, shape);
  return launch_adjoint_outer(g, dim == 2 ? 3 : 6, eps, mu, da,
                              (hipStream_t)stream);
}

int jx_solver_profile(jx_solver *S, double *eps, const double *E, int iters,
                      double *ms_out, void *stream) {
  if (!S || !eps || !E || !ms_out || iters < 1)
    return fail(JX_EINVAL, "jx_solver_profile: invalid argument");
  hipStream_t s = (hipStream_t)stream;
  jx_green *G = S->green;
  const Grid &g = G->g;
  int C = G->sym;
  if (S->res_cap < iters + 1) {
    if (S->res_dev) (void)hipFree(S->res_dev);
    S->res_dev = nullptr;
    S->res_cap = 0;
    JX_HIP(hipMalloc((void **)&S->res_dev, sizeof(double) * (iters + 1)));
    S->res_cap = iters + 1;
  }
  SolveState init;
  init.iter = 0;
  init.done = 0;
  init.tol = 0.;
  init.res_scale = 1.;
  init.max_iter = iters + 1;
  init.pad = 0;
  for (int c = 0; c < 6; ++c) init.E[c] = c < C ? E[c] : 0.;
  *S->st_host = init;
  JX_HIP(hipMemcpyAsync(S->st, S->st_host, sizeof(SolveState), hipMemcpyHostToDevice, s));
  RowsInverseLs ls;
  ls.enabled = true;
  ls.partials = S->partials;
  const int NP = 6;
  const bool fused = S->fused && rows_fusable(g, C);
  std::vector<hipEvent_t> ev((size_t)iters * (NP + 1));
  for (auto &e : ev) JX_HIP(hipEventCreate(&e));
  // one untimed iteration (it carries pass A in the fused schedule)
  int nb0 = 0;
  int rc = solver_iteration(S, eps, G->spec, S->st, ls, true, &nb0, s);
  if (!rc) rc = launch_finalize(S->partials, nb0, S->st, S->res_dev, s);
  for (int it = 0; it < iters && !rc; ++it) {
    hipEvent_t *e = &ev[(size_t)it * (NP + 1)];
    int nb = 0;
    (void)hipEventRecord(e[0], s);
    if (!fused) rc = launch_rows_forward(g, C, eps, G->spec, S->local->dev, S->st, s);
    (void)hipEventRecord(e[1], s);
    if (!rc && g.dim == 3) rc = launch_axis(g, C, plain_axis(g), -1, G->spec, G->spec, nullptr, S->st, s);
    (void)hipEventRecord(e[2], s);
    if (!rc) rc = launch_axis(g, C, pivot_axis(g), 0, G->spec, G->spec, &G->dev, S->st, s);
    (void)hipEventRecord(e[3], s);
    if (!rc && g.dim == 3) rc = launch_axis(g, C, plain_axis(g), +1, G->spec, G->spec, nullptr, S->st, s);
    (void)hipEventRecord(e[4], s);
    if (!rc)
      rc = fused ? launch_rows_update_forward(g, C, G->spec, eps, ls,
                                              S->local->dev, S->st, &nb, s)
                 : launch_rows_inverse(g, C, G->spec, eps, ls, S->st, &nb, s);
    (void)hipEventRecord(e[5], s);
    if (!rc) rc = launch_finalize(S->partials, nb, S->st, S->res_dev, s);
    (void)hipEventRecord(e[6], s);
  }
  hipError_t herr = hipStreamSynchronize(s);
  if (!rc && herr != hipSuccess) rc = hip_fail(herr, "jx_solver_profile");
  for (int p = 0; p < NP; ++p) ms_out[p] = 0.;
  if (!rc) {
    for (int it = 0; it < iters; ++it)
      for (int p = 0; p < NP; ++p) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[(size_t)it * (NP + 1) + p],
                                  ev[(size_t)it * (NP + 1) + p + 1]);
        ms_out[p] += ms;
      }
    for (int p = 0; p < NP; ++p) ms_out[p] /= iters;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  return rc;
}

}  // extern "C"
