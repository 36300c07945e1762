"""numpy restatement of janus's serial Lippmann-Schwinger path (FP64).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by ``janus_amd``.

What each function restates, and what pins it (the tests in
``tests/test_oracle.py`` carry out every pin listed here):

===========================  ==============================================  ==========================================
function                     janus source                                    pinned by
===========================  ==============================================  ==========================================
bulk_modulus                 material/elastic/linear/isotropic.pyx:79-97     test_isotropic_linear_elastic_material.py:14-20 values
gamma0                       isotropic.pyx:133-156, 173-200, 242-308         tensor closed form (test_greenop.py:21-35) within
                                                                             325 nulp; printed matrix sphinx/materials.rst:53-68
truncated_wavevector         green.pyx:432-445, 505-523                      brute-force alias rule (test_discrete_green_operator.py:214-222)
gamma_hat_truncated          green.pyx:432-451, 505-529                      same; 18 fixtures tests/data/*.npz
gamma_hat_filtered           green.pyx:607-694, 730-910                      brute-force alias sum
                                                                             (test_discrete_green_operator.py:249-260); fixtures
green_apply                  green.pyx:182-216, 353-390, 453-495, 531-580    18 fixtures (rtol=ULP, atol=10 ULP)
iso4_apply                   operators.pyx:292-299, 315-324, 352-365         J/K projector matrices (test_operators.py:180-227)
block_diag_apply             operators.pyx:655-664, 706-716                  per-voxel loop (test_operators.py:409-450)
cubic_tensor_2d              operators.pyx:411-450                           rotation of the 4th-rank cubic tensor in
                                                                             index notation, T'_ijkl = R R R R T
                                                                             (test_operators.py:238-320 cases)
matrix_apply                 operators.pyx:739-775, 824-861                  explicit per-voxel loops on a tiny grid;
                                                                             <y, A x> = <A^T y, x>
                                                                             (test_operators.py:452-505)
basic_scheme                 sphinx/tutorials/square_basic/square_basic.py   published C_1212 = 1.41903971282
                             :17-97                                          (square_basic.rst:227); matrix form (mats)
                                                                             = isotropic form on iso4_matrix blocks;
                                                                             laminate closed form (Voigt/Reuss mix)
rfft / irfft                 fft/serial/_serial_fft.pyx:152-177, 296-325     numpy.fft (test_fft.py:37-70) -- the
                                                                             oracle's FFT *is* numpy's pocketfft
adjoint_gradient             (no janus counterpart) implicit                 central finite differences of basic_scheme
                             differentiation of the fixed point
willot_wavevector /          green.pyx:976-1036 (willot2015)                 Fourier symbol of the rotated-grid
gamma_hat_willot                                                             finite-difference gradient (brute-force DFT
                                                                             of the stencil); gamma0 pins
cg_scheme                    examples/square_inclusion.py:40-47, 71-86       converged solution solves the Lippmann-
                             (Krylov on the same operators; strain form      Schwinger equation and equals the basic
                             in the C0 inner product, no janus counterpart)  scheme's; C0-conjugacy of the directions;
                                                                             dense direct solve on a 4x4 grid
cg_polarization_scheme       examples/square_inclusion.py:17-52, 77-85       dense direct solve of (L + Gamma0) tau = E;
                             (PETSc CG on the polarization form)             eps = L tau equals the strain-form solution
                                                                             (cg_scheme at the same C0); the janus
                                                                             example's own configuration
gmres_scheme                 examples/square_inclusion_scipy.py:52-59, 88    dense direct solve on 4x4 grids (symmetric
                             (GMRES on the strain form, restarted; no        and non-symmetric dC, with and without
                             janus counterpart)                              restarts); Arnoldi estimate = true residual;
                                                                             finite termination; equals CG's solution
===========================  ==============================================  ==========================================
"""

import itertools

import numpy as np

SQRT2 = np.sqrt(2.0)


def sym_of(dim):
    return (dim * (dim + 1)) // 2


# --------------------------------------------------------------------------
# Materials
# --------------------------------------------------------------------------

def bulk_modulus(g, nu, dim):
    """Bulk modulus, plane strain in 2D (isotropic.pyx:79-97)."""
    if dim == 2:
        return g / (1. - 2. * nu)
    if dim == 3:
        return 2. / 3. * (1. + nu) / (1. - 2. * nu) * g
    raise ValueError('dim must be 2 or 3 (was {0})'.format(dim))


def delta_c_coefficients(g, nu, g0, nu0, dim):
    """(sph, dev) of C - C0 as built by square_basic.py:26-29."""
    k = bulk_modulus(g, nu, dim)
    k0 = bulk_modulus(g0, nu0, dim)
    return dim * (k - k0), 2. * (g - g0)


# --------------------------------------------------------------------------
# Continuous Green operator for strains
# --------------------------------------------------------------------------

def gamma0(k, g, nu):
    """Mandel matrix of Gamma0(k), shape (..., sym, sym); 0 at k = 0.

    Follows isotropic.pyx:133-156 (auxiliary constants), :173-200 (2D) and
    :242-308 (3D), with the same order of floating-point operations.
    """
    k = np.asarray(k, dtype=np.float64)
    dim = k.shape[-1]
    daux1 = 1.0 / g
    daux2 = 0.5 / (g * (1.0 - nu))
    daux3 = 0.25 / g
    daux4 = 0.5 / g
    sym = sym_of(dim)
    out = np.zeros(k.shape[:-1] + (sym, sym), dtype=np.float64)
    if dim == 2:
        k0 = k[..., 0]
        k1 = k[..., 1]
        k0k0 = k0 * k0
        k1k1 = k1 * k1
        kk = k0k0 + k1k1
        zero = kk == 0.
        with np.errstate(divide='ignore', invalid='ignore'):
            s = 1.0 / np.where(zero, 1.0, kk)
        k0k0 = k0k0 * s
        k1k1 = k1k1 * s
        k0k1 = s * k0 * k1
        g00 = k0k0 * (daux1 - daux2 * k0k0)
        g11 = k1k1 * (daux1 - daux2 * k1k1)
        dummy = -daux2 * k0k0 * k1k1
        g01 = dummy
        g22 = 2 * (daux3 + dummy)
        g02 = SQRT2 * k0k1 * (daux4 - daux2 * k0k0)
        g12 = SQRT2 * k0k1 * (daux4 - daux2 * k1k1)
        coeffs = {(0, 0): g00, (0, 1): g01, (0, 2): g02,
                  (1, 1): g11, (1, 2): g12, (2, 2): g22}
    elif dim == 3:
        k0 = k[..., 0]
        k1 = k[..., 1]
        k2 = k[..., 2]
        k0k0 = k0 * k0
        k1k1 = k1 * k1
        k2k2 = k2 * k2
        kk = k0k0 + k1k1 + k2k2
        zero = kk == 0.
        with np.errstate(divide='ignore', invalid='ignore'):
            s = 1.0 / np.where(zero, 1.0, kk)
        k0k0 = k0k0 * s
        k1k1 = k1k1 * s
        k2k2 = k2k2 * s
        k1k2 = s * k1 * k2
        k2k0 = s * k2 * k0
        k0k1 = s * k0 * k1
        coeffs = {
            (0, 0): k0k0 * (daux1 - daux2 * k0k0),
            (1, 1): k1k1 * (daux1 - daux2 * k1k1),
            (2, 2): k2k2 * (daux1 - daux2 * k2k2),
            (3, 3): 2. * (daux3 * (k1k1 + k2k2) - daux2 * k1k1 * k2k2),
            (4, 4): 2. * (daux3 * (k2k2 + k0k0) - daux2 * k2k2 * k0k0),
            (5, 5): 2. * (daux3 * (k0k0 + k1k1) - daux2 * k0k0 * k1k1),
            (0, 1): -daux2 * k0k0 * k1k1,
            (0, 2): -daux2 * k0k0 * k2k2,
            (0, 3): -SQRT2 * daux2 * k0k0 * k1k2,
            (0, 4): SQRT2 * k2k0 * (daux4 - daux2 * k0k0),
            (0, 5): SQRT2 * k0k1 * (daux4 - daux2 * k0k0),
            (1, 2): -daux2 * k1k1 * k2k2,
            (1, 3): SQRT2 * k1k2 * (daux4 - daux2 * k1k1),
            (1, 4): -SQRT2 * daux2 * k1k1 * k2k0,
            (1, 5): SQRT2 * k0k1 * (daux4 - daux2 * k1k1),
            (2, 3): SQRT2 * k1k2 * (daux4 - daux2 * k2k2),
            (2, 4): SQRT2 * k2k0 * (daux4 - daux2 * k2k2),
            (2, 5): -SQRT2 * daux2 * k2k2 * k0k1,
            (3, 4): 2 * k0k1 * (daux3 - daux2 * k2k2),
            (3, 5): 2 * k2k0 * (daux3 - daux2 * k1k1),
            (4, 5): 2 * k1k2 * (daux3 - daux2 * k0k0),
        }
    else:
        raise ValueError('dim must be 2 or 3 (was {0})'.format(dim))
    for (i, j), c in coeffs.items():
        c = np.where(zero, 0., c)
        out[..., i, j] = c
        out[..., j, i] = c
    return out


# --------------------------------------------------------------------------
# Discrete Green operators (per frequency)
# --------------------------------------------------------------------------

def _steps(n, h):
    """s_i = 2 pi / (h n_i) (green.pyx:140-141, 305-307)."""
    return [2. * np.pi / (h * ni) for ni in n]


def truncated_wavevector(b, n, h):
    """Aliased wave vector of the truncated operator (green.pyx:432-445).

    k_i = s_i (b_i - n_i) if 2 b_i > n_i else s_i b_i.  ``b`` may be an
    integer array of shape (..., dim).
    """
    b = np.asarray(b)
    s = _steps(n, h)
    k = np.empty(b.shape, dtype=np.float64)
    for i, ni in enumerate(n):
        bi = b[..., i]
        k[..., i] = np.where(2 * bi > ni, s[i] * (bi - ni), s[i] * bi)
    return k


def gamma_hat_truncated(b, n, h, g, nu):
    """Gamma-hat(b) of the truncated operator (green.pyx:432-451)."""
    return gamma0(truncated_wavevector(b, n, h), g, nu)


def _filtered_axis_tables(n, h):
    """Per-axis aliases and weights of the filtered operator.

    For axis i and index b: the two aliases k- = s_i (b - n_i), k+ = s_i b
    and their weights cos^2(h k / 4) (green.pyx:740-800).
    """
    s = _steps(n, h)
    tabs = []
    for i, ni in enumerate(n):
        b = np.arange(ni)
        km = s[i] * (b - ni)
        kp = s[i] * b
        wm = np.cos(0.25 * h * km)
        wm = wm * wm
        wp = np.cos(0.25 * h * kp)
        wp = wp * wp
        tabs.append((km, kp, wm, wp))
    return tabs


def gamma_hat_filtered(b, n, h, g, nu):
    """Gamma-hat(b) of the filtered operator (green.pyx:607-694, 730-910).

    Weighted sum over the 2^dim alias combinations, accumulated in janus's
    order: the first axis varies slowest and the "minus" alias comes first
    (k1 = (-,-,-), k2 = (-,-,+), ..., k8 = (+,+,+)).
    """
    b = np.asarray(b)
    dim = len(n)
    tabs = _filtered_axis_tables(n, h)
    acc = None
    for combo in itertools.product((0, 1), repeat=dim):
        k = np.empty(b.shape, dtype=np.float64)
        w = None
        for i in range(dim):
            km, kp, wm, wp = tabs[i]
            bi = b[..., i]
            if combo[i] == 0:
                k[..., i] = km[bi]
                wi = wm[bi]
            else:
                k[..., i] = kp[bi]
                wi = wp[bi]
            w = wi if w is None else w * wi
        term = w[..., None, None] * gamma0(k, g, nu)
        acc = term if acc is None else acc + term
    return acc


def willot_wavevector(b, n, h):
    """Willot's modified wave vector (green.pyx:984-1023):
    k_i = sin(phi_i) prod_{j != i} cos(phi_j), phi_i = s_i h b_i / 2, for
    the full-grid index b (no aliasing)."""
    b = np.asarray(b, dtype=np.float64)
    s = np.asarray(_steps(n, h))
    phi = 0.5 * s * h * b
    sn, cs = np.sin(phi), np.cos(phi)
    dim = len(n)
    k = np.empty(b.shape)
    for i in range(dim):
        k[..., i] = sn[..., i]
        for j in range(dim):
            if j != i:
                k[..., i] = k[..., i] * cs[..., j]
    return k


def gamma_hat_willot(b, n, h, g, nu):
    """Gamma-hat(b) of janus's willot2015 operator (green.pyx:976-1036)."""
    return gamma0(willot_wavevector(b, n, h), g, nu)


def gamma_hat(kind, b, n, h, g, nu):
    if kind == 'truncated':
        return gamma_hat_truncated(b, n, h, g, nu)
    if kind == 'filtered':
        return gamma_hat_filtered(b, n, h, g, nu)
    if kind == 'willot':
        return gamma_hat_willot(b, n, h, g, nu)
    raise ValueError(kind)


# --------------------------------------------------------------------------
# FFT (janus layout) and Green apply
# --------------------------------------------------------------------------

def rfft_janus(r):
    """r2c in janus's interleaved layout (_serial_fft.pyx:152-163).

    Output shape (..., 2 * (n_last // 2 + 1)); even index = Re, odd = Im
    (sphinx/fft_tutorial.rst:43).
    """
    c = np.fft.rfftn(r)
    out = np.empty(c.shape[:-1] + (2 * c.shape[-1],), dtype=np.float64)
    out[..., 0::2] = c.real
    out[..., 1::2] = c.imag
    return out


def irfft_janus(c, shape):
    """c2r from janus's interleaved layout, scaled by 1/N (:165-177)."""
    z = c[..., 0::2] + 1j * c[..., 1::2]
    return np.fft.irfftn(z, s=shape, axes=tuple(range(len(shape))))


def _half_indices(n, i0, i1=None):
    """Integer frequency indices of the half spectrum for axis-0 slab i0
    (or, given i1, for the slabs i0 ... i1 - 1, stacked on a first axis)."""
    dim = len(n)
    first = np.arange(i0, i0 + 1 if i1 is None else i1)
    axes = [first] + [np.arange(ni) for ni in n[1:-1]] \
        + [np.arange(n[-1] // 2 + 1)]
    grids = np.meshgrid(*axes, indexing='ij')
    b = np.stack(grids, axis=-1)
    if i1 is not None:
        return b
    return b.reshape(b.shape[1:]) if dim > 1 else b


def green_apply(x, kind, g, nu, h=1.0, workers=1):
    """y = Gamma0[x] on the grid (green.pyx:182-216, 353-390, 453-580).

    r2c of each component, per-frequency symmetric matvec applied to Re
    and Im, c2r of each component (c2r drops the imaginary parts of the
    DC and Nyquist planes of the last axis, as FFTW and pocketfft do).

    workers > 1 spreads the same arithmetic over host threads: scipy's
    pocketfft with that many workers, and blocks of axis-0 slabs of the
    frequency loop in a thread pool (numpy releases the GIL inside its
    array loops, which blocks of slabs make long).  janus itself is
    single-threaded; this form only times the path on every host core.
    """
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[:-1]
    dim = len(n)
    axes = tuple(range(dim))
    if workers > 1:
        import scipy.fft
        xh = scipy.fft.rfftn(x, axes=axes, workers=workers)
    else:
        xh = np.fft.rfftn(x, axes=axes)
    yh = np.empty_like(xh)

    def slab(i0):
        b = _half_indices(n, i0)
        G = gamma_hat(kind, b, n, h, g, nu)
        yh[i0] = np.einsum('...ij,...j->...i', G, xh[i0].real) \
            + 1j * np.einsum('...ij,...j->...i', G, xh[i0].imag)

    if workers > 1:
        from concurrent.futures import ThreadPoolExecutor
        step = max(1, -(-n[0] // (4 * workers)))

        def block(a):
            e = min(a + step, n[0])
            G = gamma_hat(kind, _half_indices(n, a, e), n, h, g, nu)
            yh[a:e] = np.einsum('...ij,...j->...i', G, xh[a:e].real) \
                + 1j * np.einsum('...ij,...j->...i', G, xh[a:e].imag)

        with ThreadPoolExecutor(workers) as pool:
            list(pool.map(block, range(0, n[0], step)))
        return scipy.fft.irfftn(yh, s=n, axes=axes, workers=workers)
    for i0 in range(n[0]):
        slab(i0)
    return np.fft.irfftn(yh, s=n, axes=axes)


# --------------------------------------------------------------------------
# Local operators
# --------------------------------------------------------------------------

def iso4_apply(sph, dev, x):
    """(sph J + dev K) : x in Mandel notation (operators.pyx:292-299,
    315-324, 352-365).  ``sph``/``dev`` broadcast against x[..., 0]."""
    x = np.asarray(x, dtype=np.float64)
    sym = x.shape[-1]
    dim = 2 if sym == 3 else 3
    sph = np.asarray(sph, dtype=np.float64)
    dev = np.asarray(dev, dtype=np.float64)
    tr = (sph - dev) / float(dim)
    if dim == 2:
        aux = tr * (x[..., 0] + x[..., 1])
    else:
        aux = tr * (x[..., 0] + x[..., 1] + x[..., 2])
    y = np.empty_like(x)
    for i in range(dim):
        y[..., i] = aux + dev * x[..., i]
    for i in range(dim, sym):
        y[..., i] = dev * x[..., i]
    return y


def iso4_matrix(sph, dev, dim):
    """Matrix of sph J + dev K (operators.pyx:326-340, 367-408)."""
    sym = sym_of(dim)
    a = np.zeros((sym, sym))
    a[:dim, :dim] = (sph - dev) / dim
    for i in range(dim):
        a[i, i] = (sph + (dim - 1) * dev) / dim
    for i in range(dim, sym):
        a[i, i] = dev
    return a


def cubic_tensor_2d(t1111, t1122, t1212, theta):
    """Matrix of the 2D cubic tensor rotated by theta, from the closed form
    in cos 4 theta, sin 4 theta (operators.pyx:413-426, 440-450)."""
    c = np.cos(4. * theta)
    s = np.sin(4. * theta)
    aux = t1111 - t1122 - 2. * t1212
    t11 = 0.25 * (3. * t1111 + t1122 + 2. * t1212 + c * aux)
    t12 = 0.25 * (t1111 + 3. * t1122 - 2. * t1212 - c * aux)
    t13 = 0.25 * SQRT2 * s * aux
    t23 = -t13
    t33 = 0.5 * (t1111 - t1122 + 2. * t1212 - c * aux)
    return np.array([[t11, t12, t13], [t12, t11, t23], [t13, t23, t33]])


def matrix_apply(a, x, transpose=False):
    """Block-diagonal linear operator (operators.pyx:739-775, 824-861):
    y[v]_i = sum_j a[v]_ij x[v]_j, or sum_j a[v]_ji x[v]_j transposed.
    ``a`` has shape grid + (m, n) (or (m, n), broadcast over the grid)."""
    a = np.asarray(a, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    if transpose:
        return np.einsum('...ji,...j->...i', a, x)
    return np.einsum('...ij,...j->...i', a, x)


def _local(sph, dev, mats):
    """The local operator x -> dC : x of a scheme: isotropic per-voxel
    (sph, dev), or per-voxel matrices ``mats`` (grid + (sym, sym))."""
    if mats is not None:
        return lambda x: matrix_apply(mats, x)
    return lambda x: iso4_apply(sph, dev, x)


def _grid_shape(sph, mats):
    return np.shape(sph) if mats is None else np.shape(mats)[:-2]


def block_diag_apply(phase, table, x):
    """Block-diagonal isotropic operator from a phase map (operators.pyx
    :655-664, 706-716): y[v] = iso4(table[phase[v]]) : x[v]."""
    table = np.asarray(table, dtype=np.float64)
    sph = table[phase, 0]
    dev = table[phase, 1]
    return iso4_apply(sph, dev, x)


# --------------------------------------------------------------------------
# Basic fixed-point scheme (square_basic.py:83-97)
# --------------------------------------------------------------------------

def basic_scheme(sph, dev, E, kind, g0, nu0, tol=0.0, max_iter=400, h=1.0,
                 rhs=None, eps0=None, mats=None):
    """Moulinec-Suquet iteration eps <- R - Gamma0[dC : eps].

    ``sph``/``dev`` are per-voxel fields (shape = grid) of dC = C - C0, or
    ``mats`` holds dC as per-voxel matrices (grid + (sym, sym); then sph
    and dev are ignored).  R = E (macroscopic strain, constant) plus the optional field ``rhs``.
    The residual is ||eps_{k+1} - eps_k|| / (sqrt(N) ||E||)
    (square_basic.py:86, 91); when ||E|| = 0 the plain norm / sqrt(N) is used.
    Iteration stops after the first iteration whose residual is < tol, or
    after ``max_iter`` iterations.  Returns (eps, residual history).
    """
    E = np.asarray(E, dtype=np.float64)
    shape = _grid_shape(sph, mats)
    local = _local(sph, dev, mats)
    N = int(np.prod(shape))
    nE = np.linalg.norm(E)
    normalization = 1. / np.sqrt(N) / (nE if nE > 0 else 1.0)
    eps = np.empty(shape + (E.size,), dtype=np.float64)
    eps[...] = E if eps0 is None else eps0
    res = []
    for _ in range(max_iter):
        tau = local(eps)
        new = green_apply(tau, kind, g0, nu0, h)
        new = E - new
        if rhs is not None:
            new += rhs
        r = normalization * np.linalg.norm(new - eps)
        res.append(r)
        eps = new
        if r < tol:
            break
    return eps, np.array(res)


def mean_stress(sph, dev, eps, g0, nu0, E, mats=None):
    """<sigma> = C0 : E + <dC : eps> (square_basic.rst:212-221)."""
    dim = 2 if eps.shape[-1] == 3 else 3
    k0 = bulk_modulus(g0, nu0, dim)
    c0 = iso4_matrix(dim * k0, 2 * g0, dim)
    tau = _local(sph, dev, mats)(eps)
    return c0 @ np.asarray(E) + tau.reshape(-1, eps.shape[-1]).mean(axis=0)


# --------------------------------------------------------------------------
# Adjoint (no janus counterpart): implicit differentiation of the fixed point
# --------------------------------------------------------------------------

def adjoint_gradient(sph, dev, eps, grad_eps, kind, g0, nu0, tol, max_iter,
                     h=1.0):
    """Gradient of L(eps(dC)) w.r.t. the per-voxel (sph, dev) of dC.

    eps solves (I + G dC) eps = E with G = Gamma0 (symmetric, as is dC).
    With g = dL/deps, mu solves (I + G dC) mu = G g, i.e. the fixed point
    mu <- G g - G[dC : mu]; then dL/dsph_x = -(J eps_x) . mu_x and
    dL/ddev_x = -(K eps_x) . mu_x.  Returns (dsph, ddev, mu, residuals).
    """
    sym = eps.shape[-1]
    dim = 2 if sym == 3 else 3
    Gg = green_apply(grad_eps, kind, g0, nu0, h)
    mu, res = _field_rhs_scheme(sph, dev, Gg, kind, g0, nu0, tol, max_iter, h)
    tr_e = eps[..., :dim].sum(axis=-1)
    tr_m = mu[..., :dim].sum(axis=-1)
    j_term = tr_e * tr_m / dim
    dsph = -j_term
    ddev = -((eps * mu).sum(axis=-1) - j_term)
    return dsph, ddev, mu, res


def effective_stress_gradient(eps, eps_g):
    """Gradient of gbar . <sigma(E)> w.r.t. the per-voxel (sph, dev) of dC,
    from the forward fields eps = eps(E) and eps_g = eps(gbar) (the solve
    under the load gbar), with <sigma> = C0 E + <dC : eps>.

    With A = I + G dC (G = Gamma0 and dC symmetric), eps(a) = A^-1 a for a
    uniform a, and the adjoint state of gbar . <sigma> is
    mu = -(eps_g - gbar) / N: (I + G dC) mu = G dC gbar / N is the equation
    of eps_g - gbar scaled by -1 / N.  adjoint_gradient's contraction plus
    the explicit dependence of <dC : eps> then give
    d/dsph_x = (J eps_x) . eps_g_x / N and d/ddev_x = (K eps_x) . eps_g_x / N.
    Returns (dsph, ddev).
    """
    sym = eps.shape[-1]
    dim = 2 if sym == 3 else 3
    N = eps.size // sym
    tr_e = eps[..., :dim].sum(axis=-1)
    tr_g = eps_g[..., :dim].sum(axis=-1)
    j_term = tr_e * tr_g / dim
    return j_term / N, ((eps * eps_g).sum(axis=-1) - j_term) / N


def _field_rhs_scheme(sph, dev, rhs, kind, g0, nu0, tol, max_iter, h):
    """mu <- rhs - Gamma0[dC : mu], residual ||mu_{k+1}-mu_k||/sqrt(N)/||rhs||_rms."""
    shape = np.shape(sph)
    N = int(np.prod(shape))
    nr = np.linalg.norm(rhs) / np.sqrt(N)
    normalization = 1. / np.sqrt(N) / (nr if nr > 0 else 1.0)
    mu = rhs.copy()
    res = []
    for _ in range(max_iter):
        new = rhs - green_apply(iso4_apply(sph, dev, mu), kind, g0, nu0, h)
        r = normalization * np.linalg.norm(new - mu)
        res.append(r)
        mu = new
        if r < tol:
            break
    return mu, np.array(res)


# --------------------------------------------------------------------------
# Conjugate gradients (janus drives Krylov solvers from its examples:
# examples/square_inclusion.py:40-47, 71-86 runs PETSc CG on the polarization
# form; this is CG on the strain form, which also holds for C0 = C_matrix)
# --------------------------------------------------------------------------

def c0_inner(a, b, g0, nu0):
    """Sum over voxels of a : C0 : b in Mandel notation, C0 = dim k0 J +
    2 g0 K: 2 g0 a.b + (k0 - 2 g0 / dim) tr(a) tr(b)."""
    dim = 2 if a.shape[-1] == 3 else 3
    k0 = bulk_modulus(g0, nu0, dim)
    tra = a[..., :dim].sum(axis=-1)
    trb = b[..., :dim].sum(axis=-1)
    return 2. * g0 * np.sum(a * b) + (k0 - 2. * g0 / dim) * np.sum(tra * trb)


def cg_scheme(sph, dev, E, kind, g0, nu0, tol=0.0, max_iter=400, h=1.0,
              directions=None, rhs=None, res_scale=None, mats=None):
    """Conjugate gradients on (I + Gamma0 dC) eps~ = -Gamma0[dC : E] for the
    fluctuation eps~ (eps = E + eps~), in the inner product
    <a, b> = sum a : C0 : b.

    Gamma0 C0 is the C0-orthogonal projector onto compatible zero-mean
    fields, so I + Gamma0 dC is self-adjoint there and positive definite
    whenever C = C0 + dC is, whatever the isotropic C0 (the basic scheme
    needs C0 near the middle of the phases).  The residual after iteration k
    is ||r_k|| / (sqrt(N) ||E||) with r_k = b - A eps~_k: for the basic
    scheme this quantity is its own residual ||eps_{k+1} - eps_k|| (same
    normalization, square_basic.py:86, 91).  Stops after the first
    iteration whose residual is < tol.  ``directions``: optional list
    receiving the search directions (for tests).  ``rhs``: optional field
    added to E, as in basic_scheme: eps solves (I + Gamma0 dC) eps = E + rhs
    (rhs compatible and zero-mean, e.g. a Gamma0 output); ``res_scale``
    replaces 1 / (sqrt(N) ||E||).  ``mats``: dC as symmetric per-voxel
    matrices, as in basic_scheme.  Returns (eps, residuals).
    """
    E = np.asarray(E, dtype=np.float64)
    shape = _grid_shape(sph, mats)
    local = _local(sph, dev, mats)
    N = int(np.prod(shape))
    nE = np.linalg.norm(E)
    normalization = 1. / np.sqrt(N) / (nE if nE > 0 else 1.0)
    if res_scale is not None:
        normalization = res_scale

    def op(v):
        return v + green_apply(local(v), kind, g0, nu0, h)

    Ef = np.empty(shape + (E.size,), dtype=np.float64)
    Ef[...] = E
    r = -green_apply(local(Ef), kind, g0, nu0, h)
    if rhs is not None:
        r = rhs + r
    x = np.zeros_like(r)
    p = r.copy()
    rr = c0_inner(r, r, g0, nu0)
    res = []
    if rr == 0.:
        return Ef, np.zeros(1)
    for _ in range(max_iter):
        if directions is not None:
            directions.append(p.copy())
        q = op(p)
        alpha = rr / c0_inner(p, q, g0, nu0)
        x = x + alpha * p
        r = r - alpha * q
        res_k = normalization * np.linalg.norm(r)
        res.append(res_k)
        if res_k < tol:
            break
        rr_new = c0_inner(r, r, g0, nu0)
        beta = rr_new / rr
        p = r + beta * p
        rr = rr_new
    return Ef + x, np.array(res)


def cg_polarization_scheme(sph, dev, E, kind, g0, nu0, tol=0.0, max_iter=400,
                           h=1.0, res_scale=None, mats=None, dot=None):
    """Conjugate gradients on the polarization form, as
    examples/square_inclusion.py:17-52, 77-85 runs PETSc CG (KSP 'cg', PC
    'none', rtol): (L + Gamma0) tau = E, where the local operator L (sph,
    dev or ``mats``) is the inverse of C - C0 (the example's tau2eps), in
    the Euclidean inner product from tau = 0.  The residual after
    iteration k is ||E - A tau_k|| / (sqrt(N) ||E||) (PETSc's relative
    residual; ``res_scale`` replaces the normalization).  Stops after the
    first iteration below tol.  ``dot(a, b)`` replaces the Euclidean dot
    product (default: numpy's pairwise sum of a * b); other summation orders
    give the rounding spread of the scheme.  Returns (tau, eps = L tau,
    residuals)."""
    E = np.asarray(E, dtype=np.float64)
    shape = _grid_shape(sph, mats)
    local = _local(sph, dev, mats)
    N = int(np.prod(shape))
    nE = np.linalg.norm(E)
    normalization = 1. / np.sqrt(N) / (nE if nE > 0 else 1.0)
    if res_scale is not None:
        normalization = res_scale

    def op(v):
        return local(v) + green_apply(v, kind, g0, nu0, h)

    if dot is None:
        def dot(a, b):
            return np.sum(a * b)

    r = np.empty(shape + (E.size,), dtype=np.float64)
    r[...] = E
    x = np.zeros_like(r)
    p = r.copy()
    rr = dot(r, r)
    res = []
    if rr == 0.:
        return x, local(x), np.zeros(0)
    for _ in range(max_iter):
        q = op(p)
        alpha = rr / dot(p, q)
        x = x + alpha * p
        r = r - alpha * q
        res_k = normalization * np.linalg.norm(r)
        res.append(res_k)
        if res_k < tol:
            break
        rr_new = dot(r, r)
        beta = rr_new / rr
        p = r + beta * p
        rr = rr_new
    return x, local(x), np.array(res)


def gmres_scheme(sph, dev, E, kind, g0, nu0, tol=0.0, max_iter=400,
                 restart=30, h=1.0, rhs=None, res_scale=None, mats=None):
    """Restarted GMRES on (I + Gamma0 dC) eps~ = -Gamma0[dC : E] (+ rhs) for
    the fluctuation eps~, the strain form that
    examples/square_inclusion_scipy.py:52-59, 88 hands to scipy's GMRES.

    Euclidean inner product; Arnoldi with one classical Gram-Schmidt
    projection per step; Givens rotations on the Hessenberg matrix.  The
    residual after iteration k is the Arnoldi estimate |g_{k+1}| of
    ||b - A eps~_k|| times 1 / (sqrt(N) ||E||) (``res_scale`` replaces it),
    the normalization of the basic scheme's residual (square_basic.py:86,
    91).  Stops after the first iteration below tol.  After ``restart``
    steps the iterate is updated and the residual b - A eps~ recomputed.
    No symmetry is needed: ``mats`` may hold non-symmetric per-voxel
    matrices.  Returns (eps, residuals); no iteration when b = 0.
    """
    E = np.asarray(E, dtype=np.float64)
    shape = _grid_shape(sph, mats)
    local = _local(sph, dev, mats)
    N = int(np.prod(shape))
    nE = np.linalg.norm(E)
    normalization = 1. / np.sqrt(N) / (nE if nE > 0 else 1.0)
    if res_scale is not None:
        normalization = res_scale

    def op(v):
        return v + green_apply(local(v), kind, g0, nu0, h)

    Ef = np.empty(shape + (E.size,), dtype=np.float64)
    Ef[...] = E
    b = -green_apply(local(Ef), kind, g0, nu0, h)
    if rhs is not None:
        b = rhs + b
    x = np.zeros_like(b)
    r = b
    res = []
    while len(res) < max_iter:
        beta = np.linalg.norm(r)
        if beta == 0.:
            break
        V = [r / beta]
        H = np.zeros((restart + 1, restart))
        cs = np.zeros(restart)
        sn = np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k = 0
        for j in range(restart):
            w = op(V[j])
            hj = np.array([np.sum(w * V[i]) for i in range(j + 1)])
            for i in range(j + 1):
                w = w - hj[i] * V[i]
            hn = np.linalg.norm(w)
            H[:j + 1, j] = hj
            H[j + 1, j] = hn
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            d = np.hypot(H[j, j], H[j + 1, j])
            cs[j] = H[j, j] / d
            sn[j] = H[j + 1, j] / d
            H[j, j] = d
            H[j + 1, j] = 0.
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            k = j + 1
            res.append(normalization * abs(g[j + 1]))
            if res[-1] < tol or len(res) >= max_iter or hn == 0.:
                break
            V.append(w / hn)
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):
            y[i] = (g[i] - np.dot(H[i, i + 1:k], y[i + 1:k])) / H[i, i]
        for i in range(k):
            x = x + y[i] * V[i]
        if res[-1] < tol or len(res) >= max_iter:
            break
        r = b - op(x)
    return Ef + x, np.array(res)
