"""Pins of the oracle's anisotropic local operators (oracle/janus_oracle.py):

* cubic_tensor_2d against the rotation of the fourth-rank tensor written
  in index notation, T'_ijkl = R_ip R_jq R_kr R_ls T_pqrs, on the parameter
  grid of tests/test_operators.py:238-320 (moduli and angles);
* matrix_apply against explicit per-voxel loops (operators.pyx:739-775,
  824-861) with non-square blocks, and the transpose identity
  <y, A x> = <A^T y, x> of tests/test_operators.py:501-505;
* basic_scheme / cg_scheme with per-voxel matrices against the isotropic
  form on iso4_matrix blocks, and against the closed-form solution of a
  laminate of anisotropic layers (strain piecewise constant: the components
  with a normal index jump, the others equal E; traction continuous).
"""

import itertools

import numpy as np
import pytest

from oracle import janus_oracle as jo

ULP = np.finfo(np.float64).eps
COEFFS = [(1., 0., 0.), (0., 1., 0.), (0., 0., 1.), (2.5, -3.5, 4.5)]
ANGLES = [0, np.pi / 6, np.pi / 4, np.pi / 3, np.pi / 2, 0.37]


def rotated_cubic_mandel(t1111, t1122, t1212, theta):
    """Mandel matrix of the cubic tensor of moduli t1111, t1122, t1212 in
    axes rotated by theta, from the full tensor in index notation."""
    t = np.zeros((2, 2, 2, 2))
    for i, j in itertools.product(range(2), repeat=2):
        t[i, i, j, j] = t1111 if i == j else t1122
    for i, j in ((0, 1), (1, 0)):
        t[i, j, i, j] = t[i, j, j, i] = t1212
    c, s = np.cos(theta), np.sin(theta)
    r = np.array([[c, -s], [s, c]])
    tr = np.einsum('ip,jq,kr,ls,pqrs->ijkl', r, r, r, r, t)
    pairs = [(0, 0), (1, 1), (0, 1)]
    w = [1., 1., np.sqrt(2.)]
    m = np.empty((3, 3))
    for a, (i, j) in enumerate(pairs):
        for b, (k, l) in enumerate(pairs):
            m[a, b] = w[a] * w[b] * tr[i, j, k, l]
    return m


@pytest.mark.parametrize('coeffs', COEFFS)
@pytest.mark.parametrize('theta', ANGLES)
def test_cubic_tensor_is_the_rotated_tensor(coeffs, theta):
    expected = rotated_cubic_mandel(*coeffs, theta)
    actual = jo.cubic_tensor_2d(*coeffs, theta)
    scale = max(1.0, np.abs(expected).max())
    np.testing.assert_allclose(actual, expected, rtol=0, atol=10 * ULP * scale)


def test_cubic_tensor_isotropic_case():
    """t1111 - t1122 = 2 t1212 is isotropic: no dependence on theta, and
    the matrix is iso4_matrix(2 k, 2 mu) with k = (t1111 + t1122) / 2."""
    t1122, t1212 = 1.25, 0.75
    t1111 = t1122 + 2 * t1212
    iso = jo.iso4_matrix(t1111 + t1122, 2 * t1212, 2)
    for theta in ANGLES:
        np.testing.assert_allclose(jo.cubic_tensor_2d(t1111, t1122, t1212,
                                                      theta), iso,
                                   rtol=0, atol=10 * ULP * 4)


@pytest.mark.parametrize('shape, m, n', [((3, 4), 3, 3), ((2, 3, 2), 6, 6),
                                         ((4, 3), 2, 5), ((2, 2, 3), 4, 1)])
def test_matrix_apply_per_voxel_loops(shape, m, n):
    rng = np.random.default_rng(m * 10 + n)
    a = rng.uniform(-1, 1, size=shape + (m, n))
    x = rng.random(shape + (n,))
    xt = rng.random(shape + (m,))
    y = jo.matrix_apply(a, x)
    yt = jo.matrix_apply(a, xt, transpose=True)
    for index in itertools.product(*map(range, shape)):
        for i in range(m):
            acc = 0.
            for j in range(n):
                acc += a[index][i, j] * x[index][j]
            assert abs(y[index][i] - acc) <= 8 * ULP * max(1., abs(acc))
        for i in range(n):
            acc = 0.
            for j in range(m):
                acc += a[index][j, i] * xt[index][j]
            assert abs(yt[index][i] - acc) <= 8 * ULP * max(1., abs(acc))
    # <y, A x> = <A^T y, x>
    yax = np.sum(xt * y)
    atyx = np.sum(yt * x)
    assert abs(yax - atyx) <= 100 * ULP * max(abs(yax), 1.0)


def test_matrix_scheme_equals_isotropic_scheme():
    rng = np.random.default_rng(4)
    n = 12
    sph = rng.uniform(-0.5, 2., size=(n, n))
    dev = rng.uniform(-0.5, 2., size=(n, n))
    mats = np.empty((n, n, 3, 3))
    for index in itertools.product(range(n), range(n)):
        mats[index] = jo.iso4_matrix(sph[index], dev[index], 2)
    E = np.array([0.3, -0.2, 1.0])
    e1, r1 = jo.basic_scheme(sph, dev, E, 'filtered', 1.5, 0.3, tol=1e-12,
                             max_iter=300)
    e2, r2 = jo.basic_scheme(None, None, E, 'filtered', 1.5, 0.3, tol=1e-12,
                             max_iter=300, mats=mats)
    assert len(r1) == len(r2)
    np.testing.assert_allclose(r2, r1, rtol=1e-10, atol=1e-15)
    assert np.abs(e1 - e2).max() < 1e-13


# ---------------------------------------------------------------- laminates

def laminate_strains(stiff, frac, E, normal_idx, dim):
    """Per-layer strains of a laminate with normal e0 (exact solution):
    components in ``normal_idx`` (those with an index 1: 11, 12 in 2D;
    11, 31, 12 in 3D, Mandel) differ between layers, the others equal E;
    sigma . e0 is the same in every layer; <eps> = E."""
    sym = stiff[0].shape[0]
    N = list(normal_idx)
    T = [i for i in range(sym) if i not in N]
    inv = [np.linalg.inv(c[np.ix_(N, N)]) for c in stiff]
    A = sum(f * m for f, m in zip(frac, inv))
    b = E[N] + sum(f * m @ c[np.ix_(N, T)] @ E[T]
                   for f, m, c in zip(frac, inv, stiff))
    s = np.linalg.solve(A, b)
    out = []
    for m, c in zip(inv, stiff):
        e = E.copy()
        e[N] = m @ (s - c[np.ix_(N, T)] @ E[T])
        out.append(e)
    return out


def laminate_layers(n0):
    layer = np.zeros(n0, dtype=np.int64)
    layer[n0 // 4: n0 // 4 + 3 * n0 // 8] = 1
    layer[3 * n0 // 4:] = 2
    return layer


def spd(rng, sym, scale):
    a = rng.uniform(-1, 1, size=(sym, sym))
    return scale * (a @ a.T / sym + np.eye(sym))


@pytest.mark.parametrize('scheme', ['basic', 'cg'])
def test_laminate_2d_cubic_closed_form(scheme):
    n0, n1 = 16, 6
    g0, nu0 = 2.0, 0.3
    k0 = jo.bulk_modulus(g0, nu0, 2)
    c0 = jo.iso4_matrix(2 * k0, 2 * g0, 2)
    stiff = [jo.cubic_tensor_2d(4.0, 1.5, 1.0, 0.3),
             jo.cubic_tensor_2d(7.0, 2.0, 3.0, 1.1),
             jo.iso4_matrix(2 * 2.5, 2 * 1.6, 2)]
    layer = laminate_layers(n0)
    phase = np.repeat(layer[:, None], n1, axis=1)
    mats = np.array([c - c0 for c in stiff])[phase]
    E = np.array([0.4, -0.3, 0.8])
    fn = jo.basic_scheme if scheme == 'basic' else jo.cg_scheme
    eps, res = fn(None, None, E, 'truncated', g0, nu0, tol=1e-13,
                  max_iter=2000, mats=mats)
    assert res[-1] < 1e-13
    frac = [np.mean(layer == p) for p in range(3)]
    ref = laminate_strains(stiff, frac, E, [0, 2], 2)
    expected = np.array(ref)[phase]
    assert np.abs(eps - expected).max() < 1e-11


@pytest.mark.parametrize('scheme', ['basic', 'cg'])
def test_laminate_3d_general_anisotropic_closed_form(scheme):
    n0, n1, n2 = 16, 4, 3
    g0, nu0 = 3.0, 0.25
    k0 = jo.bulk_modulus(g0, nu0, 3)
    c0 = jo.iso4_matrix(3 * k0, 2 * g0, 3)
    rng = np.random.default_rng(8)
    stiff = [spd(rng, 6, s) for s in (2.0, 5.0, 3.0)]
    layer = laminate_layers(n0)
    phase = np.broadcast_to(layer[:, None, None], (n0, n1, n2))
    mats = np.array([c - c0 for c in stiff])[phase]
    E = rng.standard_normal(6)
    fn = jo.basic_scheme if scheme == 'basic' else jo.cg_scheme
    eps, res = fn(None, None, E, 'truncated', g0, nu0, tol=1e-13,
                  max_iter=3000, mats=mats)
    assert res[-1] < 1e-13
    frac = [np.mean(layer == p) for p in range(3)]
    ref = laminate_strains(stiff, frac, E, [0, 4, 5], 3)
    expected = np.array(ref)[phase]
    assert np.abs(eps - expected).max() < 1e-11
