"""The C restatement of janus's serial path (oracle/janus_serial.c, the timed
CPU baseline) against the pins of the numpy oracle:

* janus's 18 Green-operator fixtures (sampled in
  tests/golden/green_fixture_samples.npz) at janus's tolerance
  (tests/test_discrete_green_operator.py:199: rtol = ULP, atol = 10 ULP);
* the published tutorial result C_1212 = 1.41903971282
  (sphinx/tutorials/square_basic/square_basic.rst:227) after exactly 400
  basic-scheme iterations;
* the numpy oracle (itself pinned by both) on even, odd and mixed grids, 2D
  and 3D, both operators, h != 1, and iterate for iterate through the basic
  scheme;
* the multi-threaded form gives the single-threaded values bit for bit.
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import janus_oracle as jo
from test_oracle import check_against_fixture, fixture_cases, load_fixture

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, 'oracle', 'libjanus_serial.so')
KINDS = {'truncated': 0, 'filtered': 1}


@pytest.fixture(scope='module')
def lib():
    if not os.path.exists(LIB):
        subprocess.run(['make', '-C', os.path.join(ROOT, 'oracle')],
                       check=True, stdout=subprocess.DEVNULL)
    return ctypes.CDLL(LIB)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def c_apply(lib, x, kind, g0, nu0, h=1.0, nthreads=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    dim = x.ndim - 1
    shape = (ctypes.c_int64 * dim)(*x.shape[:-1])
    y = np.empty_like(x)
    rc = lib.js_green_apply(dim, shape, KINDS[kind], ctypes.c_double(g0),
                            ctypes.c_double(nu0), ctypes.c_double(h), _p(x),
                            _p(y), nthreads)
    assert rc == 0, rc
    return y


def c_basic(lib, phase, table, E, kind, g0, nu0, iters, nthreads=1):
    phase = np.ascontiguousarray(phase, dtype=np.uint8)
    table = np.ascontiguousarray(table, dtype=np.float64)
    E = np.ascontiguousarray(E, dtype=np.float64)
    dim = phase.ndim
    shape = (ctypes.c_int64 * dim)(*phase.shape)
    eps = np.empty(phase.shape + (E.size,))
    eps[...] = E
    res = np.zeros(iters)
    rc = lib.js_basic_iterations(dim, shape, KINDS[kind], ctypes.c_double(g0),
                                 ctypes.c_double(nu0), ctypes.c_double(1.0),
                                 _p(phase), _p(table), _p(E), _p(eps), iters,
                                 _p(res), nthreads)
    assert rc == 0, rc
    return eps, res


@pytest.mark.parametrize('name', fixture_cases())
def test_c_green_apply_fixtures(lib, name):
    kind, x, rec = load_fixture(name)
    check_against_fixture(c_apply(lib, x, kind, 0.75, 0.3, 1.0), rec)


@pytest.mark.parametrize('shape', [(8, 8), (9, 7), (30, 1), (12, 10, 14),
                                   (16, 16, 16), (5, 6, 7), (20, 18, 15),
                                   (32, 3, 40)])
@pytest.mark.parametrize('kind', ['truncated', 'filtered'])
def test_c_green_apply_vs_numpy_oracle(lib, shape, kind):
    rng = np.random.default_rng(sum(shape))
    sym = 3 if len(shape) == 2 else 6
    x = rng.standard_normal(shape + (sym,))
    y = c_apply(lib, x, kind, 1.3, 0.27, 0.7)
    r = jo.green_apply(x, kind, 1.3, 0.27, 0.7)
    assert np.abs(y - r).max() <= 2e-15 * np.abs(r).max()
    # in place
    xi = x.copy()
    shp = (ctypes.c_int64 * len(shape))(*shape)
    assert lib.js_green_apply(len(shape), shp, KINDS[kind],
                              ctypes.c_double(1.3), ctypes.c_double(0.27),
                              ctypes.c_double(0.7), _p(xi), _p(xi), 1) == 0
    assert np.array_equal(xi, y)


def test_c_tutorial_c1212(lib):
    """square_basic.py:60-97 with the published result."""
    n, dim = 256, 2
    sph_i, dev_i = jo.delta_c_coefficients(100., 0.2, 50., 0.3, dim)
    sph_m, dev_m = jo.delta_c_coefficients(1., 0.3, 50., 0.3, dim)
    phase = np.zeros((n, n), dtype=np.uint8)
    imax = int(np.ceil(n * 0.5 - 0.5))
    phase[:imax, :imax] = 1
    tab = np.array([[sph_m, dev_m], [sph_i, dev_i]])
    E = np.array([0., 0., 1.])
    eps, res = c_basic(lib, phase, tab, E, 'truncated', 50., 0.3, 400)
    tau = jo.block_diag_apply(phase, tab, eps)
    c1212 = 50. + 0.5 * tau.reshape(-1, 3).mean(axis=0)[-1] / E[-1]
    assert abs(c1212 - 1.41903971282) < 6e-12


@pytest.mark.parametrize('shape, kind', [((24, 20), 'filtered'),
                                         ((10, 12, 14), 'truncated'),
                                         ((16, 16, 16), 'filtered')])
def test_c_basic_scheme_iterates(lib, shape, kind):
    rng = np.random.default_rng(7)
    phase = (rng.random(shape) < 0.3).astype(np.uint8)
    dim = len(shape)
    sym = 3 if dim == 2 else 6
    g0, nu0 = 5.5, 0.3
    tab = np.array([jo.delta_c_coefficients(1.0, 0.3, g0, nu0, dim),
                    jo.delta_c_coefficients(10.0, 0.25, g0, nu0, dim)])
    E = rng.standard_normal(sym)
    eps, res = c_basic(lib, phase, tab, E, kind, g0, nu0, 12)
    eps_r, res_r = jo.basic_scheme(tab[phase, 0], tab[phase, 1], E, kind, g0,
                                   nu0, tol=0.0, max_iter=12)
    np.testing.assert_allclose(res, res_r, rtol=1e-12)
    assert np.abs(eps - eps_r).max() <= 1e-13 * np.abs(eps_r).max()


def test_c_threads_bitwise(lib):
    rng = np.random.default_rng(3)
    shape = (24, 20, 18)
    phase = (rng.random(shape) < 0.3).astype(np.uint8)
    tab = np.array([[-1.0, -2.0], [3.0, 4.0]])
    E = np.array([0., 0., 0., 0., 0., 1.])
    a, ra = c_basic(lib, phase, tab, E, 'filtered', 25.5, 0.3, 3, nthreads=1)
    b, rb = c_basic(lib, phase, tab, E, 'filtered', 25.5, 0.3, 3, nthreads=4)
    assert np.array_equal(a, b)
    # the residual reduction's order depends on the thread split
    np.testing.assert_allclose(ra, rb, rtol=1e-14)
