"""janus-compatible API checks that need no GPU: materials, the continuous
and discrete per-frequency Green operators (host code of libjanus_amd.so),
isotropic tensors, shape validation.  Mirrors janus's
tests/test_isotropic_linear_elastic_material.py, test_greenop.py,
test_discrete_green_operator.py (per-frequency part) and test_operators.py.
"""

import itertools

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd.green import filtered, truncated
from janus_amd.mandelvoigt import MandelVoigt
from janus_amd.operators import isotropic_4
from oracle import janus_oracle as jo

from test_oracle import GRID_SIZES, brute_force_gamma_hat, tensor_green, \
    wave_vectors

ULP = np.finfo(np.float64).eps
MU, NU = 0.75, 0.3


# ------------------------------------------------------------ materials

@pytest.mark.parametrize('g, nu, dim', [(2.0, 0.3, 2), (2.0, 0.3, 3)])
def test_poisson_from_bulk_and_shear_moduli(g, nu, dim):
    mat = material.create(g, nu, dim)
    assert material.poisson_from_bulk_and_shear_moduli(mat.k, g, dim) == nu


@pytest.mark.parametrize('g, k, nu, dim', [(1.5, 3.75, 0.3, 2),
                                           (1.5, 3.25, 0.3, 3)])
def test_material_init(g, k, nu, dim):
    mat = material.create(g, nu, dim)
    assert (mat.g, mat.k, mat.nu) == (g, k, nu)


@pytest.mark.parametrize('g, nu, dim', [(-1.0, 0.3, 2), (-1.0, 0.3, 3),
                                        (1.0, 0.5 + 1E-14, 2), (1.0, 1.0, 2),
                                        (1.0, -1.0 - 1E-14, 3), (1.0, -1.5, 3),
                                        (1.0, 0.5 + 1E-14, 3), (1.0, 1.0, 3),
                                        (1.0, 0.3, 0), (1.0, 0.3, 4)])
def test_material_invalid(g, nu, dim):
    with pytest.raises(ValueError):
        material.create(g, nu, dim)


@pytest.mark.parametrize('name, value', [('g', 2.0), ('k', 10.0),
                                         ('nu', 0.5), ('dim', 2)])
def test_material_read_only(name, value):
    mat = material.create(1.0, 0.3, 3)
    with pytest.raises(AttributeError):
        setattr(mat, name, value)


# ------------------------------------------------------------ Gamma0(k)

@pytest.mark.parametrize('dim', [2, 3])
@pytest.mark.parametrize('flag', [0, 1, 2])
def test_continuous_green_apply(dim, flag):
    mat = material.create(0.7, 0.3, dim)
    green = mat.green_operator()
    for k in wave_vectors(dim):
        expected = tensor_green(k, mat.g, mat.nu)
        green.set_frequency(k)
        tau = np.zeros(green.isize)
        for j in range(green.isize):
            tau[:] = 0.
            tau[j] = 1.
            base = {0: None, 1: tau, 2: np.empty(green.osize)}[flag]
            actual = green.apply(tau, base)
            if flag:
                assert actual.base is base
            np.testing.assert_array_almost_equal_nulp(expected[:, j], actual,
                                                      325)


@pytest.mark.parametrize('dim', [2, 3])
def test_continuous_green_to_memoryview(dim):
    mat = material.create(0.7, 0.3, dim)
    green = mat.green_operator()
    for k in wave_vectors(dim):
        green.set_frequency(k)
        np.testing.assert_array_almost_equal_nulp(
            tensor_green(k, mat.g, mat.nu), green.to_memoryview(), 325)
        base = np.empty((green.osize, green.isize))
        assert green.to_memoryview(base).base is base


def test_continuous_green_matches_oracle_bitwise():
    mat = material.create(1.0, 0.3, 3)
    green = mat.green_operator()
    rng = np.random.default_rng(3)
    for k in rng.standard_normal((50, 3)):
        green.set_frequency(k)
        np.testing.assert_array_equal(green.to_memoryview(),
                                      jo.gamma0(k, 1.0, 0.3))


def test_continuous_green_invalid():
    green = material.create(1.0, 0.3, 2).green_operator()
    with pytest.raises(ValueError):
        green.set_frequency(np.zeros(3))
    with pytest.raises(ValueError):
        material._GreenOperatorForStrains3D(material.create(1.0, 0.3, 2))


# ------------------------------------------------------------ Gamma-hat(b)

def multi_indices(n):
    return [np.asarray(b, dtype=np.intc)
            for b in itertools.product(*[range(ni) for ni in n])]


FACTORIES = {'truncated': truncated, 'filtered': filtered}


@pytest.mark.parametrize('kind', ['truncated', 'filtered'])
@pytest.mark.parametrize('n', GRID_SIZES)
@pytest.mark.parametrize('flag', [0, 1])
def test_to_memoryview(kind, n, flag):
    dim = len(n)
    greenc = material.create(MU, NU, dim).green_operator()
    greend = FACTORIES[kind](greenc, n, 1.0)
    for b in multi_indices(n):
        expected = brute_force_gamma_hat(kind, b, n, MU, NU)
        greend.set_frequency(b)
        if flag == 0:
            actual = greend.to_memoryview()
        else:
            base = np.empty_like(expected)
            actual = greend.to_memoryview(base)
            assert actual.base is base
        # janus asserts rtol = atol = ULP against its own continuous
        # operator; against the tensor closed form rounding differs slightly
        np.testing.assert_allclose(actual, expected, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize('kind', ['truncated', 'filtered'])
@pytest.mark.parametrize('n', GRID_SIZES)
def test_to_memoryview_matches_oracle(kind, n):
    greenc = material.create(MU, NU, len(n)).green_operator()
    greend = FACTORIES[kind](greenc, n, 1.0)
    bs = np.array(multi_indices(n))
    expected = jo.gamma_hat(kind, bs, n, 1.0, MU, NU)
    for b, e in zip(bs, expected):
        greend.set_frequency(b)
        np.testing.assert_allclose(greend.to_memoryview(), e, rtol=ULP,
                                   atol=ULP)


@pytest.mark.parametrize('kind', ['truncated', 'filtered'])
@pytest.mark.parametrize('n', GRID_SIZES)
@pytest.mark.parametrize('flag', [0, 1, 2])
def test_apply_by_freq(kind, n, flag):
    dim = len(n)
    x = np.array([0.3, -0.4, 0.5]) if dim == 2 else \
        np.array([0.1, -0.2, 0.3, -0.4, 0.5, -0.6])
    green = FACTORIES[kind](material.create(MU, NU, dim).green_operator(),
                            n, 1.)
    for b in multi_indices(n):
        green.set_frequency(b)
        expected = np.dot(np.asarray(green.to_memoryview()), x)
        xx = x.copy()
        base = {0: None, 1: xx, 2: np.empty_like(expected)}[flag]
        actual = green.apply_by_freq(xx, base)
        if flag:
            assert actual.base is base
        np.testing.assert_allclose(actual, expected, ULP, ULP)


def _invalid_init_params():
    g2 = material.create(MU, NU, 2).green_operator()
    g3 = material.create(MU, NU, 3).green_operator()
    return [(g3, (9, 9), 1., None), (g2, (-1, 9), 1., None),
            (g2, (9, -1), 1., None), (g2, (9, 9), -1., None),
            (g2, (9, 9), 1., fft.create_real((8, 9))),
            (g2, (9, 9), 1., fft.create_real((9, 8))),
            (g2, (9, 9, 9), 1., None), (g3, (-1, 9, 9), 1., None),
            (g3, (9, -1, 9), 1., None), (g3, (9, 9, -1), 1., None),
            (g3, (9, 9, 9), -1., None)]


@pytest.mark.parametrize('kind', ['truncated', 'filtered'])
@pytest.mark.parametrize('case', range(11))
def test_init_invalid_params(kind, case):
    greenc, n, h, transform = _invalid_init_params()[case]
    with pytest.raises(ValueError):
        FACTORIES[kind](greenc, n, h, transform)


@pytest.mark.parametrize('kind', ['truncated', 'filtered'])
def test_per_frequency_invalid_params(kind):
    g2 = FACTORIES[kind](material.create(MU, NU, 2).green_operator(),
                         (8, 16), 1.0)
    g3 = FACTORIES[kind](material.create(MU, NU, 3).green_operator(),
                         (8, 16, 32), 1.0)
    for g in (g2, g3):
        g.set_frequency(np.zeros(g.dim, dtype=np.intc))
        o, i = g.oshape[-1], g.ishape[-1]
        for shape in ((o, i + 1), (o + 1, i)):
            with pytest.raises(ValueError):
                g.to_memoryview(np.zeros(shape))
        for xs, ys in ((i, o + 1), (i + 1, o)):
            with pytest.raises(ValueError):
                g.apply_by_freq(np.zeros(xs), np.zeros(ys))
        for bad in ([-1] + [0] * (g.dim - 1), [g.ishape[0]] + [0] * (g.dim - 1),
                    [0] * (g.dim + 1)):
            with pytest.raises(ValueError):
                g.set_frequency(np.array(bad, dtype=np.intc))


def test_apply_invalid_params_shapes():
    g = truncated(material.create(MU, NU, 2).green_operator(), (8, 16), 1.0,
                  fft.create_real((8, 16)))
    i0, i1, i2 = g.ishape
    for xs, ys in [((i0 + 1, i1, i2), g.oshape), ((i0, i1 + 1, i2), g.oshape),
                   ((i0, i1, i2 + 1), g.oshape), (g.ishape, (i0 + 1, i1, i2)),
                   (g.ishape, (i0, i1 + 1, i2)), (g.ishape, (i0, i1, i2 + 1))]:
        with pytest.raises(ValueError):
            g.apply(np.zeros(xs), np.zeros(ys))


def test_operator_attributes():
    t = fft.create_real((8, 16, 32))
    g = filtered(material.create(MU, NU, 3).green_operator(), (8, 16, 32),
                 0.5, t)
    assert g.ishape == (8, 16, 32, 6) and g.oshape == (8, 16, 32, 6)
    assert (g.shape0, g.shape1, g.shape2, g.ishape3, g.oshape3) == \
        (8, 16, 32, 6, 6)
    assert g.h == 0.5 and g.transform is t and g.dim == 3
    assert t.oshape == (8, 16, 34) and t.global_ishape == (8, 16, 32)
    assert t.isize == 8 * 16 * 32 and t.osize == 8 * 16 * 34
    assert t.offset0 == 0 and t.idispl == 0 and t.odispl == 0


# ------------------------------------------------------------ FFT shapes

@pytest.mark.parametrize('shape, ishape, oshape', [
    ((128, 256), (127, 256), (128, 258)), ((128, 256), (128, 255), (128, 258)),
    ((128, 256), (128, 256), (129, 258)), ((128, 256), (128, 256), (128, 259)),
    ((128, 64, 32), (127, 64, 32), (128, 64, 34)),
    ((128, 64, 32), (128, 63, 32), (128, 64, 34)),
    ((128, 64, 32), (128, 64, 31), (128, 64, 34)),
    ((128, 64, 32), (128, 64, 32), (127, 64, 34)),
    ((128, 64, 32), (128, 64, 32), (128, 63, 34)),
    ((128, 64, 32), (128, 64, 32), (128, 64, 33))])
@pytest.mark.parametrize('direct', [True, False])
def test_fft_invalid_params(shape, ishape, oshape, direct):
    """tests/test_fft.py:19-31, 73-80: shape errors precede any GPU work."""
    r = np.empty(ishape)
    c = np.empty(oshape)
    with pytest.raises(ValueError):
        if direct:
            fft.create_real(shape).r2c(r, c)
        else:
            fft.create_real(shape).c2r(c, r)


def test_create_real_invalid_length():
    with pytest.raises(ValueError):
        fft.create_real((8,))
    with pytest.raises(ValueError):
        fft.create_real((8, 8, 8, 8))


# ------------------------------------------------------------ operators

@pytest.mark.parametrize('dim', [2, 3])
@pytest.mark.parametrize('sph, dev', [(1., 0.), (0., 1.), (2.5, -3.5)])
@pytest.mark.parametrize('flag', [0, 1, 2])
def test_isotropic_4(dim, sph, dev, flag):
    t = isotropic_4(sph, dev, dim)
    sym = t.isize
    expected = jo.iso4_matrix(sph, dev, dim)
    eye = np.eye(sym)
    for i in range(sym):
        x = eye[:, i].copy()
        base = {0: None, 1: x, 2: np.random.rand(sym)}[flag]
        ret = t.apply(x, base)
        if flag:
            assert ret.base is base
        np.testing.assert_allclose(ret, expected[:, i], 1 * ULP, 0 * ULP)
    np.testing.assert_allclose(t.to_memoryview(), expected, 0, 0)


def test_isotropic_4_invalid_dim():
    with pytest.raises(ValueError):
        isotropic_4(1., 1., 4)


def test_mandelvoigt():
    mv = MandelVoigt(3)
    assert mv.unravel_index(3) == (1, 2) and mv.ravel_multi_index(0, 1) == 5
    with pytest.raises(ValueError):
        mv.unravel_index(6)
    assert MandelVoigt(2).ravel_multi_index(1, 0) == 2


@pytest.mark.parametrize('n', [(8, 6), (7, 5), (6, 4, 5), (5, 8, 3)])
def test_willot_per_frequency_api(n):
    """to_memoryview / apply_by_freq of willot2015 (green.pyx:976-1036)
    against the oracle, whose wave vector is pinned by the finite-difference
    stencil's symbol (test_oracle.py)."""
    from janus_amd.green import willot2015
    from oracle import janus_oracle as jo
    dim = len(n)
    greend = willot2015(material.create(MU, NU, dim).green_operator(), n, 0.5)
    rng = np.random.default_rng(1)
    for b in multi_indices(n):
        expected = jo.gamma_hat('willot', np.array(b), n, 0.5, MU, NU)
        greend.set_frequency(b)
        np.testing.assert_allclose(greend.to_memoryview(), expected,
                                   rtol=1e-13, atol=1e-15)
        tau = rng.standard_normal(expected.shape[0])
        np.testing.assert_allclose(greend.apply_by_freq(tau), expected @ tau,
                                   rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize('shape', [(0, 8), (8, 0), (0, 4, 4), (4, 0, 4),
                                   (4, 4, 0)])
def test_empty_grids_are_rejected(shape):
    """A zero-length axis is a shape error (ValueError) for the transform
    and for both Green operator factories, as janus's negative sizes are."""
    with pytest.raises(ValueError):
        fft.create_real(shape)
    g = material.create(MU, NU, len(shape)).green_operator()
    for factory in (truncated, filtered):
        with pytest.raises(ValueError):
            factory(g, shape, 1.0)


def test_example_filtered_green_operator():
    """examples/example_filtered_green_operator.py: the class constructed
    directly (no transform), the per-frequency API over every frequency of
    the grid (the example takes 256 x 256; 32 x 48 here), eta[0] for
    tau = (0, 0, 1), against the oracle's filtered Gamma-hat."""
    import janus_amd.green as green
    mat = material.create(1.0, 0.3, 2)
    n0, n1 = 32, 48
    greend = green.FilteredGreenOperator2D(mat.green_operator(), (n0, n1), 1.)
    g = np.empty((n0, n1))
    b = np.empty((2,), dtype=np.intc)
    tau = np.array([0., 0., 1.])
    eta = np.empty_like(tau)
    for b0 in range(n0):
        b[0] = b0
        for b1 in range(n1):
            b[1] = b1
            greend.set_frequency(b)
            greend.apply_by_freq(tau, eta)
            g[b0, b1] = eta[0]
    bs = np.stack(np.meshgrid(np.arange(n0), np.arange(n1), indexing='ij'),
                  axis=-1).reshape(-1, 2)
    ref = jo.gamma_hat('filtered', bs, (n0, n1), 1.0, 1.0, 0.3)[:, 0, 2]
    np.testing.assert_allclose(g.ravel(), ref, rtol=ULP, atol=ULP)
