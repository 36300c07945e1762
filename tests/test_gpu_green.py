"""GPU discrete Green operators vs janus's 18 fixtures (at janus's own
tolerance, tests/test_discrete_green_operator.py:178-199, with the three
calling sequences apply(x), apply(x, x), apply(x, y)) and vs the CPU oracle
on odd and mixed-radix grids."""

import zlib

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd.green import filtered, truncated, willot2015
from oracle import janus_oracle as jo

from test_oracle import check_against_fixture, fixture_cases, load_fixture

pytestmark = pytest.mark.gpu

MU, NU = 0.75, 0.3
FACTORIES = {'truncated': truncated, 'filtered': filtered,
             'willot': willot2015}


@pytest.mark.parametrize('name', fixture_cases())
@pytest.mark.parametrize('flag', [0, 1, 2])
def test_apply_fixtures(name, flag):
    kind, x, rec = load_fixture(name)
    n = x.shape[:-1]
    transform = fft.create_real(n)
    green = FACTORIES[kind](material.create(MU, NU, len(n)).green_operator(),
                            n, 1., transform=transform)
    if flag == 0:
        actual = green.apply(x)
    elif flag == 1:
        actual = green.apply(x, x)
        assert actual.base is x
    else:
        base = np.zeros(transform.ishape + (green.oshape[-1],))
        actual = green.apply(x, base)
        assert actual.base is base
    check_against_fixture(np.asarray(actual), rec)


@pytest.mark.parametrize('kind', ['truncated', 'filtered', 'willot'])
@pytest.mark.parametrize('n', [(7, 9), (12, 10), (16, 16), (1, 8), (9, 1),
                               (5, 6, 7), (9, 8, 10), (8, 8, 8), (3, 1, 4),
                               (16, 12, 20), (80, 96), (160, 80),
                               (12, 80, 96), (96, 10, 160), (24, 16, 80),
                               (1024, 12), (12, 1024), (1024, 8, 16),
                               (8, 1024, 16), (8, 16, 1024)])
@pytest.mark.parametrize('h', [1.0, 0.25])
def test_apply_vs_oracle(kind, n, h):
    rng = np.random.default_rng(zlib.crc32(repr((kind, n)).encode()))
    dim = len(n)
    sym = 3 if dim == 2 else 6
    x = rng.standard_normal(n + (sym,))
    green = FACTORIES[kind](material.create(MU, NU, dim).green_operator(), n,
                            h, transform=fft.create_real(n))
    actual = np.asarray(green.apply(x))
    expected = jo.green_apply(x, kind, MU, NU, h)
    scale = np.abs(expected).max() + 1e-300
    np.testing.assert_allclose(actual, expected, rtol=0, atol=1e-14 * scale)


@pytest.mark.parametrize('kind', ['truncated', 'filtered', 'willot'])
@pytest.mark.parametrize('n', [(64, 128, 256), (8, 16, 512), (512, 12),
                               (128, 64), (256, 10, 64), (160, 80, 64),
                               (16, 160, 80), (16, 80, 160), (160, 80),
                               (80, 160), (96, 96, 32), (16, 32, 96),
                               (96, 96)])
@pytest.mark.parametrize('variant', [0, 1])
def test_apply_plan_variants_vs_oracle(kind, n, variant):
    """Every pass on plan variant 0 or 1 (8 values per thread for lengths
    64 ... 512, 10 for 80 and 160, 6 for 96, each length in the Gamma-hat,
    axis-1 and last-axis passes): the same operator up to rounding."""
    rng = np.random.default_rng(zlib.crc32(repr((kind, n, 'v')).encode()))
    dim = len(n)
    sym = 3 if dim == 2 else 6
    x = rng.standard_normal(n + (sym,))
    green = FACTORIES[kind](material.create(MU, NU, dim).green_operator(), n,
                            1.0, transform=fft.create_real(n))
    for p in green.PLAN_PASSES:
        green.set_plan(p, variant)
    assert set(green.plans().values()) == {variant}
    actual = np.asarray(green.apply(x))
    expected = jo.green_apply(x, kind, MU, NU, 1.0)
    scale = np.abs(expected).max()
    np.testing.assert_allclose(actual, expected, rtol=0, atol=1e-14 * scale)


def test_set_plan_rejects_bad_arguments():
    green = filtered(material.create(MU, NU, 3).green_operator(), (8, 8, 8),
                     1.0, transform=fft.create_real((8, 8, 8)))
    with pytest.raises(ValueError):
        green.set_plan(5, 0)
    with pytest.raises(ValueError):
        green.set_plan('gamma', 2)


def test_apply_device_tensors():
    import torch
    n = (10, 12, 14)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(n + (6,))
    green = filtered(material.create(2.0, 0.25, 3).green_operator(), n, 1.0,
                     fft.create_real(n))
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    assert green.apply(xd, yd) is yd
    np.testing.assert_array_equal(yd.cpu().numpy(), np.asarray(green.apply(x)))
    # in place on the device
    green.apply(xd, xd)
    np.testing.assert_array_equal(xd.cpu().numpy(), yd.cpu().numpy())


def test_apply_is_projector_on_compatible_fields():
    """Gamma0 annihilates constant fields and Gamma0 C0 Gamma0 = Gamma0 for
    the truncated operator on odd grids (no Nyquist planes, where janus's
    real-output convention symmetrizes Gamma-hat)."""
    import torch
    n = (45, 35, 27)
    g0, nu0 = 1.3, 0.27
    mat = material.create(g0, nu0, 3)
    green = truncated(mat.green_operator(), n, 1.0, fft.create_real(n))
    rng = np.random.default_rng(8)
    x = torch.from_numpy(rng.standard_normal(n + (6,))).cuda()
    const = torch.ones(n + (6,), dtype=torch.float64, device='cuda')
    assert green.apply(const).abs().max().item() < 1e-15
    gx = green.apply(x)
    # C0 : (Gamma0 x) with C0 = 3 k0 J + 2 g0 K
    c0 = torch.from_numpy(
        jo.iso4_matrix(3 * mat.k, 2 * mat.g, 3)).cuda()
    c0gx = (gx @ c0.T.contiguous()).contiguous()
    ggx = green.apply(c0gx)
    err = (ggx - gx).abs().max().item() / gx.abs().max().item()
    assert err < 1e-13
