"""GPU real FFT vs numpy.fft at the shapes and error bounds of janus's
tests/test_fft.py:7-70 (relative L1 error), through the janus API."""

import numpy as np
import pytest

import janus_amd.fft.serial as fft

pytestmark = pytest.mark.gpu

SHAPE_DELTA = [((256, 128), 2E-15), ((256, 129), 2E-14), ((257, 128), 2E-15),
               ((257, 129), 2E-14), ((128, 32, 64), 3E-15),
               ((127, 32, 64), 3E-15), ((128, 31, 64), 3.5E-15),
               ((128, 32, 63), 3.5E-15)]


def rel_l1(actual, expected):
    return np.sum(np.abs(actual - expected)) / np.sum(np.abs(expected))


@pytest.mark.parametrize('shape, delta', SHAPE_DELTA)
@pytest.mark.parametrize('inplace', [True, False])
def test_r2c(shape, delta, inplace):
    transform = fft.create_real(shape)
    np.random.seed(20150312)
    a = 2. * np.random.rand(*transform.ishape) - 1.
    out = np.empty(transform.oshape) if inplace else None
    res = transform.r2c(a, out)
    if inplace:
        assert res.base is out
    res = np.asarray(res)
    actual = res[..., 0::2] + 1j * res[..., 1::2]
    assert rel_l1(actual, np.fft.rfftn(a)) <= delta


@pytest.mark.parametrize('shape, delta', SHAPE_DELTA)
@pytest.mark.parametrize('inplace', [True, False])
def test_c2r(shape, delta, inplace):
    transform = fft.create_real(shape)
    np.random.seed(20150312)
    expected = 2. * np.random.rand(*transform.ishape) - 1.
    a = np.asarray(transform.r2c(expected))
    a_copy = a.copy()
    if inplace:
        actual = transform.c2r(a, np.empty(transform.ishape))
    else:
        actual = transform.c2r(a)
    assert rel_l1(np.asarray(actual), expected) <= delta
    np.testing.assert_array_equal(a, a_copy)   # input preserved


@pytest.mark.parametrize('shape', [(8, 8), (3, 5), (1, 6), (6, 1), (2, 2),
                                   (4, 4, 4), (5, 3, 7), (2, 9, 1), (1, 1, 1),
                                   (30, 50, 64), (40, 50, 60), (200, 300),
                                   (80, 96), (160, 80), (96, 160),
                                   (16, 96, 80), (20, 160, 24), (80, 96, 160),
                                   (1024, 24), (6, 1024, 10), (12, 10, 1024)])
def test_roundtrip_and_numpy(shape):
    rng = np.random.default_rng(11)
    transform = fft.create_real(shape)
    a = rng.standard_normal(shape)
    c = np.asarray(transform.r2c(a))
    ref = np.fft.rfftn(a)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(c[..., 0::2], ref.real, rtol=0, atol=2e-15 * scale * np.log2(a.size + 2))
    np.testing.assert_allclose(c[..., 1::2], ref.imag, rtol=0, atol=2e-15 * scale * np.log2(a.size + 2))
    back = np.asarray(transform.c2r(c))
    np.testing.assert_allclose(back, a, rtol=0, atol=1e-14)


def test_c2r_drops_dc_and_nyquist_imaginary_parts():
    """Non-Hermitian input: c2r keeps only Re of the DC / Nyquist entries of
    the last axis, as numpy.fft.irfftn (pocketfft) and FFTW do."""
    rng = np.random.default_rng(5)
    shape = (6, 8)
    transform = fft.create_real(shape)
    c = rng.standard_normal(transform.oshape)
    z = c[..., 0::2] + 1j * c[..., 1::2]
    expected = np.fft.irfftn(z, s=shape, axes=(0, 1))
    np.testing.assert_allclose(np.asarray(transform.c2r(c)), expected,
                               rtol=0, atol=1e-15)


def test_device_tensors_in_place():
    import torch
    shape = (16, 12, 10)
    transform = fft.create_real(shape)
    rng = np.random.default_rng(2)
    a = rng.standard_normal(shape)
    ad = torch.from_numpy(a).cuda()
    cd = torch.empty(transform.oshape, dtype=torch.float64, device='cuda')
    ret = transform.r2c(ad, cd)
    assert ret is cd
    np.testing.assert_allclose(cd.cpu().numpy(), np.asarray(transform.r2c(a)),
                               rtol=0, atol=0)
    back = transform.c2r(cd)
    np.testing.assert_allclose(back.cpu().numpy(), a, atol=1e-14)
