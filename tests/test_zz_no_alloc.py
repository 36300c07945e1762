"""jx_green_apply_static, the allocation-free variant of jx_green_apply,
against jx_green_apply through the C ABI.

Two calls in a row on one GPU share the static spectrum scratch; the second
is on a smaller grid (and other FFT kernels: lengths without a register
plan).  Each static result must equal the handle path's bit for bit (same
kernels, same tables), in and out of place, and a grid whose spectrum does
not fit the static array is refused before any launch.
"""

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd import _lib
from janus_amd.green import filtered, truncated

pytestmark = pytest.mark.gpu


def _apply(fn, green, x, y):
    _lib.check(fn(green._handle, x.data_ptr(), y.data_ptr(),
                  _lib.current_stream()))


@pytest.mark.parametrize('case', [
    # (factory, shape): the second grid is smaller than the first
    [(filtered, (96, 80, 64)), (truncated, (50, 40))],
])
def test_static_apply_equals_handle_apply(case):
    import torch
    L = _lib.lib()
    for factory, shape in case:
        dim = len(shape)
        sym = 3 if dim == 2 else 6
        mat0 = material.create(1.7, 0.27, dim)
        green = factory(mat0.green_operator(), shape, 1.0,
                        fft.create_real(shape))
        rng = np.random.default_rng(sum(shape))
        x = torch.from_numpy(rng.standard_normal(shape + (sym,))).cuda()
        x0 = x.clone()
        y_static = torch.full_like(x, np.nan)
        # the static call first: it must not need anything the handle path
        # would have allocated on its first use
        _apply(L.jx_green_apply_static, green, x, y_static)
        torch.cuda.synchronize()
        assert torch.equal(x, x0)
        y_ref = torch.empty_like(x)
        _apply(L.jx_green_apply, green, x, y_ref)
        torch.cuda.synchronize()
        print('static apply %s %s: max |diff| = %.3g, max |y| = %.3g'
              % (factory.__name__, shape,
                 float((y_static - y_ref).abs().max()),
                 float(y_ref.abs().max())), flush=True)
        assert torch.equal(y_static, y_ref)
        # in place
        xi = x.clone()
        _apply(L.jx_green_apply_static, green, xi, xi)
        torch.cuda.synchronize()
        assert torch.equal(xi, y_ref)


def test_static_apply_refuses_grid_beyond_scratch():
    import torch
    shape = (128, 128, 128)   # spectrum 128 * 128 * 65 * 96 B > 64 MiB
    mat0 = material.create(1.0, 0.3, 3)
    green = filtered(mat0.green_operator(), shape, 1.0,
                     fft.create_real(shape))
    x = torch.zeros(shape + (6,), dtype=torch.float64, device='cuda')
    y = torch.full_like(x, 7.0)
    with pytest.raises(ValueError, match='static'):
        _apply(_lib.lib().jx_green_apply_static, green, x, y)
    torch.cuda.synchronize()
    assert float(y.min()) == 7.0 and float(y.max()) == 7.0
