"""jx_tile_probe (the strided passes' access pattern without arithmetic,
bench.py --workload patterns) copies every value of the padded spectrum
layout, leaves the padding alone, and refuses what it does not handle."""

import ctypes

import pytest

from janus_amd import _lib

pytestmark = pytest.mark.gpu


def _shape(*n):
    return (ctypes.c_int64 * 3)(*n)


def test_tile_probe_copies_every_value():
    import torch
    lib = _lib.lib()
    shape = _shape(256, 40, 28)   # slab 40 * 15 * 96 B is not 4 KB-aligned
    nbytes = ctypes.c_int64()
    _lib.check(lib.jx_tile_probe_bytes(shape, ctypes.byref(nbytes)))
    assert nbytes.value == 8 * 12 * 256 * 40 * 15
    src = torch.arange(nbytes.value // 8, dtype=torch.float64, device='cuda')
    for pts in (1, 2, 4):
        dst = torch.zeros_like(src)
        _lib.check(lib.jx_tile_probe(shape, 0, pts, src.data_ptr(),
                                     dst.data_ptr(), nbytes.value,
                                     _lib.current_stream()))
        torch.cuda.synchronize()
        assert torch.equal(src, dst), pts
    # padded layout (slab 256 * 129 * 96 B = 774 * 4 KB): 8448 B after each
    # of the 256 slabs, untouched by the copy
    shape = _shape(64, 256, 256)
    _lib.check(lib.jx_tile_probe_bytes(shape, ctypes.byref(nbytes)))
    slab = 12 * 256 * 129 + 1056
    assert nbytes.value == 8 * 64 * slab
    src = torch.ones(nbytes.value // 8, dtype=torch.float64, device='cuda')
    dst = torch.zeros_like(src)
    _lib.check(lib.jx_tile_probe(shape, 1, 2, src.data_ptr(), dst.data_ptr(),
                                 nbytes.value, _lib.current_stream()))
    torch.cuda.synchronize()
    d = dst.view(64, slab)
    assert float(d[:, :slab - 1056].min()) == 1.0
    assert float(d[:, slab - 1056:].abs().max()) == 0.0


def test_tile_probe_refusals():
    import torch
    lib = _lib.lib()
    shape = _shape(256, 40, 28)
    nbytes = ctypes.c_int64()
    _lib.check(lib.jx_tile_probe_bytes(shape, ctypes.byref(nbytes)))
    n = nbytes.value
    buf = torch.zeros(n // 8 + 2, dtype=torch.float64, device='cuda')
    p = buf.data_ptr()
    s = _lib.current_stream()
    for args in ((shape, 2, 1, p, p, n),                 # axis 2
                 (_shape(128, 40, 28), 0, 1, p, p, n),   # 128-point axis
                 (shape, 0, 3, p, p, n),                 # 3-point tiles
                 (shape, 0, 1, p, p, n - 8),             # small buffers
                 (shape, 0, 1, p + 8, p, n)):            # misaligned
        assert lib.jx_tile_probe(*args, s) == _lib.JX_EINVAL, args
    torch.cuda.synchronize()
    assert float(buf.abs().max()) == 0.0
