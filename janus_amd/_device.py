"""Host <-> device plumbing for the janus-compatible facade.

janus's Python API takes and returns host arrays (memoryviews).  The facade
keeps that contract for numpy inputs (copy in, run the HIP path, copy out)
and also accepts torch tensors already on the GPU, which are used in place.
"""

import numpy as np

from . import _lib


def torch():
    import torch as _torch
    return _torch


def is_device_tensor(a):
    try:
        import torch as _torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(a, _torch.Tensor) and a.is_cuda


def shape_of(a):
    return tuple(int(n) for n in a.shape)


def to_device(a):
    """Contiguous float64 CUDA tensor holding the values of ``a``."""
    t = torch()
    _lib.require_gpu()
    if isinstance(a, t.Tensor):
        if a.dtype != t.float64:
            raise ValueError('float64 data expected (was {0})'.format(a.dtype))
        return a.to('cuda').contiguous()
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return t.from_numpy(arr).to('cuda')


def empty_device(shape, dtype=None):
    t = torch()
    _lib.require_gpu()
    return t.empty(shape, dtype=dtype or t.float64, device='cuda')


def to_host_into(dev, out=None):
    """Copy a device tensor to host.  Into ``out`` (numpy) when given, and
    return a view of ``out`` (janus returns views of the caller's arrays,
    tests/test_operators.py:77-81)."""
    host = dev.cpu().numpy()
    if out is None:
        return host
    out[...] = host
    return out.view()


def require_f64(a, name='array'):
    """Raise ValueError unless the torch tensor ``a`` holds float64: the
    kernels read and write 8 bytes per value, so a narrower tensor would be
    overrun (janus's typed memoryviews refuse other dtypes with ValueError
    too)."""
    t = torch()
    if isinstance(a, t.Tensor) and a.dtype != t.float64:
        raise ValueError('{0} must be float64 (was {1})'.format(name, a.dtype))
    return a


def check_float_array(a, name='array'):
    if isinstance(a, np.ndarray) and a.dtype != np.float64:
        raise ValueError('{0} must be float64 (was {1})'.format(name, a.dtype))
