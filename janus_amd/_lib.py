"""ctypes binding of ``libjanus_amd.so`` (the C ABI in include/janus_amd.h).

The library is built in-tree by :func:`janus_amd.build.build_library`.  If it
is missing or no GPU is visible, every compute entry point raises: there is
no CPU fallback on the product path.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# JANUS_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get('JANUS_AMD_LIB',
                          os.path.join(_HERE, 'libjanus_amd.so'))

JX_OK = 0
JX_EINVAL = -1
JX_EHIP = -2
JX_ENOMEM = -3
JX_EUNSUPPORTED = -4

JX_GREEN_TRUNCATED = 0
JX_GREEN_FILTERED = 1
JX_GREEN_WILLOT = 2

JX_SOLVE_CONVERGED = 1
JX_SOLVE_MAXITER = 2
JX_SOLVE_NONFINITE = 3

# (name, restype, argtypes) of every exported symbol of include/janus_amd.h
_c_int64_p = ctypes.POINTER(ctypes.c_int64)
_c_double_p = ctypes.c_void_p
_VP = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = [
    ('jx_last_error', ctypes.c_char_p, []),
    ('jx_version', ctypes.c_int, []),
    ('jx_fft_create', ctypes.c_int, [ctypes.c_int, _c_int64_p, _PP]),
    ('jx_fft_destroy', ctypes.c_int, [_VP]),
    ('jx_fft_r2c', ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ('jx_fft_c2r', ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ('jx_green_create', ctypes.c_int,
     [ctypes.c_int, _c_int64_p, ctypes.c_double, ctypes.c_int,
      ctypes.c_double, ctypes.c_double, _PP]),
    ('jx_green_destroy', ctypes.c_int, [_VP]),
    ('jx_green_apply', ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ('jx_green_apply_static', ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ('jx_green_matrix_at', ctypes.c_int, [_VP, _VP, _VP]),
    ('jx_gamma0_at', ctypes.c_int,
     [ctypes.c_int, ctypes.c_double, ctypes.c_double, _VP, _VP]),
    ('jx_local_create_phases', ctypes.c_int,
     [ctypes.c_int, _c_int64_p, _VP, _VP, ctypes.c_int, _PP]),
    ('jx_local_create_fields', ctypes.c_int,
     [ctypes.c_int, _c_int64_p, _VP, _VP, _PP]),
    ('jx_local_destroy', ctypes.c_int, [_VP]),
    ('jx_local_apply', ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ('jx_local_means', ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    ('jx_local_create_phase_matrices', ctypes.c_int,
     [ctypes.c_int, _c_int64_p, _VP, _VP, ctypes.c_int, _PP]),
    ('jx_local_create_matrix_field', ctypes.c_int,
     [ctypes.c_int, _c_int64_p, _VP, _PP]),
    ('jx_local_is_symmetric', ctypes.c_int,
     [_VP, ctypes.POINTER(ctypes.c_int)]),
    ('jx_local_apply_transpose', ctypes.c_int, [_VP, _VP, _VP, _VP]),
    ('jx_block_linear_apply', ctypes.c_int,
     [ctypes.c_int, _c_int64_p, ctypes.c_int, ctypes.c_int, _VP, _VP, _VP,
      ctypes.c_int, _VP]),
    ('jx_solver_create', ctypes.c_int, [_VP, _VP, _PP]),
    ('jx_solver_destroy', ctypes.c_int, [_VP]),
    ('jx_solver_solve', ctypes.c_int,
     [_VP, _VP, _VP, _VP, ctypes.c_double, ctypes.c_int, ctypes.c_double,
      _VP, _VP, _VP, _VP]),
    ('jx_adjoint_contract', ctypes.c_int,
     [ctypes.c_int, _c_int64_p, _VP, _VP, _VP, _VP, _VP]),
    ('jx_adjoint_outer', ctypes.c_int,
     [ctypes.c_int, _c_int64_p, _VP, _VP, _VP, _VP]),
    ('jx_stream_copy', ctypes.c_int, [_VP, _VP, ctypes.c_int64, _VP]),
    ('jx_tile_probe_bytes', ctypes.c_int, [_c_int64_p, _c_int64_p]),
    ('jx_tile_probe', ctypes.c_int,
     [_c_int64_p, ctypes.c_int, ctypes.c_int, _VP, _VP, ctypes.c_int64, _VP]),
    ('jx_solver_set_fused', ctypes.c_int, [_VP, ctypes.c_int]),
    ('jx_solver_get_fused', ctypes.c_int, [_VP]),
    ('jx_solver_profile', ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, _VP, _VP]),
    ('jx_solver_solve_cg', ctypes.c_int,
     [_VP, _VP, _VP, _VP, ctypes.c_double, ctypes.c_int, ctypes.c_double,
      _VP, _VP, _VP, _VP]),
    ('jx_solver_solve_cg_polarization', ctypes.c_int,
     [_VP, _VP, _VP, _VP, ctypes.c_double, ctypes.c_int, ctypes.c_double,
      _VP, _VP, _VP, _VP]),
    ('jx_solver_solve_gmres', ctypes.c_int,
     [_VP, _VP, _VP, _VP, ctypes.c_double, ctypes.c_int, ctypes.c_int,
      ctypes.c_double, _VP, _VP, _VP, _VP]),
    ('jx_green_set_plan', ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    ('jx_green_get_plan', ctypes.c_int, [_VP, ctypes.c_int]),
    ('jx_solver_autotune', ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP]),
    ('jx_batch_create', ctypes.c_int, [_PP]),
    ('jx_batch_destroy', ctypes.c_int, [_VP]),
    ('jx_batch_add', ctypes.c_int,
     [_VP, _VP, _VP, _VP, _VP, ctypes.c_double, ctypes.c_int, ctypes.c_double,
      _VP]),
    ('jx_batch_add_cg', ctypes.c_int,
     [_VP, _VP, _VP, _VP, _VP, ctypes.c_double, ctypes.c_int, ctypes.c_double,
      _VP]),
    ('jx_batch_ngroups', ctypes.c_int, [_VP, _VP]),
    ('jx_batch_group', ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP, _VP, _VP, _VP]),
    ('jx_batch_tuning', ctypes.c_int,
     [_VP, ctypes.c_int, ctypes.c_int, _VP, _VP, _VP]),
    ('jx_batch_set_width', ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    ('jx_batch_set_schedule', ctypes.c_int,
     [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ('jx_batch_set_streams', ctypes.c_int, [_VP, ctypes.c_int]),
    ('jx_batch_run', ctypes.c_int, [_VP, _VP]),
    ('jx_batch_result', ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP, _VP]),
    ('jx_batch_stats', ctypes.c_int, [_VP, _VP, _VP, _VP]),
]

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the shared library and declare its signatures (no GPU needed)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(
            'janus_amd: {0} is missing; build it with '
            'python -c "import janus_amd.build as b; b.build_library()"'
            .format(path))
    # torch's HIP runtime must be the process's one: loaded first, the
    # library would bind libamdhip64 from the ROCm install and torch would
    # then share it, and the first device call fails with hipErrorNoDevice.
    import torch  # noqa: F401
    lib = ctypes.CDLL(path)
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    if path == LIB_PATH:
        _lib = lib
    return lib


def lib():
    return load()


def check(rc):
    """Map a JX_* status to a Python exception (janus raises ValueError for
    invalid arguments, janus/utils/checkarray.pyx:3-53)."""
    if rc == JX_OK:
        return
    msg = lib().jx_last_error().decode('utf-8', 'replace')
    if rc == JX_EINVAL:
        raise ValueError(msg)
    if rc == JX_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == JX_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError('janus_amd HIP error: ' + msg)


def shape_array(shape):
    return (ctypes.c_int64 * len(shape))(*[int(n) for n in shape])


def ptr(a):
    """Raw pointer of a torch tensor or numpy array (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('janus_amd: no GPU visible; the HIP path is the '
                           'only implementation')
    torch.cuda.init()
    lib()


def current_stream():
    import torch
    return torch.cuda.current_stream().cuda_stream
