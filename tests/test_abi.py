"""The C-ABI library loads and exports every symbol include/janus_amd.h
declares (no compute calls: this runs without a GPU)."""

import ctypes
import os
import re

from janus_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'janus_amd.h')


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(jx_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert 'jx_solver_solve' in names and 'jx_green_apply' in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert bound == set(declared_symbols())


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.jx_version() >= 1
    assert isinstance(lib.jx_last_error(), bytes)


def test_invalid_arguments_map_to_value_error():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.jx_green_create(4, _lib.shape_array((8, 8, 8, 8)), 1.0, 0, 1.0,
                             0.3, ctypes.byref(h))
    assert rc == _lib.JX_EINVAL
    assert b'dim' in lib.jx_last_error()


def test_load_imports_torch_first():
    """The library binds the HIP runtime torch loads (loaded first, it would
    bind the ROCm install's and the first device call would find no GPU)."""
    import subprocess
    import sys
    code = ('import sys; from janus_amd import _lib; '
            'assert "torch" not in sys.modules; _lib.load(); '
            'assert "torch" in sys.modules')
    subprocess.run([sys.executable, '-c', code], cwd=ROOT, check=True)
