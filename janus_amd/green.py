"""Discrete Green operators with the interface of janus.green.

``truncated`` and ``filtered`` (janus/green.pyx:22-36) build operators whose
``apply`` runs the fused GPU passes of ``libjanus_amd.so``: last-axis r2c,
the other axes' forward transforms, Gamma-hat recomputed on the fly from the
frequency index and applied between the forward and inverse transforms of
axis 0, then the inverse transforms and the last-axis c2r.  Gamma-hat is
never stored.

* truncated: Gamma0 of the aliased wave vector k_i = s_i (b_i - n_i) if
  2 b_i > n_i else s_i b_i, s_i = 2 pi / (h n_i) (green.pyx:432-451).
* filtered: sum over the 2^dim aliases (b_i - n_i or b_i) of
  prod_i cos^2(h k_i / 4) Gamma0(k) (green.pyx:607-694, 730-910).
* willot2015: Gamma0 of Willot's modified wave vector
  k_i = sin(phi_i) prod_{j != i} cos(phi_j), phi_i = s_i h b_i / 2
  (green.pyx:976-1036): the symbol of the rotated-grid finite-difference
  gradient, up to a common complex factor.

The per-frequency API (``set_frequency``, ``to_memoryview``,
``apply_by_freq``; green.pyx:146-180, 312-351) is evaluated by the host code
of the same library and works without a GPU.
"""

import ctypes

import numpy as np

from . import _device, _lib
from .operators import AbstractStructuredOperator2D, \
    AbstractStructuredOperator3D


def truncated(green, n, h, transform=None):
    if green.dim == 2:
        return TruncatedGreenOperator2D(green, n, h, transform)
    elif green.dim == 3:
        return TruncatedGreenOperator3D(green, n, h, transform)
    else:
        raise ValueError('dim must be 2 or 3 (was {0})'.format(green.dim))


def filtered(green, n, h, transform=None):
    if green.dim == 2:
        return FilteredGreenOperator2D(green, n, h, transform)
    elif green.dim == 3:
        return FilteredGreenOperator3D(green, n, h, transform)
    else:
        raise ValueError('dim must be 2 or 3 (was {0})'.format(green.dim))


def willot2015(green, n, h, transform=None):
    if green.dim == 2:
        return FiniteDifferences2D(green, n, h, transform)
    elif green.dim == 3:
        return FiniteDifferences3D(green, n, h, transform)
    else:
        raise ValueError('dim must be 2 or 3 (was {0})'.format(green.dim))


class _DiscreteGreenOperator:
    """Shared implementation (green.pyx:89-216, 219-390)."""

    _kind = None

    def __init__(self, green, shape, h, transform=None):
        dim = self.dim
        if len(shape) != dim:
            raise ValueError('length of shape must be {0} (was {1})'
                             .format(dim, len(shape)))
        if green.dim != dim:
            raise ValueError('continuous green operator must operate in '
                             'a {0}D space'.format(dim))
        h = float(h)
        if h <= 0.:
            raise ValueError('h must be > 0 (was {0})'.format(h))
        if transform is not None:
            # tuple comparison, as green.pyx:113-116
            if transform.global_ishape != shape:
                raise ValueError('shape of transform must be {0} [was {1}]'
                                 .format(shape, transform.global_ishape))
            grid = tuple(transform.ishape)
        else:
            grid = tuple(int(n) for n in shape)
        self.init_shapes(*(grid + (green.isize, green.osize)))
        self.green = green
        self.h = h
        self.transform = transform
        self._grid = grid
        self._b = None
        self._matrix = None
        mat = green.mat
        hdl = ctypes.c_void_p()
        _lib.check(_lib.lib().jx_green_create(
            dim, _lib.shape_array(grid), h, self._kind, float(mat.g),
            float(mat.nu), ctypes.byref(hdl)))
        self._handle = hdl

    @property
    def handle(self):
        return self._handle

    # register FFT plan variant of each pass (no janus counterpart): 0 last-
    # axis r2c, 1 axis-1 transforms, 2 Gamma-hat pass, 3 last-axis c2r,
    # 4 fused update pass; variant 0 (16 values per thread) or 1 (8)
    PLAN_PASSES = ('rows_r2c', 'axis', 'gamma', 'rows_c2r', 'fused')

    def set_plan(self, pass_, variant):
        """Plan variant of one pass (name or index); results change by
        rounding only."""
        p = self.PLAN_PASSES.index(pass_) if isinstance(pass_, str) else pass_
        _lib.check(_lib.lib().jx_green_set_plan(self._handle, int(p),
                                                int(variant)))

    def get_plan(self, pass_):
        p = self.PLAN_PASSES.index(pass_) if isinstance(pass_, str) else pass_
        return _lib.lib().jx_green_get_plan(self._handle, int(p))

    def plans(self):
        """{pass name: variant}."""
        return {name: self.get_plan(i)
                for i, name in enumerate(self.PLAN_PASSES)}

    def __del__(self):
        hdl = getattr(self, '_handle', None)
        if hdl is not None and _lib._lib is not None:
            _lib._lib.jx_green_destroy(hdl)
            self._handle = None

    def __repr__(self):
        return '{0}({1}, shape={2}, h={3})'.format(
            type(self).__name__, self.green, self._grid, self.h)

    # ------------------------------------------------ per-frequency API
    def set_frequency(self, b):
        """Set the current frequency b (integer multi-index)."""
        b = np.asarray(b)
        if b.shape != (self.dim,):
            raise ValueError('invalid shape: expected ({0},), actual ({1},)'
                             .format(self.dim, b.shape[0] if b.ndim else 0))
        bb = np.ascontiguousarray(b, dtype=np.int32)
        for i in range(self.dim):
            if bb[i] < 0 or bb[i] >= self._grid[i]:
                raise ValueError('index must be >= 0 and < {0} (was {1})'
                                 .format(self._grid[i], bb[i]))
        sym = self.green.osize
        m = np.empty((sym, sym), dtype=np.float64)
        _lib.check(_lib.lib().jx_green_matrix_at(
            self._handle, bb.ctypes.data, m.ctypes.data))
        self._b = bb
        self._matrix = m

    def _current(self):
        if self._matrix is None:
            raise ValueError('set_frequency() must be called first')
        return self._matrix

    def to_memoryview(self, out=None):
        m = self._current()
        shape = (self.oshape[-1], self.ishape[-1])
        if out is None:
            return m.copy()
        if tuple(out.shape) != shape:
            raise ValueError('invalid shape: expected {0}, actual {1}'
                             .format(shape, tuple(out.shape)))
        out[...] = m
        return out.view()

    def apply_by_freq(self, tau, eta=None):
        """Gamma-hat(b) tau for a real Mandel vector; eta may be tau."""
        m = self._current()
        if np.shape(tau) != (self.ishape[-1],):
            raise ValueError('invalid shape of tau: expected ({0},), actual {1}'
                             .format(self.ishape[-1], np.shape(tau)))
        r = m.dot(np.asarray(tau, dtype=np.float64))
        if eta is None:
            return r
        if np.shape(eta) != (self.oshape[-1],):
            raise ValueError('invalid shape of eta: expected ({0},), actual {1}'
                             .format(self.oshape[-1], np.shape(eta)))
        eta[...] = r
        return eta.view()

    # ------------------------------------------------ grid apply (GPU)
    def _apply(self, x, y):
        if self.transform is None:
            raise ValueError('this operator was created without a transform; '
                             'only the per-frequency API is available')
        L = _lib.lib()

        def run(xd, yd):
            _lib.check(L.jx_green_apply(self._handle, xd.data_ptr(),
                                        yd.data_ptr(), _lib.current_stream()))
        if _device.is_device_tensor(x):
            xd = _device.require_f64(x, 'x').contiguous()
            if y is not None:
                _device.require_f64(y, 'y')
            if y is not None and _device.is_device_tensor(y) and \
                    y.is_contiguous():
                run(xd, y)
                return y
            yd = _device.empty_device(self.oshape)
            run(xd, yd)
            if y is not None:
                y.copy_(yd)
                return y
            return yd
        _device.check_float_array(np.asarray(x), 'x')
        xd = _device.to_device(x)
        run(xd, xd)   # in place on the device copy
        return _device.to_host_into(xd, y)


class DiscreteGreenOperator2D(_DiscreteGreenOperator,
                              AbstractStructuredOperator2D):
    pass


class DiscreteGreenOperator3D(_DiscreteGreenOperator,
                              AbstractStructuredOperator3D):
    pass


class TruncatedGreenOperator2D(DiscreteGreenOperator2D):
    _kind = _lib.JX_GREEN_TRUNCATED


class TruncatedGreenOperator3D(DiscreteGreenOperator3D):
    _kind = _lib.JX_GREEN_TRUNCATED


class FilteredGreenOperator2D(DiscreteGreenOperator2D):
    _kind = _lib.JX_GREEN_FILTERED


class FilteredGreenOperator3D(DiscreteGreenOperator3D):
    _kind = _lib.JX_GREEN_FILTERED


class FiniteDifferences2D(DiscreteGreenOperator2D):
    _kind = _lib.JX_GREEN_WILLOT


class FiniteDifferences3D(DiscreteGreenOperator3D):
    _kind = _lib.JX_GREEN_WILLOT
