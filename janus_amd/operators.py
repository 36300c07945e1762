"""Operators with the interface of janus.operators.

* :func:`isotropic_4` and :class:`FourthRankIsotropicTensor2D/3D` -- the
  per-voxel tensor ``sph J + dev K`` (janus/operators.pyx:53-67, 257-408).
  These single-voxel objects are parameter holders: their ``apply`` acts on
  one Mandel vector on the host, as janus's does.
* :class:`BlockDiagonalOperator2D/3D` -- the voxel-wise local operator on a
  grid (janus/operators.pyx:617-716).  Its ``apply`` runs on the GPU.  The
  janus constructor takes an object array of local operators; every entry
  must be an isotropic tensor, and the array is turned into a phase map plus
  a (sph, dev) table (or per-voxel coefficient fields when it has more than
  256 distinct tensors).  The array-based constructors
  :meth:`BlockDiagonalOperator3D.from_phases` and ``from_fields`` avoid the
  object array, which is not viable at 512^3.
"""

import ctypes

import numpy as np

from . import _device, _lib


def isotropic_4(sph, dev, dim):
    """Create a fourth-rank isotropic tensor sph J + dev K
    (operators.pyx:53-67)."""
    if dim == 2:
        return FourthRankIsotropicTensor2D(sph, dev, dim)
    elif dim == 3:
        return FourthRankIsotropicTensor3D(sph, dev, dim)
    else:
        raise ValueError('dim must be 2 or 3 (was {0})'.format(dim))


def block_diagonal_operator(loc):
    """Create a block-diagonal operator from a 2D or 3D array of local
    operators (operators.pyx:70-87)."""
    dim = len(loc.shape)
    if dim == 2:
        return BlockDiagonalOperator2D(loc)
    elif dim == 3:
        return BlockDiagonalOperator3D(loc)
    else:
        raise ValueError('number of dimensions of loc must be 2 or 3 '
                         '(was {0})'.format(dim))


def _check_vector(x, n, what='x'):
    if np.shape(x) != (n,):
        raise ValueError('invalid shape of {0}: expected ({1},), actual {2}'
                         .format(what, n, np.shape(x)))


class FourthRankIsotropicTensor:
    """sph J + dev K with tr = (sph - dev) / dim (operators.pyx:257-303)."""

    def __init__(self, sph, dev, dim):
        sym = (dim * (dim + 1)) // 2
        self._isize = sym
        self._osize = sym
        self._dim = dim
        self._sph = float(sph)
        self._dev = float(dev)
        self._tr = (self._sph - self._dev) / float(dim)

    isize = property(lambda self: self._isize)
    osize = property(lambda self: self._osize)
    dim = property(lambda self: self._dim)
    sph = property(lambda self: self._sph)
    dev = property(lambda self: self._dev)
    tr = property(lambda self: self._tr)

    def __repr__(self):
        return ('FourthRankIsotropicTensor(sph={0}, dev={1}, dim={2})'
                .format(self._sph, self._dev, self._dim))

    def _apply(self, x):
        d, tr, dev = self._dim, self._tr, self._dev
        y = np.empty(self._osize)
        if d == 2:
            aux = tr * (x[0] + x[1])
        else:
            aux = tr * (x[0] + x[1] + x[2])
        for i in range(d):
            y[i] = aux + dev * x[i]
        for i in range(d, self._osize):
            y[i] = dev * x[i]
        return y

    def apply(self, x, y=None):
        """y = T : x for one Mandel vector (operators.pyx:315-324, 352-365);
        in place allowed."""
        _check_vector(x, self._isize)
        r = self._apply(np.asarray(x, dtype=np.float64))
        if y is None:
            return r
        _check_vector(y, self._osize, 'y')
        y[...] = r
        return y.view()

    def to_memoryview(self, out=None):
        """Matrix of the tensor (operators.pyx:326-340, 367-408)."""
        d, sym = self._dim, self._osize
        a = np.zeros((sym, sym), dtype=np.float64)
        if d == 2:
            diag = (self._sph + self._dev) / 2.0
            off = 0.5 * (self._sph - self._dev)
        else:
            diag = (self._sph + 2.0 * self._dev) / 3.0
            off = (self._sph - self._dev) / 3.0
        a[:d, :d] = off
        for i in range(d):
            a[i, i] = diag
        for i in range(d, sym):
            a[i, i] = self._dev
        if out is None:
            return a
        if tuple(out.shape) != (sym, sym):
            raise ValueError('invalid shape: expected {0}, actual {1}'
                             .format((sym, sym), out.shape))
        out[...] = a
        return out.view()


class FourthRankIsotropicTensor2D(FourthRankIsotropicTensor):
    pass


class FourthRankIsotropicTensor3D(FourthRankIsotropicTensor):
    pass


# ---------------------------------------------------------------------------
# Structured operators
# ---------------------------------------------------------------------------

class AbstractStructuredOperator:
    """Operator on vector data laid out on a 2D/3D grid
    (operators.pyx:452-614).  Input shape ``ishape = grid + (isize,)``,
    output ``oshape = grid + (osize,)``."""

    dim = None

    def init_shapes(self, *args):
        """init_shapes(shape0, shape1, [shape2,] ishapeD, oshapeD)."""
        if len(args) != self.dim + 2:
            raise ValueError('expected {0} arguments'.format(self.dim + 2))
        names = ['shape{0}'.format(i) for i in range(self.dim)] + \
            ['ishape{0}'.format(self.dim), 'oshape{0}'.format(self.dim)]
        for name, v in zip(names, args):
            if int(v) <= 0:
                raise ValueError('{0} should be > 0 (was {1})'.format(name, v))
        for name, v in zip(names, args):
            setattr(self, name, int(v))
        grid = tuple(int(v) for v in args[:self.dim])
        self.ishape = grid + (int(args[self.dim]),)
        self.oshape = grid + (int(args[self.dim + 1]),)

    def _check_io(self, x, y):
        if tuple(np.shape(x)) != self.ishape:
            raise ValueError('invalid shape of x: expected {0}, actual {1}'
                             .format(self.ishape, tuple(np.shape(x))))
        if y is not None and tuple(np.shape(y)) != self.oshape:
            raise ValueError('invalid shape of y: expected {0}, actual {1}'
                             .format(self.oshape, tuple(np.shape(y))))

    def apply(self, x, y=None):
        """Return the image of x; stored in y (and a view of y returned)
        when y is given."""
        self._check_io(x, y)
        return self._apply(x, y)

    def _apply(self, x, y):
        raise NotImplementedError


class AbstractStructuredOperator2D(AbstractStructuredOperator):
    dim = 2


class AbstractStructuredOperator3D(AbstractStructuredOperator):
    dim = 3


def _run_device(fn, x, y, oshape):
    """Run fn(x_dev, y_dev) on device copies of host arrays, or in place on
    device tensors; returns what janus's apply returns."""
    if _device.is_device_tensor(x):
        xd = x.contiguous()
        if y is not None and _device.is_device_tensor(y) and y.is_contiguous():
            yd = y
        else:
            yd = _device.empty_device(oshape)
        fn(xd, yd)
        if y is not None and yd is not y:
            y.copy_(yd)
            return y
        return yd
    _device.check_float_array(np.asarray(x), 'x')
    xd = _device.to_device(x)
    yd = xd if y is x else _device.empty_device(oshape)
    fn(xd, yd)
    return _device.to_host_into(yd, y)


class _BlockDiagonalIsotropic:
    """Shared implementation of the block-diagonal isotropic operator."""

    def _setup_phases(self, phase, table):
        t = _device.torch()
        _lib.require_gpu()
        phase = np.asarray(phase)
        table = np.ascontiguousarray(table, dtype=np.float64).reshape(-1, 2)
        if phase.shape != self.ishape[:-1]:
            raise ValueError('phase map shape {0} does not match the grid {1}'
                             .format(phase.shape, self.ishape[:-1]))
        if table.shape[0] < 1 or table.shape[0] > 256:
            raise ValueError('between 1 and 256 phases are supported')
        if phase.size and (phase.min() < 0 or phase.max() >= table.shape[0]):
            raise ValueError('phase index out of range of the table')
        self._phase_dev = t.from_numpy(
            np.ascontiguousarray(phase, dtype=np.uint8)).to('cuda')
        self._table = table.copy()
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().jx_local_create_phases(
            self.dim, _lib.shape_array(self.ishape[:-1]),
            self._phase_dev.data_ptr(), self._table.ctypes.data,
            int(table.shape[0]), ctypes.byref(h)))
        self._handle = h

    def _setup_fields(self, sph, dev):
        _lib.require_gpu()
        self._sph_dev = _device.to_device(sph)
        self._dev_dev = _device.to_device(dev)
        if tuple(self._sph_dev.shape) != self.ishape[:-1] or \
                tuple(self._dev_dev.shape) != self.ishape[:-1]:
            raise ValueError('coefficient fields must have the grid shape {0}'
                             .format(self.ishape[:-1]))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().jx_local_create_fields(
            self.dim, _lib.shape_array(self.ishape[:-1]),
            self._sph_dev.data_ptr(), self._dev_dev.data_ptr(),
            ctypes.byref(h)))
        self._handle = h

    def _setup_from_loc(self, loc):
        if loc is None:
            raise ValueError('loc should not be None')
        loc = np.asarray(loc, dtype=object)
        if loc.ndim != self.dim:
            raise ValueError('loc must be a {0}D array'.format(self.dim))
        first = loc.flat[0]
        self.init_shapes(*(loc.shape + (first.isize, first.osize)))
        ids = {}
        table = []
        phase = np.empty(loc.shape, dtype=np.int64)
        flat = phase.reshape(-1)
        for i, op in enumerate(loc.flat):
            if op.isize != self.ishape[-1]:
                raise ValueError('invalid dimension block operator input: '
                                 'expected {0}, actual {1}'
                                 .format(self.ishape[-1], op.isize))
            if op.osize != self.oshape[-1]:
                raise ValueError('invalid dimension block operator output: '
                                 'expected {0}, actual {1}'
                                 .format(self.oshape[-1], op.osize))
            if not isinstance(op, FourthRankIsotropicTensor) or \
                    op.dim != self.dim:
                raise NotImplementedError(
                    'janus_amd block-diagonal operators take isotropic '
                    'tensors (isotropic_4) of dimension {0}'.format(self.dim))
            key = (op.sph, op.dev)
            p = ids.get(key)
            if p is None:
                p = len(table)
                ids[key] = p
                table.append(key)
            flat[i] = p
        self.loc = loc
        if len(table) <= 256:
            self._setup_phases(phase, np.array(table))
        else:
            tab = np.array(table)
            self._setup_fields(tab[phase, 0], tab[phase, 1])

    def _init_arrays(self, grid, sym):
        self.init_shapes(*(tuple(grid) + (sym, sym)))
        self.loc = None

    def _apply(self, x, y):
        L = _lib.lib()

        def run(xd, yd):
            _lib.check(L.jx_local_apply(self._handle, xd.data_ptr(),
                                        yd.data_ptr(), _lib.current_stream()))
        return _run_device(run, x, y, self.oshape)

    def means(self, x):
        """(<T:x>, <x>) over the grid, T the local operator (device x)."""
        xd = _device.to_device(x)
        sym = self.ishape[-1]
        tau = np.empty(sym)
        eps = np.empty(sym)
        _lib.check(_lib.lib().jx_local_means(
            self._handle, xd.data_ptr(), tau.ctypes.data, eps.ctypes.data,
            _lib.current_stream()))
        return tau, eps

    @property
    def handle(self):
        return self._handle

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and _lib._lib is not None:
            _lib._lib.jx_local_destroy(h)
            self._handle = None


class BlockDiagonalOperator2D(_BlockDiagonalIsotropic,
                              AbstractStructuredOperator2D):
    """Block-diagonal operator with 2D layout (operators.pyx:617-664)."""

    def __init__(self, loc=None, _arrays=None):
        self._handle = None
        if _arrays is None:
            self._setup_from_loc(loc)
        else:
            kind, grid, a, b = _arrays
            self._init_arrays(grid, 3)
            if kind == 'phases':
                self._setup_phases(a, b)
            else:
                self._setup_fields(a, b)

    @classmethod
    def from_phases(cls, phase, table):
        """From a (n0, n1) phase map and an (nphase, 2) table of (sph, dev)."""
        return cls(_arrays=('phases', np.shape(phase), phase, table))

    @classmethod
    def from_fields(cls, sph, dev):
        """From per-voxel coefficient fields sph, dev of shape (n0, n1)."""
        return cls(_arrays=('fields', tuple(sph.shape), sph, dev))


class BlockDiagonalOperator3D(_BlockDiagonalIsotropic,
                              AbstractStructuredOperator3D):
    """Block-diagonal operator with 3D layout (operators.pyx:667-716)."""

    def __init__(self, loc=None, _arrays=None):
        self._handle = None
        if _arrays is None:
            self._setup_from_loc(loc)
        else:
            kind, grid, a, b = _arrays
            self._init_arrays(grid, 6)
            if kind == 'phases':
                self._setup_phases(a, b)
            else:
                self._setup_fields(a, b)

    @classmethod
    def from_phases(cls, phase, table):
        """From a (n0, n1, n2) phase map and an (nphase, 2) table."""
        return cls(_arrays=('phases', np.shape(phase), phase, table))

    @classmethod
    def from_fields(cls, sph, dev):
        """From per-voxel coefficient fields of shape (n0, n1, n2)."""
        return cls(_arrays=('fields', tuple(sph.shape), sph, dev))
