"""Operators with the interface of janus.operators.

* :func:`isotropic_4` and :class:`FourthRankIsotropicTensor2D/3D` -- the
  per-voxel tensor ``sph J + dev K`` (janus/operators.pyx:53-67, 257-408).
  These single-voxel objects are parameter holders: their ``apply`` acts on
  one Mandel vector on the host, as janus's does.
* :class:`BlockDiagonalOperator2D/3D` -- the voxel-wise local operator on a
  grid (janus/operators.pyx:617-716).  Its ``apply`` runs on the GPU.  The
  janus constructor takes an object array of local operators.  Isotropic
  entries become a phase map plus a (sph, dev) table (or per-voxel
  coefficient fields when there are more than 256 distinct tensors); any
  other sym x sym linear entries (:class:`FourthRankCubicTensor2D`,
  :class:`LinearOperator`) become a phase map plus a table of matrices (or
  a per-voxel matrix field).  The array-based constructors ``from_phases``,
  ``from_fields``, ``from_phase_matrices`` and ``from_matrix_field`` avoid
  the object array, which is not viable at 512^3.
* :class:`LinearOperator`, :class:`FourthRankCubicTensor2D` -- single-voxel
  matrices (operators.pyx:195-254, 411-450), host-side like the isotropic
  tensors.
* :class:`BlockDiagonalLinearOperator2D/3D` -- a per-voxel matrix field of
  any block size with ``apply`` and ``apply_transpose`` on the GPU
  (operators.pyx:90-110, 719-890).
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


def _matrix_out(a, out):
    if out is None:
        return a.copy()
    if tuple(np.shape(out)) != a.shape:
        raise ValueError('invalid shape: expected {0}, actual {1}'
                         .format(a.shape, tuple(np.shape(out))))
    out[...] = a
    return out.view()


def _matvec_out(a, x, y, isize, osize):
    _check_vector(x, isize)
    r = a @ np.asarray(x, dtype=np.float64)
    if y is None:
        return r
    _check_vector(y, osize, 'y')
    y[...] = r
    return y.view()


class LinearOperator:
    """Single-voxel operator of matrix ``a`` (osize, isize)
    (operators.pyx:195-254); keeps a reference to ``a`` when it already is
    a float64 array, as janus keeps a shallow copy."""

    def __init__(self, a):
        a = np.asarray(a, dtype=np.float64)
        if a.ndim != 2:
            raise ValueError('a must be a 2D array')
        if a.shape[0] <= 0 or a.shape[1] <= 0:
            raise ValueError('isize and osize should be > 0')
        self.a = a
        self._isize = a.shape[1]
        self._osize = a.shape[0]

    isize = property(lambda self: self._isize)
    osize = property(lambda self: self._osize)

    def apply(self, x, y=None):
        """y = a x (operators.pyx:211-218)."""
        return _matvec_out(self.a, x, y, self._isize, self._osize)

    def apply_transpose(self, x, y=None):
        """y = a^T x (operators.pyx:220-247)."""
        return _matvec_out(self.a.T, x, y, self._osize, self._isize)

    def to_memoryview(self, out=None):
        return _matrix_out(self.a, out)


class FourthRankCubicTensor2D:
    """Fourth-rank tensor with cubic symmetry in 2D, of moduli t1111,
    t1122, t1212 in its material axes, rotated by ``theta``
    (operators.pyx:411-450).

    With e1 = (cos theta, sin theta), e2 = (-sin theta, cos theta), the
    Mandel vectors of e1 (x) e1, e2 (x) e2 and sqrt(2) sym(e1 (x) e2) are an
    orthonormal basis in which the tensor is [[t1111, t1122, 0],
    [t1122, t1111, 0], [0, 0, 2 t1212]]; the matrix is that block changed
    to the grid basis, Q T Q^T, with Q's columns the basis vectors."""

    def __init__(self, t1111, t1122, t1212, theta):
        c = np.cos(theta)
        s = np.sin(theta)
        r2 = np.sqrt(2.0)
        q = np.array([[c * c, s * s, -r2 * c * s],
                      [s * s, c * c, r2 * c * s],
                      [r2 * c * s, -r2 * c * s, c * c - s * s]])
        t0 = np.array([[t1111, t1122, 0.],
                       [t1122, t1111, 0.],
                       [0., 0., 2. * t1212]])
        a = q @ t0 @ q.T
        self._a = 0.5 * (a + a.T)  # exactly symmetric
        self.t1111 = float(t1111)
        self.t1122 = float(t1122)
        self.t1212 = float(t1212)
        self.theta = float(theta)

    isize = property(lambda self: 3)
    osize = property(lambda self: 3)
    dim = property(lambda self: 2)

    def apply(self, x, y=None):
        return _matvec_out(self._a, x, y, 3, 3)

    def to_memoryview(self, out=None):
        return _matrix_out(self._a, out)


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
        xd = _device.require_f64(x, 'x').contiguous()
        if y is not None:
            _device.require_f64(y, 'y')
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


class _BlockDiagonal:
    """Shared implementation of the block-diagonal operators with sym x sym
    blocks (a jx_local handle)."""

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

    def _setup_phase_matrices(self, phase, mats):
        t = _device.torch()
        _lib.require_gpu()
        sym = self.ishape[-1]
        phase = np.asarray(phase)
        mats = np.ascontiguousarray(mats, dtype=np.float64)
        if phase.shape != self.ishape[:-1]:
            raise ValueError('phase map shape {0} does not match the grid {1}'
                             .format(phase.shape, self.ishape[:-1]))
        if mats.ndim != 3 or mats.shape[1:] != (sym, sym):
            raise ValueError('matrices must have shape (nphase, {0}, {0})'
                             .format(sym))
        if mats.shape[0] < 1 or mats.shape[0] > 256:
            raise ValueError('between 1 and 256 phases are supported')
        if phase.size and (phase.min() < 0 or phase.max() >= mats.shape[0]):
            raise ValueError('phase index out of range of the table')
        self._phase_dev = t.from_numpy(
            np.ascontiguousarray(phase, dtype=np.uint8)).to('cuda')
        self._mats = mats.copy()
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().jx_local_create_phase_matrices(
            self.dim, _lib.shape_array(self.ishape[:-1]),
            self._phase_dev.data_ptr(), self._mats.ctypes.data,
            int(mats.shape[0]), ctypes.byref(h)))
        self._handle = h

    def _setup_matrix_field(self, a):
        _lib.require_gpu()
        self._mat_dev = _device.to_device(a)
        sym = self.ishape[-1]
        if tuple(self._mat_dev.shape) != self.ishape[:-1] + (sym, sym):
            raise ValueError('matrix field must have shape {0}'
                             .format(self.ishape[:-1] + (sym, sym)))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().jx_local_create_matrix_field(
            self.dim, _lib.shape_array(self.ishape[:-1]),
            self._mat_dev.data_ptr(), ctypes.byref(h)))
        self._handle = h

    def _setup_from_loc(self, loc):
        if loc is None:
            raise ValueError('loc should not be None')
        loc = np.asarray(loc, dtype=object)
        if loc.ndim != self.dim:
            raise ValueError('loc must be a {0}D array'.format(self.dim))
        first = loc.flat[0]
        self.init_shapes(*(loc.shape + (first.isize, first.osize)))
        sym = (self.dim * (self.dim + 1)) // 2
        iso = True
        for op in loc.flat:
            if op.isize != self.ishape[-1]:
                raise ValueError('invalid dimension block operator input: '
                                 'expected {0}, actual {1}'
                                 .format(self.ishape[-1], op.isize))
            if op.osize != self.oshape[-1]:
                raise ValueError('invalid dimension block operator output: '
                                 'expected {0}, actual {1}'
                                 .format(self.oshape[-1], op.osize))
            if isinstance(op, FourthRankIsotropicTensor):
                if op.dim != self.dim:
                    raise ValueError('isotropic tensor of dimension {0} in a '
                                     '{1}D operator'.format(op.dim, self.dim))
            else:
                iso = False
                if not hasattr(op, 'to_memoryview'):
                    raise NotImplementedError(
                        'janus_amd block-diagonal operators take linear local '
                        'operators (with to_memoryview)')
        if self.ishape[-1] != sym or self.oshape[-1] != sym:
            raise NotImplementedError(
                'janus_amd block-diagonal operators take {0} x {0} local '
                'operators in {1}D'.format(sym, self.dim))
        ids = {}
        table = []
        phase = np.empty(loc.shape, dtype=np.int64)
        flat = phase.reshape(-1)
        for i, op in enumerate(loc.flat):
            if iso:
                key = (op.sph, op.dev)
                value = key
            else:
                value = np.asarray(op.to_memoryview(), dtype=np.float64)
                key = value.tobytes()
            p = ids.get(key)
            if p is None:
                p = len(table)
                ids[key] = p
                table.append(value)
            flat[i] = p
        self.loc = loc
        tab = np.array(table)
        if iso and len(table) <= 256:
            self._setup_phases(phase, tab)
        elif iso:
            self._setup_fields(tab[phase, 0], tab[phase, 1])
        elif len(table) <= 256:
            self._setup_phase_matrices(phase, tab)
        else:
            self._setup_matrix_field(tab[phase])

    def _init_arrays(self, grid, sym):
        self.init_shapes(*(tuple(grid) + (sym, sym)))
        self.loc = None

    def _setup_arrays(self, kind, a, b):
        if kind == 'phases':
            self._setup_phases(a, b)
        elif kind == 'fields':
            self._setup_fields(a, b)
        elif kind == 'phase_matrices':
            self._setup_phase_matrices(a, b)
        else:
            self._setup_matrix_field(a)

    @classmethod
    def from_phase_matrices(cls, phase, mats):
        """From a phase map and an (nphase, sym, sym) table of matrices."""
        return cls(_arrays=('phase_matrices', np.shape(phase), phase, mats))

    @classmethod
    def from_matrix_field(cls, a):
        """From a per-voxel (grid + (sym, sym)) matrix field, host array or
        device tensor (a device tensor is kept by reference)."""
        grid = tuple(int(n) for n in a.shape[:-2])
        return cls(_arrays=('matrix_field', grid, a, None))

    def _apply(self, x, y):
        L = _lib.lib()

        def run(xd, yd):
            _lib.check(L.jx_local_apply(self._handle, xd.data_ptr(),
                                        yd.data_ptr(), _lib.current_stream()))
        return _run_device(run, x, y, self.oshape)

    def apply_transpose(self, x, y=None):
        """Image of x by the transposed operator (the isotropic modes are
        symmetric); y as in apply."""
        self._check_io(x, y)
        L = _lib.lib()

        def run(xd, yd):
            _lib.check(L.jx_local_apply_transpose(
                self._handle, xd.data_ptr(), yd.data_ptr(),
                _lib.current_stream()))
        return _run_device(run, x, y, self.oshape)

    @property
    def symmetric(self):
        """True when every block is a symmetric matrix (needed by
        conjugate gradients)."""
        out = ctypes.c_int()
        _lib.check(_lib.lib().jx_local_is_symmetric(self._handle,
                                                    ctypes.byref(out)))
        return bool(out.value)

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


class BlockDiagonalOperator2D(_BlockDiagonal,
                              AbstractStructuredOperator2D):
    """Block-diagonal operator with 2D layout (operators.pyx:617-664)."""

    def __init__(self, loc=None, _arrays=None):
        self._handle = None
        if _arrays is None:
            self._setup_from_loc(loc)
        else:
            kind, grid, a, b = _arrays
            self._init_arrays(grid, 3)
            self._setup_arrays(kind, a, b)

    @classmethod
    def from_phases(cls, phase, table):
        """From a (n0, n1) phase map and an (nphase, 2) table of (sph, dev)."""
        return cls(_arrays=('phases', np.shape(phase), phase, table))

    @classmethod
    def from_fields(cls, sph, dev):
        """From per-voxel coefficient fields sph, dev of shape (n0, n1)."""
        return cls(_arrays=('fields', tuple(sph.shape), sph, dev))


class BlockDiagonalOperator3D(_BlockDiagonal,
                              AbstractStructuredOperator3D):
    """Block-diagonal operator with 3D layout (operators.pyx:667-716)."""

    def __init__(self, loc=None, _arrays=None):
        self._handle = None
        if _arrays is None:
            self._setup_from_loc(loc)
        else:
            kind, grid, a, b = _arrays
            self._init_arrays(grid, 6)
            self._setup_arrays(kind, a, b)

    @classmethod
    def from_phases(cls, phase, table):
        """From a (n0, n1, n2) phase map and an (nphase, 2) table."""
        return cls(_arrays=('phases', np.shape(phase), phase, table))

    @classmethod
    def from_fields(cls, sph, dev):
        """From per-voxel coefficient fields of shape (n0, n1, n2)."""
        return cls(_arrays=('fields', tuple(sph.shape), sph, dev))


# ---------------------------------------------------------------------------
# Block-diagonal linear operators (matrix fields)
# ---------------------------------------------------------------------------

def block_diagonal_linear_operator(a):
    """Create a block-diagonal linear operator from a 4D or 5D array of
    local matrices (operators.pyx:90-110)."""
    dim = len(a.shape) - 2
    if dim == 2:
        return BlockDiagonalLinearOperator2D(a)
    elif dim == 3:
        return BlockDiagonalLinearOperator3D(a)
    else:
        raise ValueError('number of dimensions of a must be 4 or 5 '
                         '(was {0})'.format(dim))


class _BlockDiagonalLinear:
    """y[v] = a[v] x[v] for a per-voxel (osize, isize) matrix field
    (operators.pyx:719-890).  A host array is copied to the device once; a
    device tensor is kept by reference (janus keeps a shallow copy of a).
    When the blocks are sym x sym the operator also has a jx_local handle,
    so it can be the local operator of an LSSolver."""

    def _setup(self, a):
        if a is None:
            raise ValueError('a should not be None')
        shape = tuple(int(n) for n in a.shape)
        if len(shape) != self.dim + 2:
            raise ValueError('a must be a {0}D array'.format(self.dim + 2))
        grid, m, n = shape[:self.dim], shape[self.dim], shape[self.dim + 1]
        self.init_shapes(*(grid + (n, m)))
        self.a = a
        _lib.require_gpu()
        self._a_dev = _device.to_device(a)
        self._handle = None
        sym = (self.dim * (self.dim + 1)) // 2
        if m == sym and n == sym:
            h = ctypes.c_void_p()
            _lib.check(_lib.lib().jx_local_create_matrix_field(
                self.dim, _lib.shape_array(grid), self._a_dev.data_ptr(),
                ctypes.byref(h)))
            self._handle = h

    def _run(self, transpose, x, y, oshape):
        L = _lib.lib()
        grid = self.ishape[:-1]

        def run(xd, yd):
            if self._handle is not None:
                fn = L.jx_local_apply_transpose if transpose else \
                    L.jx_local_apply
                _lib.check(fn(self._handle, xd.data_ptr(), yd.data_ptr(),
                              _lib.current_stream()))
                return
            src = xd.clone() if xd.data_ptr() == yd.data_ptr() else xd
            _lib.check(L.jx_block_linear_apply(
                self.dim, _lib.shape_array(grid), self.oshape[-1],
                self.ishape[-1], self._a_dev.data_ptr(), src.data_ptr(),
                yd.data_ptr(), 1 if transpose else 0, _lib.current_stream()))
        return _run_device(run, x, y, oshape)

    def _apply(self, x, y):
        return self._run(False, x, y, self.oshape)

    def apply_transpose(self, x, y=None):
        """y[v] = a[v]^T x[v]; x has the output shape, y the input shape
        (operators.pyx:776-800, 863-890)."""
        if tuple(np.shape(x)) != self.oshape:
            raise ValueError('invalid shape of x: expected {0}, actual {1}'
                             .format(self.oshape, tuple(np.shape(x))))
        if y is not None and tuple(np.shape(y)) != self.ishape:
            raise ValueError('invalid shape of y: expected {0}, actual {1}'
                             .format(self.ishape, tuple(np.shape(y))))
        return self._run(True, x, y, self.ishape)

    def means(self, x):
        """(<a x>, <x>) over the grid (sym x sym blocks only)."""
        xd = _device.to_device(x)
        sym = self.ishape[-1]
        tau = np.empty(sym)
        eps = np.empty(sym)
        _lib.check(_lib.lib().jx_local_means(
            self.handle, xd.data_ptr(), tau.ctypes.data, eps.ctypes.data,
            _lib.current_stream()))
        return tau, eps

    @property
    def handle(self):
        if self._handle is None:
            raise NotImplementedError(
                'only sym x sym blocks make a local operator of a solver')
        return self._handle

    @property
    def symmetric(self):
        out = ctypes.c_int()
        _lib.check(_lib.lib().jx_local_is_symmetric(self.handle,
                                                    ctypes.byref(out)))
        return bool(out.value)

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and _lib._lib is not None:
            _lib._lib.jx_local_destroy(h)
            self._handle = None


class BlockDiagonalLinearOperator2D(_BlockDiagonalLinear,
                                    AbstractStructuredOperator2D):
    """Block-diagonal linear operator with 2D layout; a has shape
    (n0, n1, osize, isize) (operators.pyx:719-800)."""

    def __init__(self, a):
        self._handle = None
        self._setup(a)


class BlockDiagonalLinearOperator3D(_BlockDiagonalLinear,
                                    AbstractStructuredOperator3D):
    """Block-diagonal linear operator with 3D layout; a has shape
    (n0, n1, n2, osize, isize) (operators.pyx:802-890)."""

    def __init__(self, a):
        self._handle = None
        self._setup(a)
