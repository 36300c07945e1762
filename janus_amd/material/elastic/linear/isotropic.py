"""Isotropic, linear elastic materials (janus.material.elastic.linear.isotropic).

``dim == 2`` is plane strain, ``dim == 3`` classical 3D elasticity.  The
continuous Green operator for strains Gamma0(k) is evaluated by the host code
of ``libjanus_amd.so`` -- the same routine the GPU kernels inline -- following
janus/material/elastic/linear/isotropic.pyx:133-156, 173-200 and 242-308.
"""

import ctypes

import numpy as np

from .... import _lib


def create(g, nu, dim=3):
    """Create a new isotropic, linear and elastic material
    (isotropic.pyx:22-30)."""
    return IsotropicLinearElasticMaterial(g, nu, dim)


def poisson_from_bulk_and_shear_moduli(k, g, dim=3):
    """Poisson ratio from the bulk and shear moduli (isotropic.pyx:33-46)."""
    if dim == 2:
        return (k - g) / (2. * k)
    elif dim == 3:
        return (3 * k - 2 * g) / (2 * (3 * k + g))
    else:
        raise ValueError('dim must be 2 or 3 (was {0})'.format(dim))


class IsotropicLinearElasticMaterial:
    """Isotropic, linear and elastic material (isotropic.pyx:49-114).

    Attributes ``g``, ``nu``, ``dim`` and ``k`` are read-only.
    """

    __slots__ = ('_g', '_nu', '_dim', '_k')

    def __init__(self, g, nu, dim=3):
        g = float(g)
        nu = float(nu)
        if g < 0.0:
            raise ValueError('g must be >= 0 (was {0})'.format(g))
        if dim == 2 and nu > 0.5:
            raise ValueError('nu must be <= 1/2 (was {0})'.format(nu))
        if dim == 3 and (nu < -1.0 or nu > 0.5):
            raise ValueError('nu must be >= -1 and <= 1/2 (was {0})'
                             .format(nu))
        if dim != 2 and dim != 3:
            raise ValueError('dim must be 2 or 3 (was {0})'.format(dim))
        self._g = g
        self._nu = nu
        self._dim = int(dim)
        if dim == 2:
            self._k = g / (1. - 2. * nu)
        else:
            self._k = 2. / 3. * (1. + nu) / (1. - 2. * nu) * g

    @property
    def g(self):
        return self._g

    @property
    def nu(self):
        return self._nu

    @property
    def dim(self):
        return self._dim

    @property
    def k(self):
        return self._k

    def __repr__(self):
        return ('IsotropicLinearElasticMaterial'
                '(g={0}, nu={1}, dim={2})').format(self.g, self.nu, self.dim)

    def green_operator(self):
        """Return the periodic Green operator for strains of this material."""
        if self.dim == 2:
            return _GreenOperatorForStrains2D(self)
        return _GreenOperatorForStrains3D(self)


class _GreenOperatorForStrains:
    """Gamma0(k) of an isotropic reference material, Mandel notation."""

    def __init__(self, mat):
        sym = (mat.dim * (mat.dim + 1)) // 2
        self._isize = sym
        self._osize = sym
        self._dim = mat.dim
        self._mat = mat
        self._matrix = np.zeros((sym, sym), dtype=np.float64)

    @property
    def isize(self):
        return self._isize

    @property
    def osize(self):
        return self._osize

    @property
    def dim(self):
        return self._dim

    @property
    def mat(self):
        return self._mat

    def __repr__(self):
        return 'Green Operator({0})'.format(self._mat)

    def set_frequency(self, k):
        """Set the current wave vector (green.pyx:63-77)."""
        k = np.ascontiguousarray(k, dtype=np.float64)
        if k.shape != (self._dim,):
            raise ValueError('invalid shape: expected ({0},), actual {1}'
                             .format(self._dim, k.shape))
        L = _lib.lib()
        _lib.check(L.jx_gamma0_at(self._dim, self._mat.g, self._mat.nu,
                                  k.ctypes.data, self._matrix.ctypes.data))

    def to_memoryview(self, out=None):
        """Matrix of Gamma0 at the current frequency (osize, isize)."""
        if out is None:
            return self._matrix.copy()
        if tuple(out.shape) != (self._osize, self._isize):
            raise ValueError('invalid shape: expected {0}, actual {1}'
                             .format((self._osize, self._isize), out.shape))
        out[...] = self._matrix
        return out.view()

    def apply(self, x, y=None):
        """y = Gamma0(k) x at the current frequency (isotropic.pyx:202-211,
        310-331); y may be x."""
        x = np.asarray(x)
        if x.shape != (self._isize,):
            raise ValueError('invalid shape: expected ({0},), actual {1}'
                             .format(self._isize, x.shape))
        r = self._matrix.dot(x)
        if y is None:
            return r
        if tuple(y.shape) != (self._osize,):
            raise ValueError('invalid shape: expected ({0},), actual {1}'
                             .format(self._osize, y.shape))
        y[...] = r
        return y.view()


class _GreenOperatorForStrains2D(_GreenOperatorForStrains):
    def __init__(self, mat):
        if mat.dim != 2:
            raise ValueError('plane strain material expected')
        super().__init__(mat)


class _GreenOperatorForStrains3D(_GreenOperatorForStrains):
    def __init__(self, mat):
        if mat.dim != 3:
            raise ValueError('3D material expected')
        super().__init__(mat)
