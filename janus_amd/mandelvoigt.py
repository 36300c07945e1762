"""Mandel notation of second-rank symmetric tensors (janus/mandelvoigt.py).

2D order (xx, yy, sqrt2 xy); 3D order (xx, yy, zz, sqrt2 yz, sqrt2 zx,
sqrt2 xy) -- janus/mandelvoigt.py:26-40, sphinx/mandel.rst:29-32.
"""

import numpy as np

SQRT_2 = np.sqrt(2.)

_PAIRS = {2: [(0, 0), (1, 1), (0, 1)],
          3: [(0, 0), (1, 1), (2, 2), (1, 2), (2, 0), (0, 1)]}


class MandelVoigt:
    """Index maps between Mandel indices and tensor multi-indices."""

    def __init__(self, dim):
        if dim not in _PAIRS:
            raise ValueError('dim must be 2 or 3 (was {0})'.format(dim))
        self.dim = dim
        self.sym = (dim * (dim + 1)) // 2
        self._pairs = _PAIRS[dim]
        self._index = np.empty((dim, dim), dtype=int)
        for ij, (i, j) in enumerate(self._pairs):
            self._index[i, j] = ij
            self._index[j, i] = ij

    def _check(self, i, n):
        if i < 0 or i >= n:
            raise ValueError('index must be >= 0 and < {0} (was {1})'
                             .format(n, i))

    def unravel_index(self, index):
        self._check(index, self.sym)
        return self._pairs[index]

    def ravel_multi_index(self, i, j):
        self._check(i, self.dim)
        self._check(j, self.dim)
        return int(self._index[i, j])

    def as_array(self, coeff, *args):
        """Mandel matrix of the fourth-rank tensor coeff(i, j, k, l, *args)
        with minor symmetries (factor sqrt2 per shear index)."""
        a = np.empty((self.sym, self.sym), dtype=np.float64)
        for ij in range(self.sym):
            i, j = self._pairs[ij]
            for kl in range(self.sym):
                k, l = self._pairs[kl]
                v = coeff(i, j, k, l, *args)
                if ij >= self.dim and kl >= self.dim:
                    v *= 2.
                elif ij >= self.dim or kl >= self.dim:
                    v *= SQRT_2
                a[ij, kl] = v
        return a
