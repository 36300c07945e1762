"""Seeded periodic microstructures for the benchmark and test workloads.

The voxel rule is janus's notebooks/gen_ellipsoid.py:32-38: a voxel belongs
to an inclusion when its cell centre lies inside it.  Distances use the
periodic minimum image.  All generators take a ``seed`` for
``numpy.random.default_rng``.
"""

import math

import numpy as np


def square_inclusion(n, a=0.5, centered=True):
    """2D phase map: 1 in a square of side ceil(n a - 0.5) cells, 0 elsewhere.

    ``centered=False`` puts it at the corner, as
    sphinx/tutorials/square_basic/square_basic.py:35-37 does.
    """
    imax = int(np.ceil(n * a - 0.5))
    phase = np.zeros((n, n), dtype=np.uint8)
    o = (n - imax) // 2 if centered else 0
    phase[o:o + imax, o:o + imax] = 1
    return phase


def _stamp_ball(phase, center, radii, rot, label):
    """Set ``label`` on voxels whose centre is inside the (rotated)
    ellipsoid with semi-axes ``radii`` around ``center`` (periodic)."""
    shape = phase.shape
    dim = len(shape)
    r = max(radii)
    idx = []
    for d in range(dim):
        lo = int(math.floor(center[d] - r - 0.5))
        hi = int(math.ceil(center[d] + r + 0.5))
        idx.append(np.arange(lo, hi + 1))
    grids = np.meshgrid(*idx, indexing='ij')
    x = np.stack([g + 0.5 - center[d] for d, g in enumerate(grids)], axis=-1)
    if rot is not None:
        x = x @ rot
    q = sum((x[..., d] / radii[d]) ** 2 for d in range(dim))
    inside = q <= 1.0
    sel = tuple(np.mod(g[inside], shape[d]) for d, g in enumerate(grids))
    phase[sel] = label


def spheres(n, radius, volume_fraction, seed, dim=3):
    """Boolean model of overlapping spheres (Poisson number of uniform
    centres) with expected volume fraction ``volume_fraction``."""
    rng = np.random.default_rng(seed)
    vol = math.pi * radius ** 2 if dim == 2 else 4. / 3. * math.pi * radius ** 3
    lam = -math.log(1. - volume_fraction) * n ** dim / vol
    count = int(rng.poisson(lam))
    centers = rng.uniform(0., n, size=(count, dim))
    phase = np.zeros((n,) * dim, dtype=np.uint8)
    for c in centers:
        _stamp_ball(phase, c, (radius,) * dim, None, 1)
    return phase


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def ellipsoids(n, semi_axes, count, seed, labels=(1, 2)):
    """Randomly oriented ellipsoids with the given semi-axes (3D); the
    inclusions alternate over ``labels`` (phase 0 is the matrix)."""
    rng = np.random.default_rng(seed)
    phase = np.zeros((n, n, n), dtype=np.uint8)
    for i in range(count):
        c = rng.uniform(0., n, size=3)
        rot = _random_rotation(rng)
        _stamp_ball(phase, c, semi_axes, rot, labels[i % len(labels)])
    return phase


def smoothed_density(n, sigma, seed, dim=3):
    """Periodic Gaussian-smoothed N(0, 1) noise mapped affinely onto [0, 1]."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((n,) * dim)
    k = [np.fft.fftfreq(n) for _ in range(dim)]
    k[-1] = np.fft.rfftfreq(n)
    kk = sum(np.meshgrid(*[ki ** 2 for ki in k], indexing='ij'))
    filt = np.exp(-2. * (math.pi * sigma) ** 2 * kk)
    axes = tuple(range(dim))
    rho = np.fft.irfftn(np.fft.rfftn(noise, axes=axes) * filt, s=(n,) * dim,
                        axes=axes)
    lo, hi = rho.min(), rho.max()
    return (rho - lo) / (hi - lo)
