"""Seeded periodic microstructures for the benchmark and test workloads.

The voxel rule is janus's notebooks/gen_ellipsoid.py:32-38: a voxel belongs
to an inclusion when its cell centre lies inside it.  Distances use the
periodic minimum image.  All generators take a ``seed`` for
``numpy.random.default_rng``.

The workloads of the measured configurations are built from them:

* :func:`spheres` -- the 128^3 two-phase stiffness configuration and the
  256^3 benchmark grid (radius scaled with the grid);
* :func:`batched_cases` -- the batched configuration's 32 (grid, volume
  fraction, contrast, seed) draws;
* :func:`ellipsoids` / :func:`ellipsoids_device` -- the 512^3 three-phase
  configuration (the device version draws the same numbers);
* :func:`smoothed_density` -- the differentiable SIMP configuration.
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


def batched_cases(seed=3, count=32):
    """The batched configuration's microstructures: ``count`` draws of
    (n, volume fraction, contrast, seed) with n ~ U{64, 80, 96, 128, 160},
    vf ~ U[0.1, 0.35], inclusion/matrix contrast ~ logU[2, 100]; each is
    built with ``spheres(n, 8, vf, seed)``."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice([64, 80, 96, 128, 160]))
        vf = float(rng.uniform(0.1, 0.35))
        c = float(np.exp(rng.uniform(np.log(2.0), np.log(100.0))))
        out.append((n, vf, c, 1000 + i))
    return out


def ellipsoids_device(n, semi_axes, count, seed, labels=(1, 2)):
    """:func:`ellipsoids` evaluated on the GPU (same draws, same voxel rule;
    the numpy version takes minutes at 512^3).  Returns a uint8 CUDA
    tensor."""
    import torch
    rng = np.random.default_rng(seed)
    phase = torch.zeros((n, n, n), dtype=torch.uint8, device='cuda')
    r = max(semi_axes)
    ax = torch.tensor(semi_axes, dtype=torch.float64, device='cuda')
    for i in range(count):
        c = rng.uniform(0., n, size=3)
        rot = torch.from_numpy(_random_rotation(rng)).cuda()
        idx = [torch.arange(int(math.floor(c[d] - r - 0.5)),
                            int(math.ceil(c[d] + r + 0.5)) + 1, device='cuda')
               for d in range(3)]
        g = torch.meshgrid(*idx, indexing='ij')
        x = torch.stack([g[d].double() + 0.5 - c[d] for d in range(3)],
                        dim=-1) @ rot
        inside = ((x / ax) ** 2).sum(dim=-1) <= 1.0
        sel = tuple(torch.remainder(g[d][inside], n) for d in range(3))
        phase[sel] = labels[i % len(labels)]
    return phase


def large_three_phase(n=512, seed=2014, device=True):
    """The 512^3 three-phase configuration's phase map (scaled to n):
    randomly oriented ellipsoids with semi-axes (8, 8, 32) n / 512
    (aspect ratio 1:1:4), a Poisson-free count giving an expected
    inclusion volume fraction of 0.3, alternating over phases 1 and 2;
    phase 0 is the matrix.  Returns (phase, count, semi_axes)."""
    vol = 4. / 3. * math.pi * 8 * 8 * 32 * (n / 512.) ** 3
    count = int(-math.log(1 - 0.3) * n ** 3 / vol)
    axes = tuple(a * n / 512. for a in (8., 8., 32.))
    gen = ellipsoids_device if device else ellipsoids
    return gen(n, axes, count, seed), count, axes
