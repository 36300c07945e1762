"""GPU path at the configurations' full 3D sizes, through checks that do not
need the CPU oracle to run at that size, plus the oracle itself where it
finishes in seconds:

* plane waves on 256^3 and 512^3 grids: a field tau cos(2 pi b.x / n) is
  mapped by the discrete Green operator to (Gamma-hat(b) tau) cos(...),
  with Gamma-hat(b) taken from the host per-frequency API (janus's
  set_frequency / to_memoryview, green.pyx:146-180, 312-351), for the
  truncated, filtered and Willot operators;
* the 256^3 benchmark workload (spheres r = 16, contrast 50, filtered):
  the first two iterates against the oracle's basic scheme
  (square_basic.py:83-92), 1e-12 relative;
* the 128^3 stiffness configuration (spheres r = 8, contrast 50, filtered,
  midpoint reference medium): the first iterates against the oracle, and
  the converged field through size-independent properties -- the
  Lippmann-Schwinger equation eps + Gamma0[dC : eps] = E evaluated with the
  separate Green and block-diagonal operators (not the solver's fused
  kernels), the mean strain <eps> = E, and the fused and five-pass
  schedules agreeing iterate for iterate.
"""

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd import microstructure as ms
from janus_amd.green import filtered, truncated, willot2015
from janus_amd.operators import BlockDiagonalOperator3D
from janus_amd.solver import LSSolver, delta_stiffness
from oracle import janus_oracle as jo

pytestmark = pytest.mark.gpu

FACTORIES = {'truncated': truncated, 'filtered': filtered,
             'willot': willot2015}

# modes with no Nyquist component (there Gamma-hat(-b) and Gamma-hat(b)
# differ for the truncated operator, and a cosine is not an eigenfunction)
MODES = {256: [(3, 250, 7), (129, 5, 100), (0, 0, 1), (255, 1, 127),
               (40, 128 + 3, 0)],
         512: [(3, 500, 7), (300, 5, 200), (1, 0, 0), (511, 257, 255)]}


def _plane_wave(n, b):
    import torch
    theta = None
    for axis in range(3):
        i = torch.arange(n, dtype=torch.int64, device='cuda')
        t = (2 * np.pi / n) * ((b[axis] * i) % n).to(torch.float64)
        shape = [1, 1, 1]
        shape[axis] = n
        t = t.reshape(shape)
        theta = t if theta is None else theta + t
    return torch.cos(theta)


@pytest.mark.parametrize('n', [256, 512])
@pytest.mark.parametrize('kind', ['truncated', 'filtered', 'willot'])
def test_plane_waves_full_grid(n, kind):
    import torch
    grid = (n, n, n)
    mat0 = material.create(1.3, 0.28, 3)
    green = FACTORIES[kind](mat0.green_operator(), grid, 1.0,
                           fft.create_real(grid))
    rng = np.random.default_rng(n)
    x = torch.empty(grid + (6,), dtype=torch.float64, device='cuda')
    y = torch.empty_like(x)
    for b in MODES[n]:
        tau = rng.standard_normal(6)
        c = _plane_wave(n, b)
        x.copy_(c[..., None] * torch.from_numpy(tau).cuda())
        green.apply(x, y)
        green.set_frequency(np.array(b, dtype=np.int32))
        expected_amp = green.to_memoryview().dot(tau)
        expected = c[..., None] * torch.from_numpy(expected_amp).cuda()
        err = float((y - expected).abs().max()) / np.abs(expected_amp).max()
        print('plane wave n=%d %s b=%s: max err / max|y| = %.3g'
              % (n, kind, b, err), flush=True)
        assert err < 1e-13, (b, err)
        del c, expected
    del x, y
    torch.cuda.empty_cache()


def _two_phase(n, radius, kind):
    phase = ms.spheres(n, radius, 0.2, seed=2014)
    mat_m = material.create(1.0, 0.3, 3)
    mat_i = material.create(50.0, 0.25, 3)
    mat0 = material.create(25.5, 0.3, 3)
    table = np.array([delta_stiffness(mat_m, mat0),
                      delta_stiffness(mat_i, mat0)])
    green = FACTORIES[kind](mat0.green_operator(), (n,) * 3, 1.0,
                            fft.create_real((n,) * 3))
    local = BlockDiagonalOperator3D.from_phases(phase, table)
    return LSSolver(green, local), green, local, phase, table, mat0


def _iterates_vs_oracle(solver, phase, table, mat0, E, iters):
    r = solver.solve(E, tol=0.0, max_iter=iters)
    eps_ref, res_ref = jo.basic_scheme(table[phase, 0], table[phase, 1], E,
                                       'filtered', mat0.g, mat0.nu, tol=0.0,
                                       max_iter=iters)
    eps = r.eps.cpu().numpy()
    err = np.abs(eps - eps_ref).max() / np.abs(eps_ref).max()
    print('%d iterates at %s: max rel err %.3g' % (iters, phase.shape, err),
          flush=True)
    assert r.iterations == iters
    np.testing.assert_allclose(r.residuals, res_ref, rtol=1e-10)
    assert err < 1e-12, err


def test_benchmark_workload_256_iterates_match_oracle():
    solver, _, _, phase, table, mat0 = _two_phase(256, 16.0, 'filtered')
    E = np.array([0., 0., 0., 0., 0., 1.])
    _iterates_vs_oracle(solver, phase, table, mat0, E, 2)


def test_stiffness_128_properties():
    import torch
    n = 128
    solver, green, local, phase, table, mat0 = _two_phase(n, 8.0, 'filtered')
    E = np.array([0., 0., 0., 0., 0., 1.])
    _iterates_vs_oracle(solver, phase, table, mat0, E, 3)

    r = solver.solve(E, tol=1e-10, max_iter=2000)
    assert r.status == 'converged'
    eps = r.eps
    Ed = torch.from_numpy(E).cuda()
    mean = eps.reshape(-1, 6).mean(dim=0)
    assert float((mean - Ed).abs().max()) < 1e-13
    # the Lippmann-Schwinger equation with the separate operators
    tau = torch.empty_like(eps)
    local.apply(eps, tau)
    g = torch.empty_like(eps)
    green.apply(tau, g)
    lhs = eps + g - Ed
    res = float(torch.linalg.vector_norm(lhs)) / (np.sqrt(n ** 3) *
                                                  np.linalg.norm(E))
    print('128^3: %d iterations, LS equation residual %.3g'
          % (r.iterations, res), flush=True)
    assert res < 2e-10
    # the five-pass schedule gives the same iterates to rounding
    solver.fused = False
    r5 = solver.solve(E, tol=1e-10, max_iter=2000)
    solver.fused = True
    assert r5.iterations == r.iterations
    d = float((r5.eps - eps).abs().max()) / float(eps.abs().max())
    assert d < 1e-12, d
