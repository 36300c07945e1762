"""Batched solves (jx_batch_*) vs the same solves run one after another.

Every entry of a batch runs the kernels and arithmetic of
LSSolver.solve, so its iterates, residual history and iteration count must
be bit-for-bit those of the sequential solve -- whatever the grouping,
width, graph capture or autotuning.  The sequential solves themselves are
pinned against the oracle in test_gpu_solver.py; one batched entry is also
checked against the oracle here."""

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd import microstructure as ms
from janus_amd.batch import BatchSolver, effective_stiffness_batched
from janus_amd.green import filtered, truncated, willot2015
from janus_amd.operators import BlockDiagonalOperator2D, \
    BlockDiagonalOperator3D
from janus_amd.solver import LSSolver, delta_stiffness, effective_stiffness
from oracle import janus_oracle as jo

pytestmark = pytest.mark.gpu

FACTORIES = {'truncated': truncated, 'filtered': filtered,
             'willot': willot2015}


def make_solver(phase, contrast, kind, nu_i=0.25, fused=True):
    dim = phase.ndim
    mat_m = material.create(1.0, 0.3, dim)
    mat_i = material.create(contrast, nu_i, dim)
    mat0 = material.create(0.5 * (1.0 + contrast), 0.3, dim)
    table = np.array([delta_stiffness(m, mat0) for m in (mat_m, mat_i)])
    cls = BlockDiagonalOperator2D if dim == 2 else BlockDiagonalOperator3D
    local = cls.from_phases(phase, table)
    n = phase.shape
    green = FACTORIES[kind](mat0.green_operator(), n, 1.0, fft.create_real(n))
    S = LSSolver(green, local)
    S.fused = fused
    return S, table, mat0


def field_solver(shape, seed):
    """Per-voxel SIMP-like moduli (field local operator)."""
    import torch
    dim = len(shape)
    rho = np.random.default_rng(seed).random(shape)
    g = 1.0 + 4.0 * rho ** 3
    mat0 = material.create(3.0, 0.3, dim)
    k = 2. / 3. * 1.3 / 0.4 * g
    sph = torch.from_numpy(dim * (k - mat0.k)).cuda()
    dev = torch.from_numpy(2. * (g - mat0.g)).cuda()
    local = BlockDiagonalOperator3D.from_fields(sph, dev)
    green = filtered(mat0.green_operator(), shape, 1.0, fft.create_real(shape))
    S = LSSolver(green, local)
    S._keep = (sph, dev)
    return S


def unit(sym, j):
    E = np.zeros(sym)
    E[j] = 1.0
    return E


@pytest.fixture(scope='module')
def cases():
    """(solver, E, tol, max_iter, rhs) for a mix of groups: two 32^3
    filtered microstructures sharing a group, a truncated one, a five-pass
    one, a field local operator on 16x32x64, a 2D grid, a grid without
    register FFT lengths (per-entry launches), an entry with a rhs field and
    one stopped by max_iter."""
    import torch
    out = []
    sA, _, _ = make_solver(ms.spheres(32, 4.0, 0.2, seed=3), 10.0, 'filtered')
    sA2, _, _ = make_solver(ms.spheres(32, 3.0, 0.3, seed=4), 4.0, 'filtered')
    sB, _, _ = make_solver(ms.spheres(32, 4.0, 0.15, seed=5), 6.0, 'truncated')
    sF, _, _ = make_solver(ms.spheres(32, 4.0, 0.2, seed=6), 5.0, 'filtered',
                           fused=False)
    sC = field_solver((16, 32, 64), 7)
    s2, _, _ = make_solver(ms.spheres(80, 6.0, 0.2, seed=8, dim=2)[:, :64]
                           .copy(), 8.0, 'filtered')
    sG, _, _ = make_solver(ms.spheres(30, 3.0, 0.2, seed=9)[:, :20, :24]
                           .copy(), 3.0, 'truncated')
    for j in range(6):
        out.append((sA, unit(6, j), 1e-10, 1000, None))
    for j in (0, 5):
        out.append((sA2, unit(6, j), 1e-10, 1000, None))
    for j in (1, 3):
        out.append((sB, unit(6, j), 1e-10, 1000, None))
    for j in (2, 4):
        out.append((sF, unit(6, j), 1e-10, 1000, None))
    for j in (0, 5, 3):
        out.append((sC, unit(6, j), 1e-10, 1000, None))
    for j in range(3):
        out.append((s2, unit(3, j), 1e-10, 1000, None))
    for j in (0, 4):
        out.append((sG, unit(6, j), 1e-10, 1000, None))
    rhs = torch.from_numpy(np.random.default_rng(10).standard_normal(
        sA.shape) * 0.1).cuda()
    out.append((sA, unit(6, 2), 1e-10, 1000, rhs))
    out.append((sA2, unit(6, 1), 1e-10, 5, None))
    return out


@pytest.fixture(scope='module')
def sequential(cases):
    res = []
    for S, E, tol, mi, rhs in cases:
        r = S.solve(E, tol=tol, max_iter=mi, rhs=rhs)
        res.append((r.eps.cpu().numpy(), r.residuals.copy(), r.iterations,
                    r.status))
    return res


SCHEDULES = [
    dict(iters_per_launch=8, graph=True, autotune=False),
    dict(iters_per_launch=3, graph=False, autotune=False),
    dict(iters_per_launch=5, graph=True, autotune=True),
    # groups forked onto concurrent streams (captured and enqueued)
    dict(iters_per_launch=8, graph=True, autotune=False, streams=4),
    dict(iters_per_launch=3, graph=False, autotune=False, streams=3),
    dict(iters_per_launch=5, graph=False, autotune=True, streams=8),
]


@pytest.mark.parametrize('schedule', SCHEDULES)
def test_batch_matches_sequential_bitwise(cases, sequential, schedule):
    batch = BatchSolver(**schedule)
    for S, E, tol, mi, rhs in cases:
        batch.add(S, E, tol=tol, max_iter=mi, rhs=rhs)
    res = batch.run()
    for (eps, hist, its, status), r in zip(sequential, res):
        assert r.iterations == its
        assert r.status == status
        np.testing.assert_array_equal(r.residuals, hist)
        np.testing.assert_array_equal(r.eps.cpu().numpy(), eps)
    assert res[-1].status == 'max_iter' and res[-1].iterations == 5
    ms_, replays, captures = batch.stats()
    assert replays >= 1
    if schedule['graph']:
        assert captures >= 1


def test_groups_and_widths(cases):
    batch = BatchSolver()
    for S, E, tol, mi, rhs in cases:
        batch.add(S, E, tol=tol, max_iter=mi, rhs=rhs)
    groups = batch.groups()
    # sA + sA2 (8 + 1 max_iter entry), sB, sF, sC, 2D, generic, rhs
    sizes = {g[0]: [] for g in groups}
    for shape, size, width, batched in groups:
        sizes[shape].append((size, batched))
    assert sorted(sizes[(32, 32, 32)]) == [(1, True), (2, True), (2, True),
                                           (9, True)]
    assert sizes[(30, 20, 24)] == [(2, False)]
    assert sizes[(16, 32, 64)] == [(3, True)]
    batch.set_width(0, 2)
    batch.run()
    assert batch.groups()[0][2] == 2


def test_widths_give_identical_results(cases, sequential):
    batch = BatchSolver(iters_per_launch=4)
    for S, E, tol, mi, rhs in cases:
        batch.add(S, E, tol=tol, max_iter=mi, rhs=rhs)
    for g in range(len(batch.groups())):
        batch.set_width(g, 1 + g % 3)
    res = batch.run()
    for (eps, hist, its, status), r in zip(sequential, res):
        assert r.iterations == its
        np.testing.assert_array_equal(r.eps.cpu().numpy(), eps)


def test_batch_rerun_restarts_from_eps(cases, sequential):
    """A second run continues from the last iterates: one more iteration
    below tol each."""
    S, E, tol, mi, rhs = cases[0]
    batch = BatchSolver()
    batch.add(S, E, tol=tol, max_iter=mi)
    r1 = batch.run()[0]
    np.testing.assert_array_equal(r1.eps.cpu().numpy(), sequential[0][0])
    r2 = batch.run()[0]
    assert r2.status == 'converged' and r2.iterations == 1


def test_batched_entry_matches_oracle():
    phase = ms.spheres(32, 4.0, 0.2, seed=11)
    S, table, mat0 = make_solver(phase, 8.0, 'filtered')
    batch = BatchSolver()
    E = np.array([0.1, -0.2, 0.05, 0.3, 0.0, 1.0])
    batch.add(S, E, tol=1e-10, max_iter=1000)
    r = batch.run()[0]
    eps_ref, res_ref = jo.basic_scheme(table[phase, 0], table[phase, 1], E,
                                       'filtered', mat0.g, mat0.nu, tol=1e-10,
                                       max_iter=1000)
    assert r.iterations == len(res_ref)
    np.testing.assert_allclose(r.residuals, res_ref, rtol=1e-9, atol=1e-15)
    eps = r.eps.cpu().numpy()
    assert np.abs(eps - eps_ref).max() / np.abs(eps_ref).max() < 1e-12


def test_effective_stiffness_batched_matches_sequential():
    solvers, mats0 = [], []
    for seed, n, c in ((1, 32, 5.0), (2, 16, 20.0), (3, 32, 2.0)):
        S, _, m0 = make_solver(ms.spheres(n, 3.0, 0.2, seed=seed), c,
                               'filtered')
        solvers.append(S)
        mats0.append(m0)
    Cs, its, sts, batch = effective_stiffness_batched(
        solvers, mats0, tol=1e-10, max_iter=2000, autotune=True)
    for S, m0, C, it in zip(solvers, mats0, Cs, its):
        Cseq, its_seq, sts_seq = effective_stiffness(S, m0, tol=1e-10,
                                                     max_iter=2000)
        assert it == its_seq
        np.testing.assert_array_equal(C, Cseq)
        # symmetric to the convergence level
        assert np.abs(C - C.T).max() < 1e-8 * np.abs(C).max()
