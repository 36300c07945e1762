"""Restarted GMRES on the GPU (jx_solver_solve_gmres) vs the oracle's
gmres_scheme, iterate for iterate: residual histories, iteration counts and
the solution field, with and without restarts, for isotropic phase maps and
for non-symmetric per-voxel matrices (which conjugate gradients refuse)."""

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd import microstructure as ms
from janus_amd.green import filtered, truncated, willot2015
from janus_amd.operators import BlockDiagonalOperator2D, \
    BlockDiagonalOperator3D
from janus_amd.solver import LSSolver, delta_stiffness
from oracle import janus_oracle as jo

pytestmark = pytest.mark.gpu

FACTORIES = {'truncated': truncated, 'filtered': filtered,
             'willot': willot2015}
CLS = {2: BlockDiagonalOperator2D, 3: BlockDiagonalOperator3D}


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def iso_solver(dim, n, kind, mu_i, g0):
    if dim == 2:
        phase = ms.square_inclusion(n, 0.5)
    else:
        phase = ms.spheres(n, 3.0, 0.2, seed=3)
    mats = [material.create(1.0, 0.3, dim), material.create(mu_i, 0.25, dim)]
    mat0 = material.create(g0, 0.3, dim)
    table = np.array([delta_stiffness(m, mat0) for m in mats])
    local = CLS[dim].from_phases(phase, table)
    shape = phase.shape
    green = FACTORIES[kind](mat0.green_operator(), shape, 1.0,
                            fft.create_real(shape))
    return LSSolver(green, local), table[phase, 0], table[phase, 1]


CASES = [
    # dim, n, kind, mu_i, g0, restart
    (2, 32, 'truncated', 10.0, 1.0, 20),
    (2, 33, 'filtered', 20.0, 1.0, 8),
    (2, 25, 'willot', 10.0, 5.5, 30),
    (3, 16, 'filtered', 50.0, 1.0, 12),
    (3, 15, 'truncated', 5.0, 3.0, 5),
    (3, 32, 'filtered', 10.0, 0.9, 20),
]


@pytest.mark.parametrize('dim, n, kind, mu_i, g0, restart', CASES)
def test_gmres_matches_oracle(dim, n, kind, mu_i, g0, restart):
    solver, sph, dev = iso_solver(dim, n, kind, mu_i, g0)
    E = np.random.default_rng(n).standard_normal(solver.sym)
    r = solver.solve_gmres(E, tol=1e-10, max_iter=500, restart=restart)
    assert r.status == 'converged'
    eps_ref, res_ref = jo.gmres_scheme(sph, dev, E, kind, g0, 0.3, tol=1e-10,
                                       max_iter=500, restart=restart)
    assert r.iterations == len(res_ref)
    np.testing.assert_allclose(r.residuals, res_ref, rtol=1e-8, atol=1e-15)
    assert rel_err(r.eps.cpu().numpy(), eps_ref) < 1e-12


def nonsymmetric_solver(dim, shape, kind, seed):
    """dC = isotropic inclusion contrast + a non-symmetric per-voxel
    perturbation (a matrix field)."""
    rng = np.random.default_rng(seed)
    sym = 3 if dim == 2 else 6
    phase = (ms.square_inclusion(shape[0], 0.5) if dim == 2 else
             ms.spheres(shape[0], 3.0, 0.2, seed=seed))
    mat0 = material.create(2.0, 0.3, dim)
    iso = np.array([jo.iso4_matrix(*delta_stiffness(material.create(g, 0.3,
                                                                    dim),
                                                    mat0), dim)
                    for g in (1.0, 8.0)])
    field = iso[phase] + 0.2 * rng.standard_normal(phase.shape + (sym, sym))
    local = CLS[dim].from_matrix_field(field)
    assert not local.symmetric
    green = FACTORIES[kind](mat0.green_operator(), phase.shape, 1.0,
                            fft.create_real(phase.shape))
    return LSSolver(green, local), field, mat0


@pytest.mark.parametrize('dim, n, kind, restart', [
    (2, 32, 'filtered', 10), (3, 16, 'truncated', 30), (3, 12, 'willot', 6)])
def test_gmres_nonsymmetric_matches_oracle(dim, n, kind, restart):
    S, field, mat0 = nonsymmetric_solver(dim, (n,) * dim, kind, seed=n)
    E = np.random.default_rng(9).standard_normal(S.sym)
    with pytest.raises(Exception, match='symmetric'):
        S.solve_cg(E, tol=1e-8)
    r = S.solve_gmres(E, tol=1e-10, max_iter=500, restart=restart)
    assert r.status == 'converged'
    eps_ref, res_ref = jo.gmres_scheme(None, None, E, kind, mat0.g, mat0.nu,
                                       tol=1e-10, max_iter=500,
                                       restart=restart, mats=field)
    assert r.iterations == len(res_ref)
    np.testing.assert_allclose(r.residuals, res_ref, rtol=1e-8, atol=1e-15)
    assert rel_err(r.eps.cpu().numpy(), eps_ref) < 1e-12
    # the field solves the Lippmann-Schwinger equation of that operator, to
    # the stopping rule's own norm ||.|| / (sqrt(N) ||E||)
    eps = r.eps.cpu().numpy()
    lhs = eps + jo.green_apply(jo.matrix_apply(field, eps), kind, mat0.g,
                               mat0.nu)
    true = np.linalg.norm(lhs - E) / np.sqrt(n ** dim) / np.linalg.norm(E)
    assert true < 1.5e-10 and abs(true - r.residuals[-1]) < 1e-12


def test_gmres_equals_cg_and_basic_scheme():
    """Symmetric dC, midpoint C0: the three solvers reach the same field."""
    solver, sph, dev = iso_solver(3, 32, 'filtered', 10.0, 5.5)
    E = np.array([0., 0., 0., 0., 0., 1.])
    a = solver.solve(E, tol=1e-12, max_iter=2000).eps.cpu().numpy()
    b = solver.solve_cg(E, tol=1e-12, max_iter=500).eps.cpu().numpy()
    r = solver.solve_gmres(E, tol=1e-12, max_iter=500, restart=20)
    assert r.status == 'converged'
    c = r.eps.cpu().numpy()
    assert rel_err(c, b) < 1e-10 and rel_err(c, a) < 1e-10


def test_gmres_rhs_and_stops():
    """A field right-hand side (the adjoint solve's form), the max_iter
    stop, restart = 1, and the argument checks."""
    import torch
    solver, sph, dev = iso_solver(2, 32, 'truncated', 10.0, 1.0)
    rng = np.random.default_rng(2)
    g = rng.standard_normal((32, 32, 3))
    rhs = np.asarray(solver.green.apply(g))
    E = np.array([0.3, -0.2, 1.0])
    r = solver.solve_gmres(E, tol=1e-10, max_iter=500, restart=15,
                           rhs=torch.from_numpy(rhs).cuda())
    eps_ref, res_ref = jo.gmres_scheme(sph, dev, E, 'truncated', 1.0, 0.3,
                                       tol=1e-10, max_iter=500, restart=15,
                                       rhs=rhs)
    assert r.status == 'converged' and r.iterations == len(res_ref)
    assert rel_err(r.eps.cpu().numpy(), eps_ref) < 1e-12
    r = solver.solve_gmres(E, tol=1e-14, max_iter=7, restart=3)
    eps_ref, res_ref = jo.gmres_scheme(sph, dev, E, 'truncated', 1.0, 0.3,
                                       tol=1e-14, max_iter=7, restart=3)
    assert r.status == 'max_iter' and r.iterations == 7 == len(res_ref)
    np.testing.assert_allclose(r.residuals, res_ref, rtol=1e-8)
    assert rel_err(r.eps.cpu().numpy(), eps_ref) < 1e-12
    r = solver.solve_gmres(E, tol=1e-14, max_iter=5, restart=1)
    _, res_ref = jo.gmres_scheme(sph, dev, E, 'truncated', 1.0, 0.3,
                                 tol=1e-14, max_iter=5, restart=1)
    np.testing.assert_allclose(r.residuals, res_ref, rtol=1e-8)
    with pytest.raises(ValueError):
        solver.solve_gmres(E, restart=0)
    with pytest.raises(ValueError):
        solver.solve_gmres(E, max_iter=0)


def test_effective_stiffness_methods_agree():
    """effective_stiffness by the basic scheme, CG and GMRES (midpoint C0,
    where all three converge) gives the same 6 x 6 C_eff; with C0 = the
    matrix only the Krylov methods converge, and they agree."""
    from janus_amd.solver import effective_stiffness
    solver, _, _ = iso_solver(3, 16, 'filtered', 10.0, 5.5)
    mat0 = material.create(5.5, 0.3, 3)
    Cs = [effective_stiffness(solver, mat0, tol=1e-12, max_iter=2000,
                              method=m)
          for m in ('basic', 'cg', 'gmres')]
    for C, its, sts in Cs:
        assert sts == ['converged'] * 6
    for C, _, _ in Cs[1:]:
        assert rel_err(C, Cs[0][0]) < 1e-10
    solver, _, _ = iso_solver(3, 16, 'filtered', 10.0, 1.0)
    mat0 = material.create(1.0, 0.3, 3)
    Ccg, _, s1 = effective_stiffness(solver, mat0, tol=1e-12, max_iter=500,
                                     method='cg')
    Cgm, _, s2 = effective_stiffness(solver, mat0, tol=1e-12, max_iter=500,
                                     method='gmres', restart=30)
    assert s1 == s2 == ['converged'] * 6
    assert rel_err(Cgm, Ccg) < 1e-10
    with pytest.raises(ValueError):
        effective_stiffness(solver, mat0, method='bicgstab')
