"""Gradients of ls_solve_matrix (per-voxel matrix local operators,
jx_adjoint_outer) against central finite differences of the oracle's solve,
for symmetric and non-symmetric matrix fields, and against the isotropic
adjoint (ls_solve) on isotropic blocks."""

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd import microstructure as ms
from janus_amd.autograd import ls_solve, ls_solve_matrix
from janus_amd.green import filtered, truncated
from janus_amd.solver import delta_stiffness
from oracle import janus_oracle as jo

pytestmark = pytest.mark.gpu

FACTORIES = {'truncated': truncated, 'filtered': filtered}


def setup(dim, n, kind, symmetric, seed):
    rng = np.random.default_rng(seed)
    sym = 3 if dim == 2 else 6
    phase = (ms.square_inclusion(n, 0.5) if dim == 2 else
             ms.spheres(n, 2.0, 0.2, seed=seed))
    mat0 = material.create(2.5, 0.3, dim)
    iso = np.array([jo.iso4_matrix(*delta_stiffness(
        material.create(g, 0.3, dim), mat0), dim) for g in (1.0, 4.0)])
    field = iso[phase] + 0.1 * rng.standard_normal(phase.shape + (sym, sym))
    if symmetric:
        field = 0.5 * (field + np.swapaxes(field, -1, -2))
    shape = phase.shape
    green = FACTORIES[kind](mat0.green_operator(), shape, 1.0,
                            fft.create_real(shape))
    return field, green, mat0, rng


@pytest.mark.parametrize('dim, n, kind, symmetric, method', [
    (2, 12, 'truncated', True, 'cg'), (2, 12, 'filtered', True, 'basic'),
    (2, 12, 'filtered', False, 'gmres'), (2, 10, 'truncated', False, 'basic'),
    (3, 6, 'filtered', False, 'gmres')])
def test_matrix_gradient_vs_finite_differences(dim, n, kind, symmetric,
                                               method):
    import torch
    field, green, mat0, rng = setup(dim, n, kind, symmetric, seed=n + dim)
    sym = field.shape[-1]
    w = rng.standard_normal(field.shape[:-1])  # loss = sum(w * eps)
    E = rng.standard_normal(sym)
    A = torch.from_numpy(field).cuda().requires_grad_()
    Et = torch.tensor(E, device='cuda', requires_grad=True)
    info = {}
    eps = ls_solve_matrix(A, Et, green, tol=1e-13, max_iter=3000,
                          method=method, info=info)
    loss = (eps * torch.from_numpy(w).cuda()).sum()
    loss.backward()
    assert info['forward'].status == 'converged'
    assert info['adjoint'].status == 'converged'
    dA = A.grad.cpu().numpy()
    dE = Et.grad.cpu().numpy()

    def oracle_loss(f, e):
        x, res = jo.gmres_scheme(None, None, e, kind, mat0.g, mat0.nu,
                                 tol=1e-14, max_iter=3000, restart=40,
                                 mats=f)
        assert res[-1] < 1e-14
        return np.sum(w * x)

    h = 1e-6
    scale = np.abs(dA).max()
    grid = field.shape[:-2]
    for k in range(4):
        vox = tuple(int(rng.integers(0, s)) for s in grid)
        i, j = (int(v) for v in rng.integers(0, sym, size=2))
        fp, fm = field.copy(), field.copy()
        fp[vox + (i, j)] += h
        fm[vox + (i, j)] -= h
        fd = (oracle_loss(fp, E) - oracle_loss(fm, E)) / (2 * h)
        assert abs(fd - dA[vox + (i, j)]) < 1e-6 * scale, (vox, i, j, fd,
                                                          dA[vox + (i, j)])
    c = int(rng.integers(0, sym))
    ep, em = E.copy(), E.copy()
    ep[c] += h
    em[c] -= h
    fd = (oracle_loss(field, ep) - oracle_loss(field, em)) / (2 * h)
    assert abs(fd - dE[c]) < 1e-6 * max(abs(fd), np.abs(dE).max())


def test_matrix_gradient_matches_isotropic_adjoint():
    """Isotropic blocks sph J + dev K given as a matrix field: the chain rule
    dsph = sum dA . J, ddev = sum dA . K reproduces ls_solve's gradient."""
    import torch
    dim, n = 3, 8
    phase = ms.spheres(n, 2.0, 0.2, seed=5)
    mat0 = material.create(2.5, 0.3, dim)
    tab = np.array([delta_stiffness(material.create(g, 0.3, dim), mat0)
                    for g in (1.0, 4.0)])
    sph, dev = tab[phase, 0], tab[phase, 1]
    J = jo.iso4_matrix(1.0, 0.0, dim)
    K = jo.iso4_matrix(0.0, 1.0, dim)
    field = sph[..., None, None] * J + dev[..., None, None] * K
    green = filtered(mat0.green_operator(), phase.shape, 1.0,
                     fft.create_real(phase.shape))
    rng = np.random.default_rng(1)
    w = torch.from_numpy(rng.standard_normal(phase.shape + (6,))).cuda()
    E = torch.tensor(rng.standard_normal(6), device='cuda')
    A = torch.from_numpy(field).cuda().requires_grad_()
    (ls_solve_matrix(A, E, green, tol=1e-13, max_iter=2000) * w).sum() \
        .backward()
    s = torch.from_numpy(sph).cuda().requires_grad_()
    d = torch.from_numpy(dev).cuda().requires_grad_()
    (ls_solve(s, d, E, green, tol=1e-13, max_iter=2000) * w).sum().backward()
    dA = A.grad.cpu().numpy()
    dsph = np.einsum('...ij,ij->...', dA, J)
    ddev = np.einsum('...ij,ij->...', dA, K)
    scale = max(np.abs(s.grad.cpu().numpy()).max(),
                np.abs(d.grad.cpu().numpy()).max())
    np.testing.assert_allclose(dsph, s.grad.cpu().numpy(), rtol=0,
                               atol=1e-9 * scale)
    np.testing.assert_allclose(ddev, d.grad.cpu().numpy(), rtol=0,
                               atol=1e-9 * scale)


def test_cg_refuses_nonsymmetric_matrix_field():
    import torch
    field, green, mat0, rng = setup(2, 8, 'truncated', False, seed=3)
    A = torch.from_numpy(field).cuda().requires_grad_()
    E = torch.tensor([1., 0., 0.], device='cuda')
    with pytest.raises(ValueError, match='symmetric'):
        ls_solve_matrix(A, E, green, method='cg')
