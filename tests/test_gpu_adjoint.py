"""Implicit-differentiation adjoint on the GPU.

* gradients with respect to the per-voxel (sph, dev) of C - C0 and to E
  against the oracle's adjoint (oracle/janus_oracle.py, adjoint_gradient)
  on a small grid;
* dL/drho of the SIMP density problem (mu(rho) = 1 + 19 rho^3, nu = 0.3,
  L = (C_1212 - 3)^2 under E = e_xy) against central finite differences
  (step 1e-6) at seeded voxels, with both solves converged to 1e-13;
* effective_stress (no adjoint solve for losses of <sigma>) against the
  ls_solve + mean_stress gradient and the oracle's forward-field gradient,
  with dL/d<sigma> parallel to E (no second solve) and not (one forward
  solve under dL/d<sigma>).
"""

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd import microstructure as ms
from janus_amd.autograd import effective_stress, iso_delta, ls_solve, \
    mean_stress
from janus_amd.green import filtered, truncated
from oracle import janus_oracle as jo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('method', ['basic', 'cg'])
@pytest.mark.parametrize('dim, n, kind', [(2, 24, 'filtered'),
                                          (3, 12, 'truncated'),
                                          (3, 10, 'filtered')])
def test_gradient_vs_oracle(dim, n, kind, method):
    import torch
    rng = np.random.default_rng(dim * 100 + n)
    g0, nu0 = 5.0, 0.3
    mat0 = material.create(g0, nu0, dim)
    shape = (n,) * dim
    sym = 3 if dim == 2 else 6
    sph = rng.uniform(-3., 3., shape)
    dev = rng.uniform(-3., 3., shape)
    E = rng.standard_normal(sym)
    w = rng.standard_normal(shape + (sym,))
    factory = filtered if kind == 'filtered' else truncated
    green = factory(mat0.green_operator(), shape, 1.0, fft.create_real(shape))
    sph_t = torch.tensor(sph, device='cuda', requires_grad=True)
    dev_t = torch.tensor(dev, device='cuda', requires_grad=True)
    E_t = torch.tensor(E, device='cuda', requires_grad=True)
    eps = ls_solve(sph_t, dev_t, E_t, green, tol=1e-14, max_iter=3000,
                   method=method)
    loss = (torch.from_numpy(w).cuda() * eps).sum()
    loss.backward()
    eps_ref, _ = jo.basic_scheme(sph, dev, E, kind, g0, nu0, tol=1e-14,
                                 max_iter=3000)
    dsph, ddev, mu, _ = jo.adjoint_gradient(sph, dev, eps_ref, w, kind, g0,
                                            nu0, tol=1e-14, max_iter=3000)
    np.testing.assert_allclose(sph_t.grad.cpu().numpy(), dsph, rtol=0,
                               atol=1e-10 * np.abs(dsph).max())
    np.testing.assert_allclose(dev_t.grad.cpu().numpy(), ddev, rtol=0,
                               atol=1e-10 * np.abs(ddev).max())
    lam = w - jo.iso4_apply(sph, dev, mu)
    dE = lam.reshape(-1, sym).sum(axis=0)
    np.testing.assert_allclose(E_t.grad.cpu().numpy(), dE, rtol=1e-9,
                               atol=1e-9 * np.abs(dE).max())


def simp_loss(rho, green, mat0, E, tol, max_iter=4000, info=None,
              method='basic'):
    import torch
    g = 1.0 + rho ** 3 * (20.0 - 1.0)
    nu = torch.full_like(rho, 0.3)
    sph, dev = iso_delta(g, nu, mat0, 3)
    eps = ls_solve(sph, dev, E, green, tol=tol, max_iter=max_iter, info=info,
                   method=method)
    sigma = mean_stress(sph, dev, eps, E, mat0)
    c1212 = 0.5 * sigma[5] / E[5]
    return (c1212 - 3.0) ** 2


@pytest.mark.parametrize('method', ['basic', 'cg'])
def test_simp_gradient_finite_differences(method):
    import torch
    n = 24
    rho_np = ms.smoothed_density(n, 2.0, seed=4)
    mat0 = material.create(10.5, 0.3, 3)
    green = filtered(mat0.green_operator(), (n,) * 3, 1.0,
                     fft.create_real((n,) * 3))
    E = torch.zeros(6, dtype=torch.float64, device='cuda')
    E[5] = 1.0
    rho = torch.tensor(rho_np, device='cuda', requires_grad=True)
    info = {}
    loss = simp_loss(rho, green, mat0, E, 1e-14, info=info, method=method)
    loss.backward()
    assert info['forward'].status == 'converged'
    assert info['adjoint'].status == 'converged'
    grad = rho.grad.cpu().numpy()
    # both finite-difference solves run the same, fixed number of
    # iterations (past 1e-14) so that their truncation errors cancel
    iters = info['forward'].iterations + 50
    rng = np.random.default_rng(20)
    voxels = [tuple(rng.integers(0, n, 3)) for _ in range(20)]
    h = 1e-6
    errs = []
    for v in voxels:
        rp = rho_np.copy()
        rp[v] += h
        rm = rho_np.copy()
        rm[v] -= h
        with torch.no_grad():
            lp = simp_loss(torch.tensor(rp, device='cuda'), green, mat0, E,
                           0.0, iters, method=method).item()
            lm = simp_loss(torch.tensor(rm, device='cuda'), green, mat0, E,
                           0.0, iters, method=method).item()
        fd = (lp - lm) / (2 * h)
        errs.append(abs(fd - grad[v]))
    scale = np.abs(grad).max()
    assert max(errs) < 1e-5 * scale, (max(errs), scale)


@pytest.mark.parametrize('method', ['basic', 'cg'])
@pytest.mark.parametrize('parallel', [True, False])
def test_effective_stress_gradient(method, parallel):
    import torch
    n = 14
    rng = np.random.default_rng(31)
    mat0 = material.create(5.0, 0.3, 3)
    shape = (n,) * 3
    sph = rng.uniform(-3., 3., shape)
    dev = rng.uniform(-3., 3., shape)
    E = rng.standard_normal(6)
    w = 1.7 * E if parallel else rng.standard_normal(6)
    green = filtered(mat0.green_operator(), shape, 1.0, fft.create_real(shape))

    def run(fn):
        sph_t = torch.tensor(sph, device='cuda', requires_grad=True)
        dev_t = torch.tensor(dev, device='cuda', requires_grad=True)
        E_t = torch.tensor(E, device='cuda', requires_grad=True)
        info = {}
        sigma = fn(sph_t, dev_t, E_t, info)
        loss = (torch.from_numpy(w).cuda() * sigma).sum()
        if not parallel:
            loss = loss + 0.3 * sigma[5] ** 2
        loss.backward()
        return (sigma.detach().cpu().numpy(), sph_t.grad.cpu().numpy(),
                dev_t.grad.cpu().numpy(), E_t.grad.cpu().numpy(), info)

    fast = run(lambda s, d, e, info: effective_stress(
        s, d, e, green, mat0, tol=1e-14, max_iter=3000, info=info,
        method=method))
    full = run(lambda s, d, e, info: mean_stress(
        s, d, ls_solve(s, d, e, green, tol=1e-14, max_iter=3000, info=info,
                       method=method), e, mat0))
    assert full[4]['adjoint'] is not None
    np.testing.assert_allclose(fast[0], full[0], rtol=1e-12, atol=1e-14)
    for a, b in zip(fast[1:4], full[1:4]):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max())
    # dL/d<sigma> = w (parallel to E: no solve) or w + 0.6 <sigma>_5 e_5
    assert (fast[4]['adjoint'] is None) == parallel
    # against the oracle's forward-field formula
    gbar = w.copy()
    if not parallel:
        gbar[5] += 0.6 * fast[0][5]
    eps = jo.basic_scheme(sph, dev, E, 'filtered', 5.0, 0.3, tol=1e-14,
                          max_iter=3000)[0]
    eps_g = jo.basic_scheme(sph, dev, gbar, 'filtered', 5.0, 0.3,
                            tol=1e-14, max_iter=3000)[0]
    dsph, ddev = jo.effective_stress_gradient(eps, eps_g)
    np.testing.assert_allclose(fast[1], dsph, rtol=0,
                               atol=1e-9 * np.abs(dsph).max())
    np.testing.assert_allclose(fast[2], ddev, rtol=0,
                               atol=1e-9 * np.abs(ddev).max())


def test_effective_stress_parallel_load_needs_no_solve():
    """L = (C_1212 - 3)^2 under E = e_xy: dL/d<sigma> is parallel to E, so
    the backward is one contraction; it equals the ls_solve gradient."""
    import torch
    n = 16
    rho_np = ms.smoothed_density(n, 2.0, seed=5)
    mat0 = material.create(10.5, 0.3, 3)
    green = filtered(mat0.green_operator(), (n,) * 3, 1.0,
                     fft.create_real((n,) * 3))
    E = torch.zeros(6, dtype=torch.float64, device='cuda')
    E[5] = 1.0

    def grad(fast):
        rho = torch.tensor(rho_np, device='cuda', requires_grad=True)
        g = 1.0 + rho ** 3 * (20.0 - 1.0)
        sph, dev = iso_delta(g, torch.full_like(rho, 0.3), mat0, 3)
        info = {}
        if fast:
            sigma = effective_stress(sph, dev, E, green, mat0, tol=1e-14,
                                     max_iter=3000, info=info)
        else:
            eps = ls_solve(sph, dev, E, green, tol=1e-14, max_iter=3000,
                           info=info)
            sigma = mean_stress(sph, dev, eps, E, mat0)
        loss = (0.5 * sigma[5] / E[5] - 3.0) ** 2
        loss.backward()
        return rho.grad.cpu().numpy(), info

    g_fast, info = grad(True)
    g_full, _ = grad(False)
    assert info['adjoint'] is None
    np.testing.assert_allclose(g_fast, g_full, rtol=0,
                               atol=1e-9 * np.abs(g_full).max())
