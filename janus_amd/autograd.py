"""Implicit-differentiation adjoint of the fixed-point solve (PyTorch-ROCm).

janus has no gradients; this module adds them on top of the same fused GPU
iteration.  With G = Gamma0 and dC = sph J + dev K (both symmetric), the
converged strain solves F(eps, dC) = eps + G[dC : eps] - E = 0.  For a loss
L(eps) with g = dL/deps:

* the adjoint state lambda solves (I + dC G) lambda = g; mu := G lambda then
  solves (I + G dC) mu = G g, i.e. the *same* fixed point as the forward
  solve, mu <- G g - G[dC : mu], with the field G g as right-hand side;
* dL/dsph_x = -(J eps_x) . mu_x and dL/ddev_x = -(K eps_x) . mu_x (the
  contraction kernel jx_adjoint_contract);
* dL/dE = sum_x lambda_x = sum_x (g_x - dC_x : mu_x).

Only the converged eps is kept from the forward pass.  The backward costs
one Green-operator application, one fixed-point solve and one contraction.
Both solves use the same method: conjugate gradients (``method='cg'``, the
default for the isotropic and other symmetric local operators), where the
right-hand side G g enters the initial residual only, so an adjoint
iteration costs what a forward one does; GMRES (the default for
non-symmetric matrix fields); or the basic scheme (``method='basic'``),
whose adjoint fixed point reads G g in every update.

A forward or adjoint solve that stops without reaching ``tol`` (status
'max_iter' or 'nonfinite', with tol > 0) raises a RuntimeWarning: its field, and any
gradient built on it, is not the converged one.

For losses of the macroscopic stress only (effective moduli),
effective_stress needs no adjoint solve: there g = dC gbar / N and mu is
the forward fluctuation under the load gbar (see its docstring).
"""

import ctypes
import warnings

import numpy as np

from . import _device, _lib
from .operators import BlockDiagonalOperator2D, BlockDiagonalOperator3D
from .solver import LSSolver


def _torch():
    return _device.torch()


def _local_from_fields(sph, dev):
    cls = BlockDiagonalOperator2D if sph.dim() == 2 else BlockDiagonalOperator3D
    return cls.from_fields(sph, dev)


def _check_status(r, what, tol):
    """Warn when a solve stopped without converging (r: SolveResult or
    None for a solve that was not needed).  tol <= 0 asks for a fixed
    iteration count: stopping at max_iter is then the request."""
    if r is not None and r.status != 'converged' and \
            not (tol <= 0 and r.status == 'max_iter'):
        warnings.warn('janus_amd: the {0} solve stopped with status {1!r} '
                      'after {2} iterations (residual {3:.3e}); its result '
                      'is not converged'.format(
                          what, r.status, r.iterations,
                          float(r.residuals[-1]) if len(r.residuals)
                          else float('nan')),
                      RuntimeWarning, stacklevel=3)


def _resolve_method(method, symmetric, allowed):
    """'auto' -> conjugate gradients for a symmetric local operator, GMRES
    otherwise (GMRES only where ``allowed`` lists it)."""
    if method == 'auto':
        method = 'cg' if symmetric else 'gmres'
    if method not in allowed:
        raise ValueError('method must be one of {0}'.format(
            ', '.join(repr(m) for m in ('auto',) + tuple(allowed))))
    return method


def contract(eps, mu):
    """(dsph, ddev) = (-(J eps).mu, -(K eps).mu) voxel by voxel (GPU)."""
    t = _torch()
    grid = tuple(eps.shape[:-1])
    dsph = t.empty(grid, dtype=t.float64, device=eps.device)
    ddev = t.empty(grid, dtype=t.float64, device=eps.device)
    _lib.check(_lib.lib().jx_adjoint_contract(
        len(grid), _lib.shape_array(grid), eps.data_ptr(), mu.data_ptr(),
        dsph.data_ptr(), ddev.data_ptr(), _lib.current_stream()))
    return dsph, ddev


def adjoint_solve(green, local, g, tol, max_iter, solver=None,
                  method='basic'):
    """mu solving mu = G g - G[local : mu]; returns (mu, SolveResult).

    ``solver``: an LSSolver on (green, local) to reuse (e.g. the forward
    solve's); ``method``: 'basic' (fixed point), 'cg' or 'gmres'."""
    t = _torch()
    gg = green.apply(g.contiguous())
    norm = float(t.linalg.vector_norm(gg))
    if norm == 0.0:
        return t.zeros_like(gg), None
    if solver is None:
        solver = LSSolver(green, local)
    zero = np.zeros(gg.shape[-1])
    if method == 'cg':
        r = solver.solve_cg(zero, tol=tol, max_iter=max_iter, rhs=gg,
                            res_scale=1.0 / norm)
        _check_status(r, 'adjoint', tol)
        return r.eps, r
    if method == 'gmres':
        r = solver.solve_gmres(zero, tol=tol, max_iter=max_iter, rhs=gg,
                               res_scale=1.0 / norm)
        _check_status(r, 'adjoint', tol)
        return r.eps, r
    mu = gg.clone()
    r = solver.solve(zero, tol=tol, max_iter=max_iter, eps=mu, rhs=gg,
                     res_scale=1.0 / norm)
    _check_status(r, 'adjoint', tol)
    return mu, r


class _LSSolveFunction:
    """Namespace holding the torch.autograd.Function (built lazily so that
    importing this module does not import torch)."""
    fn = None


def _make_function():
    t = _torch()

    class LSSolveFunction(t.autograd.Function):
        @staticmethod
        def forward(ctx, sph, dev, E, green, tol, max_iter, info, method):
            sph = sph.detach().contiguous()
            dev = dev.detach().contiguous()
            local = _local_from_fields(sph, dev)
            solver = LSSolver(green, local)
            r = _solve(solver, E.detach().cpu().numpy(), tol, max_iter,
                       method)
            _check_status(r, 'forward', tol)
            if info is not None:
                info['forward'] = r
            ctx.save_for_backward(sph, dev, r.eps)
            ctx.green = green
            # the backward reuses the forward's operators (same sph, dev)
            ctx.local = local
            ctx.solver = solver
            ctx.tol = tol
            ctx.max_iter = max_iter
            ctx.info = info
            ctx.method = method
            return r.eps

        @staticmethod
        def backward(ctx, g):
            sph, dev, eps = ctx.saved_tensors
            local = ctx.local
            mu, r = adjoint_solve(ctx.green, local, g, ctx.tol, ctx.max_iter,
                                  solver=ctx.solver, method=ctx.method)
            if ctx.info is not None:
                ctx.info['adjoint'] = r
            dsph, ddev = contract(eps, mu)
            lam = g.contiguous() - local.apply(mu)
            dims = tuple(range(lam.dim() - 1))
            dE = lam.sum(dim=dims)
            return dsph, ddev, dE, None, None, None, None, None

    return LSSolveFunction


def ls_solve(sph, dev, E, green, tol=1e-10, max_iter=1000, info=None,
             method='auto'):
    """Differentiable Lippmann-Schwinger solve.

    Args:
        sph, dev: float64 GPU tensors of the grid shape, the per-voxel
            coefficients of C - C0 = sph J + dev K (may require grad).
        E: float64 tensor (sym,), macroscopic strain (may require grad).
        green: a janus_amd.green operator with a transform (reference C0).
        info: optional dict receiving the SolveResult of the forward
            ('forward') and adjoint ('adjoint') solves.
        method: 'cg' (conjugate gradients, LSSolver.solve_cg; 'auto', the
            default, picks it: isotropic dC is symmetric) or 'basic' (the
            fixed point), for both the forward and the adjoint solve.
    Returns:
        eps, the converged strain field (grid + (sym,)).
    """
    method = _resolve_method(method, True, ('basic', 'cg'))
    if _LSSolveFunction.fn is None:
        _LSSolveFunction.fn = _make_function()
    return _LSSolveFunction.fn.apply(sph, dev, E, green, tol, max_iter, info,
                                     method)


class _EffectiveStressFunction:
    fn = None


def _make_effective_stress():
    t = _torch()

    class EffectiveStressFunction(t.autograd.Function):
        @staticmethod
        def forward(ctx, sph, dev, E, green, mat0, tol, max_iter, info,
                    method):
            sph = sph.detach().contiguous()
            dev = dev.detach().contiguous()
            local = _local_from_fields(sph, dev)
            solver = LSSolver(green, local)
            En = E.detach().cpu().numpy()
            r = _solve(solver, En, tol, max_iter, method)
            _check_status(r, 'forward', tol)
            if info is not None:
                info['forward'] = r
                info['adjoint'] = None
            tau, _ = local.means(r.eps)
            from .solver import reference_stiffness
            sigma = reference_stiffness(mat0).dot(En) + tau
            ctx.save_for_backward(r.eps)
            ctx.E = En
            ctx.c0 = reference_stiffness(mat0)
            ctx.local = local
            ctx.solver = solver
            ctx.args = (tol, max_iter, info, method)
            return t.from_numpy(sigma).to(E.device)

        @staticmethod
        def backward(ctx, gbar):
            (eps,) = ctx.saved_tensors
            tol, max_iter, info, method = ctx.args
            gb = gbar.detach().cpu().numpy().astype(np.float64)
            E = ctx.E
            ee = float(E.dot(E))
            alpha = float(gb.dot(E)) / ee if ee > 0 else 0.0
            nb = float(np.linalg.norm(gb))
            if nb == 0.0:
                return None, None, t.zeros_like(gbar), None, None, None, \
                    None, None, None
            if ee > 0 and np.linalg.norm(gb - alpha * E) <= 1e-14 * nb:
                # gbar parallel to E: eps(gbar) = alpha eps(E) (linearity)
                eps_g, scale = eps, alpha
            else:
                r = _solve(ctx.solver, gb, tol, max_iter, method)
                _check_status(r, 'adjoint', tol)
                if info is not None:
                    info['adjoint'] = r
                eps_g, scale = r.eps, 1.0
            N = eps.numel() // eps.shape[-1]
            # contract() returns (-(J eps) . mu, -(K eps) . mu)
            dsph, ddev = contract(eps, eps_g * (-scale / N))
            # dL/dE = C_eff gbar = <sigma(gbar)> (C_eff is symmetric)
            tau_g, _ = ctx.local.means(eps_g)
            dE = ctx.c0.dot(gb) + scale * tau_g
            return dsph, ddev, t.from_numpy(dE).to(gbar.device), None, None, \
                None, None, None, None

    return EffectiveStressFunction


def _solve(solver, E, tol, max_iter, method):
    if method == 'cg':
        return solver.solve_cg(E, tol=tol, max_iter=max_iter)
    if method == 'gmres':
        return solver.solve_gmres(E, tol=tol, max_iter=max_iter)
    return solver.solve(E, tol=tol, max_iter=max_iter)


def effective_stress(sph, dev, E, green, mat0, tol=1e-10, max_iter=1000,
                     info=None, method='auto'):
    """Differentiable macroscopic stress <sigma> = C0 : E + <dC : eps(E)>.

    The same solve as ls_solve followed by mean_stress, with a backward
    that needs no adjoint solve.  For a loss L(<sigma>) with gbar =
    dL/d<sigma>, dL/deps_x = dC_x gbar / N, so the adjoint state of
    ls_solve's backward solves (I + Gamma0 dC) mu = Gamma0 dC gbar / N:
    the forward equation of the fluctuation eps(gbar) - gbar under the
    uniform load gbar, times -1 / N.  Hence mu = -(eps(gbar) - gbar) / N,
    and with the explicit term of <dC : eps>, dL/dsph_x =
    (J eps_x) . eps(gbar)_x / N and dL/ddev_x = (K eps_x) . eps(gbar)_x / N.
    When gbar is parallel to E (a loss of the effective modulus that E
    probes), eps(gbar) is a multiple of the forward field and the backward
    is one contraction; otherwise it is one forward solve under gbar (no
    field right-hand side).

    Args: as ls_solve, plus mat0 (the isotropic reference material of
    green).  info receives 'forward' and 'adjoint' (None when no solve was
    needed).  Returns <sigma>, a (sym,) tensor."""
    method = _resolve_method(method, True, ('basic', 'cg'))
    if _EffectiveStressFunction.fn is None:
        _EffectiveStressFunction.fn = _make_effective_stress()
    return _EffectiveStressFunction.fn.apply(sph, dev, E, green, mat0, tol,
                                             max_iter, info, method)


def outer(eps, mu):
    """da = -mu (x) eps voxel by voxel (GPU): the gradient with respect to a
    per-voxel matrix local operator, shape grid + (sym, sym)."""
    t = _torch()
    grid = tuple(eps.shape[:-1])
    sym = eps.shape[-1]
    da = t.empty(grid + (sym, sym), dtype=t.float64, device=eps.device)
    _lib.check(_lib.lib().jx_adjoint_outer(
        len(grid), _lib.shape_array(grid), eps.data_ptr(), mu.data_ptr(),
        da.data_ptr(), _lib.current_stream()))
    return da


class _LSSolveMatrixFunction:
    fn = None


def _make_matrix_function():
    t = _torch()

    class LSSolveMatrixFunction(t.autograd.Function):
        @staticmethod
        def forward(ctx, mats, E, green, tol, max_iter, info, method):
            mats = mats.detach().contiguous()
            cls = (BlockDiagonalOperator2D if mats.dim() == 4
                   else BlockDiagonalOperator3D)
            local = cls.from_matrix_field(mats)
            method = _resolve_method(method, local.symmetric,
                                     ('basic', 'cg', 'gmres'))
            if method == 'cg' and not local.symmetric:
                raise ValueError("method 'cg' needs symmetric matrices "
                                 "(use 'gmres' or 'basic')")
            solver = LSSolver(green, local)
            r = _solve(solver, E.detach().cpu().numpy(), tol, max_iter,
                       method)
            _check_status(r, 'forward', tol)
            if info is not None:
                info['forward'] = r
            if local.symmetric:
                local_t, solver_t = local, solver
            else:
                # the adjoint fixed point runs on the transposed matrices
                mats_t = mats.transpose(-1, -2).contiguous()
                local_t = cls.from_matrix_field(mats_t)
                solver_t = LSSolver(green, local_t)
                ctx.mats_t = mats_t
            # mats shares the caller's storage and version counter (the
            # symmetric path's adjoint reuses the forward operator, which
            # reads it): saving it makes autograd refuse a backward after an
            # in-place change of the matrices
            ctx.save_for_backward(r.eps, mats)
            ctx.green = green
            ctx.local_t = local_t
            ctx.solver_t = solver_t
            ctx.args = (tol, max_iter, info, method)
            return r.eps

        @staticmethod
        def backward(ctx, g):
            eps, _ = ctx.saved_tensors
            tol, max_iter, info, method = ctx.args
            mu, r = adjoint_solve(ctx.green, ctx.local_t, g, tol, max_iter,
                                  solver=ctx.solver_t, method=method)
            if info is not None:
                info['adjoint'] = r
            da = outer(eps, mu)
            lam = g.contiguous() - ctx.local_t.apply(mu)
            dE = lam.sum(dim=tuple(range(lam.dim() - 1)))
            return da, dE, None, None, None, None, None

    return LSSolveMatrixFunction


def ls_solve_matrix(mats, E, green, tol=1e-10, max_iter=1000, info=None,
                    method='auto'):
    """Differentiable Lippmann-Schwinger solve with a per-voxel matrix local
    operator (BlockDiagonalLinearOperator, operators.pyx:719-890).

    F(eps, A) = eps + G[A eps] - E = 0 with G = Gamma0.  For a loss L(eps)
    with g = dL/deps, the adjoint state solves (I + A^T G) lambda = g, so
    mu := G lambda solves the fixed point mu = G g - G[A^T mu] (the forward
    one with the transposed matrices and the field right-hand side G g);
    then dL/dA_x = -mu_x (x) eps_x (jx_adjoint_outer) and dL/dE =
    sum_x (g_x - A_x^T mu_x).

    Args:
        mats: float64 GPU tensor grid + (sym, sym) (may require grad); need
            not be symmetric.
        E, green, tol, max_iter, info: as ls_solve.
        method: 'basic', 'cg' (symmetric matrices only) or 'gmres', for
            the forward and the adjoint solve; 'auto' (default): 'cg' for
            symmetric matrices, 'gmres' otherwise.
    Returns:
        eps, the converged strain field."""
    if method not in ('auto', 'basic', 'cg', 'gmres'):
        raise ValueError("method must be 'auto', 'basic', 'cg' or 'gmres'")
    if _LSSolveMatrixFunction.fn is None:
        _LSSolveMatrixFunction.fn = _make_matrix_function()
    return _LSSolveMatrixFunction.fn.apply(mats, E, green, tol, max_iter,
                                           info, method)


def iso_delta(g, nu, mat0, dim):
    """(sph, dev) of C(g, nu) - C0 for tensors g, nu (differentiable):
    sph = dim (k - k0), dev = 2 (g - g0) as square_basic.py:26-29 builds it,
    with k from isotropic.pyx:79-97."""
    if dim == 2:
        k = g / (1. - 2. * nu)
    else:
        k = 2. / 3. * (1. + nu) / (1. - 2. * nu) * g
    return dim * (k - mat0.k), 2. * (g - mat0.g)


def mean_stress(sph, dev, eps, E, mat0):
    """<sigma> = C0 : E + <dC : eps> as a differentiable torch expression."""
    t = _torch()
    dim = eps.dim() - 1
    sym = eps.shape[-1]
    tr = (sph - dev) / dim
    trace = eps[..., :dim].sum(dim=-1)
    tau_n = (tr * trace).unsqueeze(-1) + dev.unsqueeze(-1) * eps[..., :dim]
    tau_s = dev.unsqueeze(-1) * eps[..., dim:]
    tau = t.cat([tau_n, tau_s], dim=-1)
    mean_tau = tau.reshape(-1, sym).mean(dim=0)
    from .solver import reference_stiffness
    c0 = t.from_numpy(reference_stiffness(mat0)).to(eps.device)
    return c0 @ E + mean_tau
