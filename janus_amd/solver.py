"""On-device basic (Moulinec-Suquet) fixed-point solver.

janus leaves the solver loop to user scripts; the canonical one is
sphinx/tutorials/square_basic/square_basic.py:83-97:

    eps = E;  repeat: new = E - Gamma0[(C - C0) : eps]
                      res = ||new - eps|| / (sqrt(N) ||E||);  eps = new

:class:`LSSolver` runs exactly that loop on the GPU through
``jx_solver_solve``.  An iteration is a chain of passes over HBM: the local
operator fused into the last-axis r2c, the other axes' forward transforms,
Gamma-hat recomputed between the forward and inverse transforms of axis 0,
the inverse transforms, and the last-axis c2r fused with ``E - .``, the
residual partial sums and -- in the default fused schedule -- the local
operator and r2c of the next iteration.  The residual is reduced on the
device and the host reads the solver state once per chunk of iterations.
The iteration count is the one the host loop would give: the loop stops
after the first iteration whose residual is below ``tol``.
"""

import ctypes
from collections import namedtuple

import numpy as np

from . import _device, _lib
from .material.elastic.linear.isotropic import IsotropicLinearElasticMaterial

SolveResult = namedtuple('SolveResult', 'eps residuals iterations status')

STATUS = {_lib.JX_SOLVE_CONVERGED: 'converged',
          _lib.JX_SOLVE_MAXITER: 'max_iter',
          _lib.JX_SOLVE_NONFINITE: 'nonfinite'}


class LSSolver:
    """Fixed point eps <- E + rhs - Gamma0[local : eps] on the GPU.

    Args:
        green: a janus_amd.green operator built with a transform.
        local: a janus_amd.operators.BlockDiagonalOperator2D/3D holding
            C - C0.
    """

    def __init__(self, green, local):
        if green.transform is None:
            raise ValueError('the Green operator needs a transform')
        if tuple(green.ishape) != tuple(local.ishape):
            raise ValueError('Green operator and local operator shapes '
                             'differ: {0} vs {1}'.format(green.ishape,
                                                         local.ishape))
        _lib.require_gpu()
        self.green = green
        self.local = local
        self.shape = tuple(green.ishape)
        self.sym = self.shape[-1]
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().jx_solver_create(green.handle, local.handle,
                                               ctypes.byref(h)))
        self._handle = h

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and _lib._lib is not None:
            _lib._lib.jx_solver_destroy(h)
            self._handle = None

    @property
    def fused(self):
        """True when the last pass of an iteration also prepares the next
        one (the default wherever the register kernels apply)."""
        return _lib.lib().jx_solver_get_fused(self._handle) == 1

    @fused.setter
    def fused(self, value):
        _lib.check(_lib.lib().jx_solver_set_fused(self._handle, int(bool(value))))

    def autotune(self, iters=5):
        """Time every pass of this grid with each register FFT plan variant
        (``iters`` launches each, on scratch fields), keep the fastest in the
        Green operator and choose the fused or five-pass schedule.

        Returns {pass name: [ms of variant 0, ms of variant 1]} (0 when a
        variant does not exist or the pass does not run on this grid)."""
        ms = np.zeros(10)
        _lib.check(_lib.lib().jx_solver_autotune(
            self._handle, int(iters), ms.ctypes.data, _lib.current_stream()))
        names = self.green.PLAN_PASSES
        return {names[p]: [float(ms[2 * p]), float(ms[2 * p + 1])]
                for p in range(len(names))}

    def solve(self, E, tol=1e-10, max_iter=200, eps=None, rhs=None,
              res_scale=None):
        """Run the fixed point.

        Args:
            E: macroscopic strain (sym,), Mandel notation.
            tol: stop after the first iteration whose residual is < tol.
            max_iter: iteration cap.
            eps: optional float64 GPU tensor of shape ``grid + (sym,)``: the
                initial iterate, overwritten by the result.  Default: E.
            rhs: optional GPU field added to E.
            res_scale: residual normalization; default 1 / (sqrt(N) ||E||).
        Returns:
            SolveResult(eps, residuals, iterations, status).
        """
        t = _device.torch()
        E = np.ascontiguousarray(E, dtype=np.float64)
        if E.shape != (self.sym,):
            raise ValueError('E must have shape ({0},)'.format(self.sym))
        if eps is None:
            eps = t.empty(self.shape, dtype=t.float64, device='cuda')
            eps[...] = t.from_numpy(E).to('cuda')
        elif not (_device.is_device_tensor(eps) and eps.is_contiguous() and
                  tuple(eps.shape) == self.shape and eps.dtype == t.float64):
            raise ValueError('eps must be a contiguous float64 GPU tensor of '
                             'shape {0}'.format(self.shape))
        rhs_ptr = None
        if rhs is not None:
            if not (_device.is_device_tensor(rhs) and rhs.is_contiguous() and
                    tuple(rhs.shape) == self.shape and
                    rhs.dtype == t.float64):
                raise ValueError('rhs must be a contiguous float64 GPU field '
                                 'of shape {0}'.format(self.shape))
            rhs_ptr = rhs.data_ptr()
        if res_scale is None:
            res_scale = 0.0
        hist = np.zeros(max_iter, dtype=np.float64)
        iters = ctypes.c_int(0)
        status = ctypes.c_int(0)
        _lib.check(_lib.lib().jx_solver_solve(
            self._handle, eps.data_ptr(), E.ctypes.data, rhs_ptr, float(tol),
            int(max_iter), float(res_scale), hist.ctypes.data,
            ctypes.byref(iters), ctypes.byref(status),
            _lib.current_stream()))
        n = iters.value
        return SolveResult(eps, hist[:n], n, STATUS.get(status.value, 'unknown'))

    def solve_cg(self, E, tol=1e-10, max_iter=200, eps=None, rhs=None,
                 res_scale=None):
        """Conjugate gradients on the same operators: the fluctuation
        eps - E solves (I + Gamma0 dC) x = -Gamma0[dC : E] in the inner
        product sum a : C0 : b.  Converges for any isotropic C0 with C
        positive definite, including C0 = the matrix phase (the basic scheme
        diverges there).  Residual and stopping rule as :meth:`solve`.

        Args:
            E: macroscopic strain (sym,), Mandel notation.
            eps: optional float64 GPU tensor of shape ``grid + (sym,)``
                receiving the solution (its content is not used).
            rhs: optional GPU field added to E (compatible and zero-mean,
                e.g. a Gamma0 output, as in the adjoint solve).
        Returns:
            SolveResult(eps, residuals, iterations, status).
        """
        t = _device.torch()
        E = np.ascontiguousarray(E, dtype=np.float64)
        if E.shape != (self.sym,):
            raise ValueError('E must have shape ({0},)'.format(self.sym))
        if eps is None:
            eps = t.empty(self.shape, dtype=t.float64, device='cuda')
        elif not (_device.is_device_tensor(eps) and eps.is_contiguous() and
                  tuple(eps.shape) == self.shape and eps.dtype == t.float64):
            raise ValueError('eps must be a contiguous float64 GPU tensor of '
                             'shape {0}'.format(self.shape))
        if rhs is not None and not (_device.is_device_tensor(rhs) and
                                    rhs.is_contiguous() and
                                    tuple(rhs.shape) == self.shape and
                                    rhs.dtype == t.float64):
            raise ValueError('rhs must be a contiguous float64 GPU field of '
                             'shape {0}'.format(self.shape))
        hist = np.zeros(max_iter, dtype=np.float64)
        iters = ctypes.c_int(0)
        status = ctypes.c_int(0)
        _lib.check(_lib.lib().jx_solver_solve_cg(
            self._handle, eps.data_ptr(), E.ctypes.data, _lib.ptr(rhs),
            float(tol),
            int(max_iter), float(res_scale or 0.0), hist.ctypes.data,
            ctypes.byref(iters), ctypes.byref(status),
            _lib.current_stream()))
        n = iters.value
        return SolveResult(eps, hist[:n], n, STATUS.get(status.value, 'unknown'))

    def solve_cg_polarization(self, E, tol=1e-10, max_iter=200, tau=None,
                              eps=None, res_scale=None):
        """Conjugate gradients on the polarization form, as janus's
        examples/square_inclusion.py:17-52, 77-85 runs PETSc CG: this
        solver's local operator must be the inverse of C - C0 (the
        example's ``tau2eps``), and tau solves (local + Gamma0) tau = E from
        tau = 0, in the Euclidean inner product.  Needs C0 strictly below
        (or above) every phase.  The residual is ||E - A tau_k|| /
        (sqrt(N) ||E||) (PETSc's relative residual without preconditioner).

        Returns (tau, SolveResult(eps, residuals, iterations, status)) with
        eps = local : tau."""
        t = _device.torch()
        E = np.ascontiguousarray(E, dtype=np.float64)
        if E.shape != (self.sym,):
            raise ValueError('E must have shape ({0},)'.format(self.sym))
        out = []
        for f in (tau, eps):
            if f is None:
                f = t.empty(self.shape, dtype=t.float64, device='cuda')
            elif not (_device.is_device_tensor(f) and f.is_contiguous() and
                      tuple(f.shape) == self.shape and f.dtype == t.float64):
                raise ValueError('tau and eps must be contiguous float64 GPU '
                                 'tensors of shape {0}'.format(self.shape))
            out.append(f)
        tau, eps = out
        hist = np.zeros(max_iter, dtype=np.float64)
        iters = ctypes.c_int(0)
        status = ctypes.c_int(0)
        _lib.check(_lib.lib().jx_solver_solve_cg_polarization(
            self._handle, tau.data_ptr(), eps.data_ptr(), E.ctypes.data,
            float(tol), int(max_iter), float(res_scale or 0.0),
            hist.ctypes.data, ctypes.byref(iters), ctypes.byref(status),
            _lib.current_stream()))
        n = iters.value
        return tau, SolveResult(eps, hist[:n], n,
                                STATUS.get(status.value, 'unknown'))

    def solve_gmres(self, E, tol=1e-10, max_iter=200, restart=20, eps=None,
                    rhs=None, res_scale=None):
        """Restarted GMRES on the same strain form as :meth:`solve_cg`
        (examples/square_inclusion_scipy.py:52-59, 88 gives it to scipy's
        GMRES, whose default restart is 20), in the Euclidean inner product:
        no symmetry of dC is needed.  The residual is the Arnoldi estimate of
        ||b - A x_k|| with the normalization of :meth:`solve`; ``restart``
        Arnoldi steps per cycle (restart + 3 fields of the grid are
        allocated for the call).  Other arguments as :meth:`solve_cg`.
        Returns SolveResult(eps, residuals, iterations, status)."""
        t = _device.torch()
        E = np.ascontiguousarray(E, dtype=np.float64)
        if E.shape != (self.sym,):
            raise ValueError('E must have shape ({0},)'.format(self.sym))
        if eps is None:
            eps = t.empty(self.shape, dtype=t.float64, device='cuda')
        elif not (_device.is_device_tensor(eps) and eps.is_contiguous() and
                  tuple(eps.shape) == self.shape and eps.dtype == t.float64):
            raise ValueError('eps must be a contiguous float64 GPU tensor of '
                             'shape {0}'.format(self.shape))
        if rhs is not None and not (_device.is_device_tensor(rhs) and
                                    rhs.is_contiguous() and
                                    tuple(rhs.shape) == self.shape and
                                    rhs.dtype == t.float64):
            raise ValueError('rhs must be a contiguous float64 GPU field of '
                             'shape {0}'.format(self.shape))
        hist = np.zeros(max_iter, dtype=np.float64)
        iters = ctypes.c_int(0)
        status = ctypes.c_int(0)
        _lib.check(_lib.lib().jx_solver_solve_gmres(
            self._handle, eps.data_ptr(), E.ctypes.data, _lib.ptr(rhs),
            float(tol), int(max_iter), int(restart), float(res_scale or 0.0),
            hist.ctypes.data, ctypes.byref(iters), ctypes.byref(status),
            _lib.current_stream()))
        n = iters.value
        return SolveResult(eps, hist[:n], n, STATUS.get(status.value, 'unknown'))

    def mean_stress(self, eps, E, mat0):
        """<sigma> = C0 : E + <(C - C0) : eps> (square_basic.rst:212-221)."""
        tau, _ = self.local.means(eps)
        return reference_stiffness(mat0).dot(np.asarray(E)) + tau


def reference_stiffness(mat0):
    """Mandel matrix of the isotropic reference stiffness C0 = dim k0 J +
    2 g0 K."""
    if not isinstance(mat0, IsotropicLinearElasticMaterial):
        raise ValueError('isotropic reference material expected')
    from .operators import isotropic_4
    return isotropic_4(mat0.dim * mat0.k, 2. * mat0.g, mat0.dim).to_memoryview()


def delta_stiffness(mat, mat0):
    """(sph, dev) of C - C0 as built in square_basic.py:26-29."""
    return mat.dim * (mat.k - mat0.k), 2. * (mat.g - mat0.g)


def effective_stiffness(solver, mat0, tol=1e-10, max_iter=200,
                        method='basic', restart=20):
    """Columns C_eff[:, j] = <sigma> under the Mandel unit loads E = e_j.

    method: 'basic' (the fixed point), 'cg' (:meth:`LSSolver.solve_cg`) or
    'gmres' (:meth:`LSSolver.solve_gmres` with ``restart``).
    Returns (C_eff, iterations per load, status per load)."""
    if method not in ('basic', 'cg', 'gmres'):
        raise ValueError("method must be 'basic', 'cg' or 'gmres'")
    sym = solver.sym
    C = np.zeros((sym, sym))
    its, sts = [], []
    for j in range(sym):
        E = np.zeros(sym)
        E[j] = 1.0
        if method == 'cg':
            r = solver.solve_cg(E, tol=tol, max_iter=max_iter)
        elif method == 'gmres':
            r = solver.solve_gmres(E, tol=tol, max_iter=max_iter,
                                   restart=restart)
        else:
            r = solver.solve(E, tol=tol, max_iter=max_iter)
        C[:, j] = solver.mean_stress(r.eps, E, mat0)
        its.append(r.iterations)
        sts.append(r.status)
    return C, its, sts
