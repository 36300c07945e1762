"""Many fixed-point solves in one sequence of launches.

janus runs one fixed point at a time from user scripts
(sphinx/tutorials/square_basic/square_basic.py:83-97); effective-stiffness
studies repeat it for every load case and microstructure.  A
:class:`BatchSolver` takes any number of right-hand sides of any
:class:`janus_amd.solver.LSSolver` -- load cases of one solver, solvers on
different grids -- and runs them together through ``jx_batch_run``:

* right-hand sides on the same grid (and Green operator kind and h) share
  every kernel launch, ``width`` of them at a time, so that small grids fill
  the GPU;
* each replay runs ``iters_per_launch`` iterations of every group chunk that
  still has a running entry, and after the first replay the whole replay is a
  captured hipGraph, captured again only when a chunk has finished;
* with ``autotune=True`` the first replays time the candidate widths (all,
  12, 6, 3, 2, 1 entries per launch) of every group and keep the one with
  the least time per entry-iteration;
* with ``streams > 1`` and ``graph=False`` the groups of different grids
  run side by side on that many streams, so that small grids' launches fill
  what large ones leave of the GPU.

Each entry stops after the first iteration whose residual is below its
``tol``, exactly as :meth:`LSSolver.solve` would, with the same iterates.
Entries added with ``method='cg'`` run :meth:`LSSolver.solve_cg` instead
(conjugate gradients on the strain form, symmetric local operators, register
FFT lengths on every axis), with its iterates; they share launches with
other CG entries of their grid only.
"""

import ctypes

import numpy as np

from . import _device, _lib
from .solver import STATUS, SolveResult


class BatchSolver:
    """A set of right-hand sides solved together on the GPU.

    Args:
        iters_per_launch: iterations per replay (host reads the states once
            per replay).
        graph: replay a captured hipGraph (True) or enqueue the launches.
        autotune: pick each group's width by timing the first replays.
        streams: with ``graph=False``, groups of different grids run on up
            to this many streams at once (1: one after another); captured
            graphs keep one stream.
    """

    def __init__(self, iters_per_launch=8, graph=True, autotune=False,
                 streams=1):
        _lib.require_gpu()
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().jx_batch_create(ctypes.byref(h)))
        self._handle = h
        self._entries = []
        _lib.check(_lib.lib().jx_batch_set_schedule(
            h, int(iters_per_launch), int(bool(graph)), int(bool(autotune))))
        _lib.check(_lib.lib().jx_batch_set_streams(h, int(streams)))

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and _lib._lib is not None:
            _lib._lib.jx_batch_destroy(h)
            self._handle = None

    def add(self, solver, E, tol=1e-10, max_iter=200, eps=None, rhs=None,
            res_scale=None, method='basic'):
        """Add one right-hand side; arguments as :meth:`LSSolver.solve`,
        ``method`` 'basic' (the fixed point) or 'cg' (as
        :meth:`LSSolver.solve_cg`).

        Returns the entry index.  ``eps`` (default: E everywhere) is the
        entry's own GPU field, overwritten by the last iterate; a fixed-point
        entry starts from its content (a second run continues), a CG entry
        from E, as :meth:`LSSolver.solve_cg`."""
        if method not in ('basic', 'cg'):
            raise ValueError("method must be 'basic' or 'cg'")
        t = _device.torch()
        E = np.ascontiguousarray(E, dtype=np.float64)
        if E.shape != (solver.sym,):
            raise ValueError('E must have shape ({0},)'.format(solver.sym))
        if eps is None:
            eps = t.empty(solver.shape, dtype=t.float64, device='cuda')
            eps[...] = t.from_numpy(E).to('cuda')
        elif not (_device.is_device_tensor(eps) and eps.is_contiguous() and
                  tuple(eps.shape) == solver.shape and
                  eps.dtype == t.float64):
            raise ValueError('eps must be a contiguous float64 GPU tensor of '
                             'shape {0}'.format(solver.shape))
        if rhs is not None and not (_device.is_device_tensor(rhs) and
                                    rhs.is_contiguous() and
                                    tuple(rhs.shape) == solver.shape and
                                    rhs.dtype == t.float64):
            raise ValueError('rhs must be a contiguous float64 GPU field of '
                             'shape {0}'.format(solver.shape))
        index = ctypes.c_int(-1)
        add = (_lib.lib().jx_batch_add_cg if method == 'cg'
               else _lib.lib().jx_batch_add)
        _lib.check(add(
            self._handle, solver._handle, eps.data_ptr(), E.ctypes.data,
            _lib.ptr(rhs), float(tol), int(max_iter),
            float(res_scale or 0.0), ctypes.byref(index)))
        # keep the solver and the fields alive while the batch holds them
        self._entries.append((solver, eps, rhs, int(max_iter)))
        return index.value

    def set_width(self, group, width):
        """Entries per launch of ``group`` (0 = the whole group)."""
        _lib.check(_lib.lib().jx_batch_set_width(self._handle, int(group),
                                                 int(width)))

    def groups(self):
        """[(shape, size, width, batched)] in order of first appearance."""
        lib = _lib.lib()
        n = ctypes.c_int(0)
        _lib.check(lib.jx_batch_ngroups(self._handle, ctypes.byref(n)))
        out = []
        for g in range(n.value):
            dim, size, width, batched = (ctypes.c_int(), ctypes.c_int(),
                                         ctypes.c_int(), ctypes.c_int())
            shape = (ctypes.c_int64 * 3)()
            _lib.check(lib.jx_batch_group(
                self._handle, g, ctypes.byref(dim), shape, ctypes.byref(size),
                ctypes.byref(width), ctypes.byref(batched)))
            out.append((tuple(shape[:dim.value]), size.value, width.value,
                        bool(batched.value)))
        return out

    def tuning(self):
        """Per group, the widths the autotuner timed in the last run and
        their ms per entry-iteration: [[(width, ms), ...], ...]."""
        lib = _lib.lib()
        out = []
        for g in range(len(self.groups())):
            n = ctypes.c_int(0)
            w = (ctypes.c_int * 16)()
            ms = (ctypes.c_double * 16)()
            _lib.check(lib.jx_batch_tuning(self._handle, g, 16,
                                           ctypes.byref(n), w, ms))
            out.append([(w[i], ms[i]) for i in range(min(n.value, 16))])
        return out

    def run(self):
        """Run every entry to its stop; returns [SolveResult] in entry
        order."""
        lib = _lib.lib()
        _lib.check(lib.jx_batch_run(self._handle, _lib.current_stream()))
        results = []
        for i, (_, eps, _, max_iter) in enumerate(self._entries):
            hist = np.zeros(max_iter, dtype=np.float64)
            iters, status = ctypes.c_int(0), ctypes.c_int(0)
            _lib.check(lib.jx_batch_result(self._handle, i, hist.ctypes.data,
                                           ctypes.byref(iters),
                                           ctypes.byref(status)))
            n = iters.value
            results.append(SolveResult(eps, hist[:n], n,
                                       STATUS.get(status.value, 'unknown')))
        return results

    def stats(self):
        """(wall ms, replays, graph captures) of the last run."""
        ms = ctypes.c_double(0.)
        replays, captures = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(_lib.lib().jx_batch_stats(self._handle, ctypes.byref(ms),
                                             ctypes.byref(replays),
                                             ctypes.byref(captures)))
        return ms.value, replays.value, captures.value


def effective_stiffness_batched(solvers, mats0, tol=1e-10, max_iter=200,
                                method='basic', **schedule):
    """Full effective stiffness of every solver (all sym Mandel unit loads)
    in one batch, by the fixed point (method 'basic') or conjugate gradients
    ('cg').  Returns ([C_eff], [[iterations]], [[status]], batch)."""
    from .solver import reference_stiffness
    batch = BatchSolver(**schedule)
    index = []
    for S in solvers:
        row = []
        for j in range(S.sym):
            E = np.zeros(S.sym)
            E[j] = 1.0
            row.append(batch.add(S, E, tol=tol, max_iter=max_iter,
                                 method=method))
        index.append(row)
    res = batch.run()
    Cs, its, sts = [], [], []
    for S, m0, row in zip(solvers, mats0, index):
        C = np.zeros((S.sym, S.sym))
        C0 = reference_stiffness(m0)
        for j, i in enumerate(row):
            E = np.zeros(S.sym)
            E[j] = 1.0
            tau, _ = S.local.means(res[i].eps)
            C[:, j] = C0.dot(E) + tau
        Cs.append(C)
        its.append([res[i].iterations for i in row])
        sts.append([res[i].status for i in row])
    return Cs, its, sts, batch
