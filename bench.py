"""Benchmark: basic-scheme Lippmann-Schwinger iterations per second at 256^3
in FP64 on one MI355X, with the HBM roofline of the dominant kernel and the
CPU baseline.

Workload (a "step" is one fixed-point iteration eps <- E - Gamma0[(C-C0):eps]
over the whole grid, one load case):
  * 256^3 voxels, two phases: Boolean model of spheres of radius 16 voxels,
    volume fraction about 0.2, seed 2014 (the 128^3 sphere generator scaled
    to 256^3);
  * matrix mu = 1, nu = 0.3; inclusions mu = 50, nu = 0.25;
  * filtered Green operator; reference medium mu0 = 25.5, nu0 = 0.3 (the
    Moulinec-Suquet midpoint -- with C0 = matrix the basic scheme diverges);
  * Mandel macroscopic strain E = e_xy.
Inputs are resident in HBM before the timed region.

The default JSON line also carries, measured in the same process:
  * roofline: the slowest pass's algorithmic bytes over its hipEvent time,
    and its HBM traffic per launch from the committed rocprofv3 counter
    summary (null when that summary was recorded for other kernel sources);
  * cpu_baseline / cpu_baseline_all_cores: the C restatement of janus's
    serial path (oracle/janus_serial.c) on the same 256^3 workload, 5
    iterations, 1 thread (janus is single-threaded) and all host cores;
  * adjoint_over_forward and configs: the differentiable 128^3 SIMP
    configuration (adjoint-gradient time / forward time, the metric's second
    half), the batched {64..160}^3 mix (speedup over sequential solves and
    the bandwidth bound of the mix) and the 512^3 three-phase configuration
    (6-load batch iterations/s, roofline fraction, per-phase adjoint);
    --no-configs skips them.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       [--no-configs]

Other workloads (one JSON line each, --workload NAME):
  plain2d    2D 256^2 square inclusion, truncated, C0 = matrix: the stated
             configuration diverges; reported with the midpoint C0 run
  stiffness  128^3 spheres r=8, contrast 50: full 6x6 C_eff at 1e-10
  batched    32 microstructures on grids {64,80,96,128,160}^3 x 6 loads:
             one batch vs 192 sequential solves at 1e-10
  adjoint    128^3 SIMP density: adjoint gradient time / forward solve time
  large      512^3 three-phase ellipsoids: 6-load batch iterations/s and
             an adjoint gradient with respect to the per-phase mu
  fft        janus's fftw_python_benchmark_serial: r2c ms per call at
             128^3, 256^3, 512^3 (device tensors and numpy arrays)
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PASS_NAMES = ['A_rows_r2c_local', 'B_axis1_fwd', 'C_axis0_gamma',
              'D_axis1_inv', 'E_rows_c2r_update', 'residual_reduce']
PASS_NAMES_FUSED = ['A_rows_r2c_local', 'B_axis1_fwd', 'C_axis0_gamma',
                    'D_axis1_inv', 'EA_rows_c2r_update_local_r2c',
                    'residual_reduce']


def pass_bytes(n, fused, sym=6, m=1):
    """Algorithmic HBM bytes of each pass of one iteration (3D, FP64):
    real field 8 sym B/voxel, spectrum 16 sym B/point, phase id m B/voxel.
    Fused schedule: pass A is folded into pass E of the previous iteration."""
    N = n[0] * n[1] * n[2]
    Nh = n[0] * n[1] * (n[2] // 2 + 1)
    f = 8 * sym
    s = 16 * sym
    if fused:
        return [0, 2 * s * Nh, 2 * s * Nh, 2 * s * Nh,
                # EA: read spectrum + eps_old + phase, write eps + spectrum
                2 * s * Nh + (2 * f + m) * N, 0]
    return [(f + m) * N + s * Nh,      # A: read eps + phase, write spectrum
            2 * s * Nh,                # B
            2 * s * Nh,                # C
            2 * s * Nh,                # D
            s * Nh + 2 * f * N,        # E: read spectrum + eps_old, write eps
            0]


def five_pass_bytes(n, sym=6, m=1):
    """B_alg = (144 + m) N + 768 Nh, the 5-pass per-iteration byte count
    the roofline target is stated against."""
    N = n[0] * n[1] * n[2]
    Nh = n[0] * n[1] * (n[2] // 2 + 1)
    return (16 * sym + 8 * sym + m) * N + 8 * 16 * sym * Nh


def build_workload(n, kind):
    import janus_amd.fft.serial as fft
    import janus_amd.material.elastic.linear.isotropic as material
    from janus_amd import microstructure as ms
    from janus_amd.green import filtered, truncated
    from janus_amd.operators import BlockDiagonalOperator3D
    from janus_amd.solver import LSSolver, delta_stiffness

    phase = ms.spheres(n, 16.0 * n / 256.0, 0.2, seed=2014)
    mat_m = material.create(1.0, 0.3, 3)
    mat_i = material.create(50.0, 0.25, 3)
    mat0 = material.create(25.5, 0.3, 3)
    table = np.array([delta_stiffness(mat_m, mat0), delta_stiffness(mat_i, mat0)])
    factory = filtered if kind == 'filtered' else truncated
    green = factory(mat0.green_operator(), (n,) * 3, 1.0,
                    fft.create_real((n,) * 3))
    local = BlockDiagonalOperator3D.from_phases(phase, table)
    return LSSolver(green, local), phase, table, mat0


def cpu_baseline(n, kind, phase, table, mat0, iters=5, threads=1):
    """The C restatement of janus's serial path (oracle/janus_serial.c:
    per-component r2c, janus's per-frequency Gamma-hat loop, c2r; the basic
    scheme of square_basic.py:83-92) on the same workload, single-threaded
    like janus (threads=1) or with OpenMP on `threads` host threads.
    Returns (iterations/s, seconds, residuals)."""
    import ctypes
    lib_path = os.path.join(ROOT, 'oracle', 'libjanus_serial.so')
    if not os.path.exists(lib_path):
        raise RuntimeError('oracle/libjanus_serial.so is missing (built by '
                           '__graft_entry__.build())')
    lib = ctypes.CDLL(lib_path)
    phase = np.ascontiguousarray(phase, dtype=np.uint8)
    table = np.ascontiguousarray(table, dtype=np.float64)
    E = np.zeros(6)
    E[5] = 1.0
    eps = np.empty((n, n, n, 6))
    eps[...] = E
    res = np.zeros(iters)
    shape = (ctypes.c_int64 * 3)(n, n, n)
    vp = ctypes.c_void_p
    t0 = time.perf_counter()
    rc = lib.js_basic_iterations(
        3, shape, 0 if kind == 'truncated' else 1, ctypes.c_double(mat0.g),
        ctypes.c_double(mat0.nu), ctypes.c_double(1.0),
        phase.ctypes.data_as(vp), table.ctypes.data_as(vp),
        E.ctypes.data_as(vp), eps.ctypes.data_as(vp), int(iters),
        res.ctypes.data_as(vp), int(threads))
    t = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError('js_basic_iterations failed (%d)' % rc)
    return iters / t, t, res


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def host_workers():
    """Host threads for the all-cores baseline: the CPUs this process may
    run on, at most 16 (one GPU's share of a GPU node)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def stream_copy_gbps(mib=2048, reps=10):
    """Bandwidth (read + write bytes / time) of the library's 16-byte
    stream-copy kernel: the practical HBM ceiling next to the 8 TB/s
    data-sheet peak."""
    import torch
    from janus_amd import _lib
    nbytes = mib * (1 << 20)
    a = torch.ones(nbytes // 8, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    lib = _lib.lib()
    stream = _lib.current_stream()

    def copy():
        _lib.check(lib.jx_stream_copy(a.data_ptr(), b.data_ptr(), nbytes,
                                      stream))
    copy()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        copy()
    end.record()
    torch.cuda.synchronize()
    sec = start.elapsed_time(end) * 1e-3 / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / sec / 1e9


TRAFFIC_TAG = 'traffic_per_launch'
SOURCES_TAG = 'kernel_sources_sha256'


def kernel_sources_hash():
    """sha256 of the HIP sources the library is built from (file names and
    contents, in name order): a counter summary recorded for other sources
    describes other kernels."""
    import hashlib
    csrc = os.path.join(ROOT, 'janus_amd', 'csrc')
    h = hashlib.sha256()
    for f in sorted(os.listdir(csrc)):
        if f.endswith(('.hip', '.h')):
            h.update(f.encode())
            with open(os.path.join(csrc, f), 'rb') as fh:
                h.update(fh.read())
    return h.hexdigest()


def traffic_file(n, kind):
    return os.path.join(ROOT, 'profiles',
                        'ls_iteration_%d_%s_traffic.md' % (n, kind))


def read_traffic(n, kind, kernel, v):
    """HBM bytes per launch of a pass run with the v-values-per-thread plan,
    from the committed counter summary (lines '<tag> <pass> <v> <bytes>'),
    and where it came from.  (None, reason) when the summary is missing,
    lists no such kernel, or was recorded for other kernel sources than
    these (its '<sources tag> <sha256>' line)."""
    path = traffic_file(n, kind)
    rel = os.path.relpath(path, ROOT)
    if not os.path.exists(path):
        return None, 'no counter summary %s' % rel
    recorded = None
    value = None
    for line in open(path):
        f = line.strip().strip('`').split()
        if len(f) == 2 and f[0] == SOURCES_TAG:
            recorded = f[1]
        if len(f) == 4 and f[0] == TRAFFIC_TAG and f[1] == kernel and \
                f[2] == str(v):
            value = int(f[3])
    if recorded != kernel_sources_hash():
        return None, ('%s was recorded for other kernel sources (%s %s); '
                      'rerun the two --pmc passes' % (rel, SOURCES_TAG,
                                                      recorded))
    if value is None:
        return None, '%s lists no %s with %d values per thread' % (rel,
                                                                   kernel, v)
    return value, '%s (%s %s)' % (rel, SOURCES_TAG, recorded[:16])


def plan_v(name):
    """Values per thread of the plan in a kernel symbol (Plan<L, V, ...>)."""
    if 'Plan<' not in name:
        return None
    return int(name.split('Plan<')[1].split(',')[1])


# bench pass -> plan pass of the Green operator
PASS_PLAN = {'A_rows_r2c_local': 'rows_r2c', 'B_axis1_fwd': 'axis',
             'C_axis0_gamma': 'gamma', 'D_axis1_inv': 'axis',
             'E_rows_c2r_update': 'rows_c2r',
             'EA_rows_c2r_update_local_r2c': 'fused'}


def kernel_pass(name):
    """Bench pass name of a kernel symbol (None for other kernels)."""
    if 'k_stream_copy' in name:
        return 'stream_copy'
    head = name.split('(')[0].replace(' ', '')
    if '>,1,' in head:  # one component: the janus.fft.serial transforms
        if 'k_rows_forward_reg' in head:
            return 'fft_rows_r2c'
        if 'k_rows_inverse_reg' in head:
            return 'fft_rows_c2r'
        if 'k_axis_reg' in head:
            return 'fft_axis_fwd' if head.endswith(',-1>') else 'fft_axis_inv'
        return None
    if 'k_rows_update_forward_reg' in name:
        return 'EA_rows_c2r_update_local_r2c'
    if 'k_rows_forward_reg' in name:
        return 'A_rows_r2c_local'
    if 'k_rows_inverse_reg' in name:
        return 'E_rows_c2r_update'
    if 'k_axis_reg' in name:
        tail = name.split('(')[0].replace(' ', '')
        if tail.endswith(',-1>'):
            return 'B_axis1_fwd'
        if tail.endswith(',1>'):
            return 'D_axis1_inv'
        if tail.endswith(',0>'):
            return 'C_axis0_gamma'
    return None


def stats_summary(stats_csv, cmd, title='256^3 filtered LS iteration'):
    """Markdown table of a rocprofv3 --stats kernel summary (jx kernels)."""
    import csv
    out = ['# Kernel durations, ' + title, '',
           'Command:', '', '    ' + cmd, '',
           'rocprofv3 kernel statistics (jx kernels; the autotuner also '
           'launches the variants it does not keep):', '',
           '| kernel | pass | calls | average us | min us | max us |',
           '|---|---|---|---|---|---|']
    for r in csv.DictReader(open(stats_csv)):
        name = r['Name']
        if 'jx::' not in name:
            continue
        short = name.replace('(anonymous namespace)::', '').split('(')[0]
        short = short.replace('void ', '')
        out.append('| %s | %s | %s | %.1f | %.1f | %.1f |'
                   % (short, kernel_pass(name) or '', r['Calls'],
                      float(r['AverageNs']) / 1e3, float(r['MinNs']) / 1e3,
                      float(r['MaxNs']) / 1e3))
    return '\n'.join(out) + '\n'


def traffic_summary(fetch_csv, write_csv, cmd,
                    title='256^3 filtered LS iteration'):
    """Markdown summary of rocprofv3 FETCH_SIZE / WRITE_SIZE runs: each
    counter is scaled by what it reads for the stream-copy kernel, whose
    bytes are known (16 B per grid thread)."""
    import csv
    import collections
    vals = {}
    copies = {}
    for counter, path in (('FETCH_SIZE', fetch_csv), ('WRITE_SIZE', write_csv)):
        agg = collections.defaultdict(list)
        ratios = []
        for r in csv.DictReader(open(path)):
            name = r['Kernel_Name']
            v = float(r['Counter_Value']) * 1024.0
            agg[name].append(v)
            if 'k_stream_copy' in name:
                ratios.append(int(r['Grid_Size']) * 16 / v)
        vals[counter] = {k: sum(v) / len(v) for k, v in agg.items()}
        copies[counter] = sum(ratios) / len(ratios)
    out = ['# HBM traffic per launch, ' + title, '',
           'Commands (one counter per run):', '',
           '    ' + cmd.replace('COUNTER', 'FETCH_SIZE'),
           '    ' + cmd.replace('COUNTER', 'WRITE_SIZE'), '',
           'Calibration on k_stream_copy (bytes known): FETCH_SIZE x %.4f, '
           'WRITE_SIZE x %.4f.' % (copies['FETCH_SIZE'], copies['WRITE_SIZE']),
           '', '| kernel | FETCH raw MB | WRITE raw MB | calibrated read MB | '
           'calibrated write MB | traffic MB |', '|---|---|---|---|---|---|']
    tags = []
    for name in vals['FETCH_SIZE']:
        if 'jx::' not in name:
            continue
        f = vals['FETCH_SIZE'][name]
        w = vals['WRITE_SIZE'].get(name, 0.)
        fr, wr = f * copies['FETCH_SIZE'], w * copies['WRITE_SIZE']
        short = name.split('(')[0].replace('void ', '')
        out.append('| %s | %.1f | %.1f | %.1f | %.1f | %.1f |'
                   % (short, f / 1e6, w / 1e6, fr / 1e6, wr / 1e6,
                      (fr + wr) / 1e6))
        pname = kernel_pass(name)
        if pname:
            tags.append('    %s %s %s %d' % (TRAFFIC_TAG, pname, plan_v(name),
                                            int(fr + wr)))
    out += ['', 'Traffic per launch (bytes) read by bench.py, for the '
            'kernels built from these sources:', '',
            '    %s %s' % (SOURCES_TAG, kernel_sources_hash())] + tags
    return '\n'.join(out) + '\n'


# ---------------------------------------------------------------- workloads
def _sync():
    import torch
    torch.cuda.synchronize()


def _log(msg):
    print('bench: ' + msg, file=sys.stderr, flush=True)


def _two_phase_solver(phase, mat_m, mat_i, mat0, kind):
    import janus_amd.fft.serial as fft
    from janus_amd.green import filtered, truncated
    from janus_amd.operators import BlockDiagonalOperator2D, \
        BlockDiagonalOperator3D
    from janus_amd.solver import LSSolver, delta_stiffness
    dim = phase.ndim
    table = np.array([delta_stiffness(m, mat0) for m in (mat_m, mat_i)])
    cls = BlockDiagonalOperator2D if dim == 2 else BlockDiagonalOperator3D
    factory = filtered if kind == 'filtered' else truncated
    green = factory(mat0.green_operator(), phase.shape, 1.0,
                    fft.create_real(phase.shape))
    return LSSolver(green, cls.from_phases(phase, table)), table


def _unit(sym, j):
    E = np.zeros(sym)
    E[j] = 1.0
    return E


def _line(metric, value, unit, config, higher=True, **extra):
    d = {'metric': metric, 'value': value, 'unit': unit, 'n_gpus': 1,
         'higher_is_better': higher, 'dtype': 'f64', 'data': 'synthetic',
         'config': config}
    d.update(extra)
    print(json.dumps(d), flush=True)


def workload_plain2d(args):
    """2D 256^2, centred square inclusion (vf 0.25), matrix mu 1 / inclusion
    mu 10 (nu 0.3), truncated Gamma0, E = (0, 0, 1), 1e-10 or 200
    iterations.  With C0 = matrix the basic scheme diverges (the iteration
    operator has eigenvalues up to about mu_i / mu_0 - 1 = 9); the GPU
    iterates are checked against the oracle's over the first iterations, then
    the same problem runs with the midpoint C0."""
    import janus_amd.material.elastic.linear.isotropic as material
    from janus_amd import microstructure as ms
    from oracle import janus_oracle as jo
    n = 256
    phase = ms.square_inclusion(n, 0.5)
    mat_m = material.create(1.0, 0.3, 2)
    mat_i = material.create(10.0, 0.3, 2)
    E = np.array([0., 0., 1.])
    out = {}
    for label, g0 in (('C0=matrix', 1.0), ('C0=midpoint', 5.5)):
        mat0 = material.create(g0, 0.3, 2)
        solver, table = _two_phase_solver(phase, mat_m, mat_i, mat0,
                                          'truncated')
        _sync()
        t0 = time.perf_counter()
        r = solver.solve(E, tol=1e-10, max_iter=200)
        _sync()
        t = time.perf_counter() - t0
        ref_iters = 12
        eps_ref, res_ref = jo.basic_scheme(table[phase, 0], table[phase, 1],
                                           E, 'truncated', g0, 0.3, tol=0.,
                                           max_iter=ref_iters)
        rp = solver.solve(E, tol=0., max_iter=ref_iters)
        e = rp.eps.cpu().numpy()
        rel = float(np.abs(e - eps_ref).max() / np.abs(eps_ref).max())
        res_rel = float(np.max(np.abs(rp.residuals - res_ref) / res_ref))
        tau, _ = solver.local.means(r.eps)
        out[label] = {
            'mu0': g0, 'status': r.status, 'iterations': r.iterations,
            'residual_first': float(r.residuals[0]),
            'residual_last': float(r.residuals[-1]),
            'seconds': round(t, 4),
            'it_per_s': round(r.iterations / t, 1),
            'C_1212': float(g0 + 0.5 * tau[2] / E[2])
            if r.status == 'converged' else None,
            'oracle_%d_iterations' % ref_iters: {
                'strain_rel_err': rel, 'residual_rel_err': res_rel}}
        _log('%s: %s after %d iterations, residual %.3e'
             % (label, r.status, r.iterations, r.residuals[-1]))
        if g0 == 1.0:
            # the stated configuration by conjugate gradients
            solver.solve_cg(E, tol=1e-10, max_iter=3)  # warm-up
            _sync()
            t0 = time.perf_counter()
            rc = solver.solve_cg(E, tol=1e-10, max_iter=500)
            _sync()
            t = time.perf_counter() - t0
            eps_cg, res_cg = jo.cg_scheme(table[phase, 0], table[phase, 1], E,
                                          'truncated', g0, 0.3, tol=1e-10,
                                          max_iter=500)
            tau, _ = solver.local.means(rc.eps)
            out['C0=matrix, CG'] = {
                'status': rc.status, 'iterations': rc.iterations,
                'residual_last': float(rc.residuals[-1]),
                'seconds': round(t, 4),
                'C_1212': float(g0 + 0.5 * tau[2] / E[2]),
                'oracle_cg': {
                    'iterations': len(res_cg),
                    'strain_rel_err': float(
                        np.abs(rc.eps.cpu().numpy() - eps_cg).max() /
                        np.abs(eps_cg).max())}}
            _log('C0=matrix, CG: %s after %d iterations'
                 % (rc.status, rc.iterations))
    # examples/square_inclusion.py as written: corner inclusion, matrix
    # (1, 0.3), inclusion (10, 0.2), C0 (0.9, 0.3), CG on the polarization
    # form with tau2eps = (C - C0)^-1 as the local operator, rtol 1e-10
    from janus_amd.operators import BlockDiagonalOperator2D
    from janus_amd.green import truncated
    import janus_amd.fft.serial as fft
    from janus_amd.solver import LSSolver, delta_stiffness
    mat0 = material.create(0.9, 0.3, 2)
    corner = ms.square_inclusion(n, 0.5, centered=False)
    inv = 1.0 / np.array([delta_stiffness(material.create(g, nu, 2), mat0)
                          for g, nu in ((1.0, 0.3), (10.0, 0.2))])
    pol = LSSolver(truncated(mat0.green_operator(), (n, n), 1.0,
                             fft.create_real((n, n))),
                   BlockDiagonalOperator2D.from_phases(corner, inv))
    pol.solve_cg_polarization(E, tol=1e-10, max_iter=3)  # warm-up
    _sync()
    t0 = time.perf_counter()
    tau, rpol = pol.solve_cg_polarization(E, tol=1e-10, max_iter=1000)
    _sync()
    t = time.perf_counter() - t0
    tau_ref, _, res_ref = jo.cg_polarization_scheme(
        inv[corner, 0], inv[corner, 1], E, 'truncated', 0.9, 0.3, tol=1e-10,
        max_iter=1000)
    out['janus square_inclusion example (polarization CG)'] = {
        'status': rpol.status, 'iterations': rpol.iterations,
        'oracle_iterations': len(res_ref),
        'seconds': round(t, 4),
        'tau_rel_err_vs_oracle': float(
            np.abs(tau.cpu().numpy() - tau_ref).max() /
            np.abs(tau_ref).max())}
    _log('square_inclusion example: %s after %d iterations, %.4f s'
         % (rpol.status, rpol.iterations, t))
    # examples/square_inclusion_scipy.py as written: 1024^2, filtered, a
    # pore (g = 0, nu = 0.2) in the corner, C0 = matrix (1, 0.3), that
    # file's local operator isotropic_4((k - k0) / 2, (g - g0) / 2), GMRES
    # with scipy's defaults (restart 20, rtol 1e-5); it prints the mean
    # strains
    import torch
    from janus_amd.green import filtered
    m_i = material.create(0.0, 0.2, 2)
    m_m = material.create(1.0, 0.3, 2)
    m_0 = material.create(1.0, 0.3, 2)
    ns = 1024
    pore = ms.square_inclusion(ns, 0.5, centered=False)
    tab = np.array([[(m.k - m_0.k) / 2., (m.g - m_0.g) / 2.]
                    for m in (m_m, m_i)])
    gm = LSSolver(filtered(m_0.green_operator(), (ns, ns), 1.0,
                           fft.create_real((ns, ns))),
                  BlockDiagonalOperator2D.from_phases(pore, tab))
    gm.solve_gmres(E, tol=1e-5, max_iter=3)  # warm-up
    _sync()
    t0 = time.perf_counter()
    rg = gm.solve_gmres(E, tol=1e-5, max_iter=2000, restart=20)
    _sync()
    t = time.perf_counter() - t0
    eps = rg.eps
    lhs = eps + torch.as_tensor(gm.green.apply(gm.local.apply(eps)))
    Ef = torch.tensor(E, dtype=torch.float64, device=eps.device)
    true_res = float(torch.linalg.vector_norm(lhs - Ef) /
                     (ns * np.linalg.norm(E)))
    means = eps.reshape(-1, 3).mean(dim=0).cpu().numpy()
    out['janus square_inclusion_scipy example (GMRES)'] = {
        'grid': [ns, ns], 'status': rg.status, 'iterations': rg.iterations,
        'seconds': round(t, 4), 'true_relative_residual': true_res,
        'avg_eps': [float(v) for v in means]}
    _log('square_inclusion_scipy example: %s after %d iterations, %.3f s, '
         'true residual %.2e' % (rg.status, rg.iterations, t, true_res))
    _line('2D 256^2 basic scheme (square inclusion, contrast 10)',
          out['C0=midpoint']['it_per_s'], 'iterations/s',
          {'workload': '2D 256^2 centred square vf 0.25, mu 1/10 nu 0.3, '
                       'truncated Gamma0, E=(0,0,1), tol 1e-10, max 200',
           'runs': out})


def workload_stiffness(args):
    """128^3 Boolean spheres r = 8, vf ~ 0.2, seed 2014; matrix mu 1 nu 0.3,
    inclusions mu 50 nu 0.25; filtered Gamma0; the six Mandel unit loads
    one after another and as one batch at 1e-10 -> the 6x6 C_eff.  C0 =
    matrix diverges (reported over 30 iterations); the solves use the
    midpoint mu0 = 25.5."""
    import torch
    import janus_amd.material.elastic.linear.isotropic as material
    from janus_amd import microstructure as ms
    from janus_amd.batch import effective_stiffness_batched
    from janus_amd.solver import effective_stiffness
    from oracle import janus_oracle as jo
    n = args.grid or 128
    phase = ms.spheres(n, 8.0 * n / 128, 0.2, seed=2014)
    mat_m = material.create(1.0, 0.3, 3)
    mat_i = material.create(50.0, 0.25, 3)
    mat_div = material.create(1.0, 0.3, 3)
    s_div, _ = _two_phase_solver(phase, mat_m, mat_i, mat_div, 'filtered')
    if not args.no_autotune:
        s_div.autotune(iters=3)
    r_div = s_div.solve(_unit(6, 5), tol=1e-10, max_iter=30)
    # the stated configuration (C0 = matrix) by conjugate gradients
    s_div.solve_cg(_unit(6, 0), tol=1e-10, max_iter=3)  # warm-up
    C_cg = np.zeros((6, 6))
    its_cg = []
    _sync()
    t0 = time.perf_counter()
    for j in range(6):
        rc = s_div.solve_cg(_unit(6, j), tol=1e-10, max_iter=2000)
        its_cg.append(rc.iterations)
        tau, _ = s_div.local.means(rc.eps)
        from janus_amd.solver import reference_stiffness
        C_cg[:, j] = reference_stiffness(mat_div).dot(_unit(6, j)) + tau
    _sync()
    t_cg = time.perf_counter() - t0
    _log('CG, C0 = matrix, 6 loads: %.3f s, iterations %s' % (t_cg, its_cg))
    # the same six loads by restarted GMRES (no symmetry needed)
    s_div.solve_gmres(_unit(6, 0), tol=1e-10, max_iter=3)  # warm-up
    C_gm = np.zeros((6, 6))
    its_gm = []
    _sync()
    t0 = time.perf_counter()
    for j in range(6):
        rg = s_div.solve_gmres(_unit(6, j), tol=1e-10, max_iter=2000,
                               restart=args.gmres_restart)
        its_gm.append(rg.iterations)
        tau, _ = s_div.local.means(rg.eps)
        C_gm[:, j] = reference_stiffness(mat_div).dot(_unit(6, j)) + tau
    _sync()
    t_gm = time.perf_counter() - t0
    _log('GMRES(%d), C0 = matrix, 6 loads: %.3f s, iterations %s'
         % (args.gmres_restart, t_gm, its_gm))
    del s_div
    mat0 = material.create(25.5, 0.3, 3)
    solver, table = _two_phase_solver(phase, mat_m, mat_i, mat0, 'filtered')
    if not args.no_autotune:
        solver.autotune(iters=3)
    # parity: first iterations vs the oracle (load e_xy)
    k = 3
    eps_ref, res_ref = jo.basic_scheme(table[phase, 0], table[phase, 1],
                                       _unit(6, 5), 'filtered', 25.5, 0.3,
                                       tol=0., max_iter=k)
    rp = solver.solve(_unit(6, 5), tol=0., max_iter=k)
    rel = float(np.abs(rp.eps.cpu().numpy() - eps_ref).max() /
                np.abs(eps_ref).max())
    solver.solve(_unit(6, 0), tol=0., max_iter=3)  # warm-up
    _sync()
    t0 = time.perf_counter()
    C_seq, its, sts = effective_stiffness(solver, mat0, tol=1e-10,
                                          max_iter=5000)
    _sync()
    t_seq = time.perf_counter() - t0
    _log('sequential 6 loads: %.3f s, iterations %s' % (t_seq, its))
    t0 = time.perf_counter()
    Cs, its_b, sts_b, batch = effective_stiffness_batched(
        [solver], [mat0], tol=1e-10, max_iter=5000, autotune=True)
    _sync()
    t_b = time.perf_counter() - t0
    C_b = Cs[0]
    _log('batched 6 loads: %.3f s, widths %s' % (t_b, batch.groups()))
    total_it = int(sum(its))
    B = five_pass_bytes((n, n, n))
    _line('6x6 effective stiffness at %d^3 (six load cases, 1e-10)' % n,
          round(t_seq, 4), 's', {
              'workload': '%d^3 spheres r=%g vf~0.2 seed 2014, mu 1/50 nu '
                          '0.3/0.25, filtered Gamma0, C0 mu0=25.5 nu0=0.3, '
                          'Mandel unit loads, tol 1e-10' % (n, 8.0 * n / 128),
              'C_eff': np.round(C_seq, 10).tolist(),
              'iterations': its, 'status': sts,
              'seconds_sequential': round(t_seq, 4),
              'seconds_batched': round(t_b, 4),
              'batched_widths': [g[2] for g in batch.groups()],
              'batched_equal_bitwise': bool(np.array_equal(C_b, C_seq)),
              'batched_iterations': its_b[0],
              'iterations_per_s': round(total_it / t_seq, 1),
              'iteration_GBps_B_alg': round(B * total_it / t_seq / 1e9, 1),
              'symmetry_rel': float(np.abs(C_seq - C_seq.T).max() /
                                    np.abs(C_seq).max()),
              'C0_matrix_divergence': {
                  'status': r_div.status,
                  'residuals': [float(x) for x in r_div.residuals[::5]]},
              'C0_matrix_cg': {
                  'C_eff': np.round(C_cg, 10).tolist(),
                  'iterations': its_cg,
                  'seconds': round(t_cg, 4),
                  'symmetry_rel': float(np.abs(C_cg - C_cg.T).max() /
                                        np.abs(C_cg).max()),
                  'rel_diff_to_midpoint_basic_C_eff': float(
                      np.abs(C_cg - C_seq).max() / np.abs(C_seq).max())},
              'C0_matrix_gmres': {
                  'restart': args.gmres_restart,
                  'iterations': its_gm,
                  'seconds': round(t_gm, 4),
                  'rel_diff_to_cg_C_eff': float(
                      np.abs(C_gm - C_cg).max() / np.abs(C_cg).max())},
              'oracle_%d_iterations_strain_rel_err' % k: rel},
          higher=False)


def iteration_bytes(shape, fused, sym=6, m=1):
    """Algorithmic HBM bytes of one iteration of a solve on a 3D grid in its
    schedule: fused (96 + m) N + 768 Nh, five passes (144 + m) N + 768 Nh."""
    N = int(np.prod(shape))
    Nh = int(np.prod(shape[:-1])) * (shape[-1] // 2 + 1)
    return ((2 if fused else 3) * 8 * sym + m) * N + 8 * 16 * sym * Nh


def cg_iteration_bytes(shape, sym=6, m=1):
    """Algorithmic HBM bytes of one fused conjugate-gradient iteration on a
    3D grid: 768 Nh for the four spectrum passes, and per voxel pass A
    (r, p, x read, p, x written, phase id), pass E (p read, q written) and
    the vector kernel (r, q read, r written): (10 * 8 sym + m) N."""
    N = int(np.prod(shape))
    Nh = int(np.prod(shape[:-1])) * (shape[-1] // 2 + 1)
    return (10 * 8 * sym + m) * N + 8 * 16 * sym * Nh


def workload_batched(args, emit=True):
    """32 seeded microstructures, n ~ U{64, 80, 96, 128, 160}, Boolean
    spheres r = 8, vf ~ U[0.1, 0.35], inclusion/matrix contrast ~
    logU[2, 100] (matrix mu 1 nu 0.3, inclusions nu 0.25), filtered Gamma0,
    midpoint C0, six Mandel unit loads each: 192 right-hand sides at 1e-10,
    one autotuned batch vs 192 sequential solves.  The bound printed beside
    them is the entries' algorithmic bytes (sum over entries of iterations x
    bytes per iteration in the solver's schedule) over the measured copy
    bandwidth: the time of a batch that streams every byte at copy speed."""
    import torch
    import janus_amd.material.elastic.linear.isotropic as material
    from janus_amd import microstructure as ms
    from janus_amd.batch import BatchSolver
    cases = ms.batched_cases(count=args.batch_count)
    solvers = []
    for n, vf, c, seed in cases:
        phase = ms.spheres(n, 8.0, vf, seed=seed)
        mat_m = material.create(1.0, 0.3, 3)
        mat_i = material.create(c, 0.25, 3)
        mat0 = material.create(0.5 * (1.0 + c), 0.3, 3)
        S = _two_phase_solver(phase, mat_m, mat_i, mat0, 'filtered')[0]
        if not args.no_autotune:
            S.autotune(iters=3)
        solvers.append(S)
    _log('built %d microstructures' % len(solvers))
    rhs = [(i, j) for i in range(len(solvers)) for j in range(6)]

    def fields():
        out = []
        for i, j in rhs:
            e = torch.empty(solvers[i].shape, dtype=torch.float64,
                            device='cuda')
            e[...] = torch.from_numpy(_unit(6, j)).cuda()
            out.append(e)
        return out
    # warm-up: every grid size once
    seen = set()
    for S in solvers:
        if S.shape not in seen:
            seen.add(S.shape)
            S.solve(_unit(6, 0), tol=0., max_iter=2)
    eps_seq = fields()
    _sync()
    t0 = time.perf_counter()
    seq = []
    for (i, j), e in zip(rhs, eps_seq):
        r = solvers[i].solve(_unit(6, j), tol=1e-10, max_iter=args.max_iter,
                             eps=e)
        seq.append((r.iterations, r.status))
    _sync()
    t_seq = time.perf_counter() - t0
    _log('sequential: %.3f s, %d iterations' % (t_seq,
                                                sum(x[0] for x in seq)))
    eps_b = fields()
    batch = BatchSolver(iters_per_launch=args.iters_per_launch,
                        graph=args.graph,
                        autotune=True, streams=args.batch_streams)
    for (i, j), e in zip(rhs, eps_b):
        batch.add(solvers[i], _unit(6, j), tol=1e-10, max_iter=args.max_iter,
                  eps=e)
    _sync()
    t0 = time.perf_counter()
    res = batch.run()
    _sync()
    t_b = time.perf_counter() - t0
    ms_run, replays, captures = batch.stats()
    _log('batched: %.3f s (%d replays, %d captures)' % (t_b, replays,
                                                         captures))
    _log('tuning (width, ms/entry-it) per group: %s' % batch.tuning())
    per_grid = {}
    if args.per_group:
        # each grid size alone: sequential solves vs its own batch
        for n in sorted(set(c[0] for c in cases)):
            idx = [k for k, (i, j) in enumerate(rhs) if cases[i][0] == n]
            t_s = 0.
            for k in idx:
                i, j = rhs[k]
                eps_seq[k][...] = torch.from_numpy(_unit(6, j)).cuda()
                _sync()
                t0 = time.perf_counter()
                solvers[i].solve(_unit(6, j), tol=1e-10,
                                 max_iter=args.max_iter, eps=eps_seq[k])
                _sync()
                t_s += time.perf_counter() - t0
            bg = BatchSolver(iters_per_launch=args.iters_per_launch,
                             graph=args.graph, autotune=True)
            for k in idx:
                i, j = rhs[k]
                eps_seq[k][...] = torch.from_numpy(_unit(6, j)).cuda()
                bg.add(solvers[i], _unit(6, j), tol=1e-10,
                       max_iter=args.max_iter, eps=eps_seq[k])
            _sync()
            t0 = time.perf_counter()
            bg.run()
            _sync()
            t_g = time.perf_counter() - t0
            its_g = sum(seq[k][0] for k in idx)
            _log('grid %d tuning (width, ms/entry-it): %s'
                 % (n, bg.tuning()))
            per_grid[n] = {'entries': len(idx), 'iterations': its_g,
                           'seconds_sequential': round(t_s, 4),
                           'seconds_batched': round(t_g, 4),
                           'speedup': round(t_s / t_g, 3),
                           'widths': [g[2] for g in bg.groups()],
                           'ms_per_entry_iteration_sequential':
                               round(1e3 * t_s / its_g, 4),
                           'ms_per_entry_iteration_batched':
                               round(1e3 * t_g / its_g, 4)}
            _log('grid %d: %s' % (n, per_grid[n]))
            del bg
    equal = sum(bool(torch.equal(a, b)) for a, b in zip(eps_seq, eps_b))
    maxdiff = max(float((a - b).abs().max()) for a, b in zip(eps_seq, eps_b))
    same_iters = all(r.iterations == s[0] for r, s in zip(res, seq))
    its = [s[0] for s in seq]
    conv = sum(s[1] == 'converged' for s in seq)
    groups = [{'grid': list(g[0]), 'entries': g[1], 'width': g[2]}
              for g in batch.groups()]
    del eps_seq, eps_b, batch
    torch.cuda.empty_cache()
    # the same 192 right-hand sides by conjugate gradients: sequential
    # solve_cg against one batch of CG entries
    cg = {}
    if not args.no_batched_cg:
        for S in solvers[:1]:
            S.solve_cg(_unit(6, 0), tol=0., max_iter=2)  # warm-up
        eps_cs = fields()
        _sync()
        t0 = time.perf_counter()
        seq_cg = []
        for (i, j), e in zip(rhs, eps_cs):
            r = solvers[i].solve_cg(_unit(6, j), tol=1e-10,
                                    max_iter=args.max_iter, eps=e)
            seq_cg.append((r.iterations, r.status))
        _sync()
        t_seq_cg = time.perf_counter() - t0
        eps_cb = fields()
        bcg = BatchSolver(iters_per_launch=args.iters_per_launch,
                          graph=args.graph, autotune=True,
                          streams=args.batch_streams)
        for (i, j), e in zip(rhs, eps_cb):
            bcg.add(solvers[i], _unit(6, j), tol=1e-10,
                    max_iter=args.max_iter, eps=e, method='cg')
        _sync()
        t0 = time.perf_counter()
        res_cg = bcg.run()
        _sync()
        t_b_cg = time.perf_counter() - t0
        eq_cg = sum(bool(torch.equal(a, b)) for a, b in zip(eps_cs, eps_cb))
        same_cg = all(r.iterations == q[0] for r, q in zip(res_cg, seq_cg))
        its_cg = [q[0] for q in seq_cg]
        _log('CG: sequential %.3f s, batched %.3f s (%d iterations); '
             'batched CG vs sequential fixed point %.2fx'
             % (t_seq_cg, t_b_cg, sum(its_cg), t_seq / t_b_cg))
        del eps_cs, eps_cb, bcg
        torch.cuda.empty_cache()
        cg = {'seconds_sequential': round(t_seq_cg, 4),
              'seconds_batched': round(t_b_cg, 4),
              'speedup': round(t_seq_cg / t_b_cg, 3),
              'speedup_vs_sequential_fixed_point': round(t_seq / t_b_cg, 3),
              'iterations_total': int(sum(its_cg)),
              'converged': sum(q[1] == 'converged' for q in seq_cg),
              'entries_bitwise_equal': eq_cg,
              'same_iteration_counts': same_cg,
              'bytes_alg_total': int(sum(
                  it * cg_iteration_bytes(solvers[i].shape[:-1])
                  for (i, j), it in zip(rhs, its_cg)))}
    copy_gbps = stream_copy_gbps()
    total_bytes = sum(it * iteration_bytes(solvers[i].shape[:-1],
                                           solvers[i].fused)
                      for (i, j), it in zip(rhs, its))
    bound = total_bytes / (copy_gbps * 1e9)
    bound_peak = total_bytes / (HBM_PEAK_GBS * 1e9)
    _log('bound: %.3e B over %.1f GB/s copy = %.3f s (%.3f s at 8 TB/s); '
         'sequential %.3f s, batched %.3f s'
         % (total_bytes, copy_gbps, bound, bound_peak, t_seq, t_b))
    summary = {
        'speedup': round(t_seq / t_b, 3),
        'seconds_sequential': round(t_seq, 4),
        'seconds_batched': round(t_b, 4),
        'bound_seconds_copy_bw': round(bound, 4),
        'bound_seconds_8TBps': round(bound_peak, 4),
        'bytes_alg_total': int(total_bytes),
        'copy_GBps': round(copy_gbps, 1),
        'speedup_bound_copy_bw': round(t_seq / bound, 3),
        'batched_over_bound': round(t_b / bound, 3),
        'entries': len(rhs),
        'entries_bitwise_equal': equal,
        'same_iteration_counts': same_iters,
        'converged': conv,
        'iterations_total': int(sum(its)),
        'streams': args.batch_streams,
    }
    if cg:
        cg['bound_seconds_copy_bw'] = round(
            cg['bytes_alg_total'] / (copy_gbps * 1e9), 4)
        cg['speedup_bound_copy_bw'] = round(
            cg['seconds_sequential'] / cg['bound_seconds_copy_bw'], 3)
        cg['batched_over_bound'] = round(
            cg['seconds_batched'] / cg['bound_seconds_copy_bw'], 3)
        summary['cg'] = cg
    if emit:
        _line('batched vs sequential plain solves, %d right-hand sides (%d '
              'microstructures x 6 loads) at 1e-10' % (len(rhs), len(solvers)),
              round(t_seq / t_b, 3), 'x (speedup)', {
                  'workload': '%d seeded microstructures, n~U{64,80,96,128,'
                              '160}, spheres r=8, vf~U[0.1,0.35], contrast~'
                              'logU[2,100], filtered Gamma0, midpoint C0, 6 '
                              'Mandel unit loads, tol 1e-10, max_iter %d'
                              % (len(solvers), args.max_iter),
                  'grids': [c[0] for c in cases],
                  'contrasts': [round(c[2], 3) for c in cases],
                  'summary': summary,
                  'iterations_min_max': [int(min(its)), int(max(its))],
                  'replays': replays, 'graph_captures': captures,
                  'iters_per_launch': args.iters_per_launch,
                  'streams': args.batch_streams,
                  'groups': groups,
                  'per_grid': per_grid,
                  'parity': {'entries_bitwise_equal': equal,
                             'entries': len(rhs),
                             'max_abs_diff': maxdiff,
                             'same_iteration_counts': same_iters}})
    return summary


def workload_adjoint(args, emit=True):
    """128^3 SIMP density: rho = periodic Gaussian-smoothed noise (sigma 4
    voxels, seed 2014) mapped to [0, 1], mu = 1 + 19 rho^3, nu = 0.3,
    midpoint C0 (mu0 = 10.5), filtered Gamma0, E = e_xy, L = (C_1212 - 3)^2.
    Time of the backward (Gamma0 g, adjoint solve, contraction) over the
    forward solve, both at 1e-10, for conjugate gradients (ls_solve's
    default) and the basic scheme; the same loss through effective_stress;
    <dL/drho, v> against a central difference of L along a smooth seeded
    direction v (step 1e-6 relative, solves at 1e-14)."""
    import torch
    import janus_amd.fft.serial as fft
    import janus_amd.material.elastic.linear.isotropic as material
    from janus_amd import microstructure as ms
    from janus_amd.autograd import (effective_stress, iso_delta, ls_solve,
                                    mean_stress)
    from janus_amd.green import filtered
    from janus_amd.operators import BlockDiagonalOperator3D
    from janus_amd.solver import LSSolver, reference_stiffness
    n = args.grid or 128
    rho_np = ms.smoothed_density(n, 4.0, seed=2014)
    mat0 = material.create(10.5, 0.3, 3)
    green = filtered(mat0.green_operator(), (n,) * 3, 1.0,
                     fft.create_real((n,) * 3))
    E = torch.zeros(6, dtype=torch.float64, device='cuda')
    E[5] = 1.0
    if not args.no_autotune:
        # tune the Green operator's plans on the same grid
        z = torch.zeros((n,) * 3, dtype=torch.float64, device='cuda')
        LSSolver(green, BlockDiagonalOperator3D.from_fields(z, z)).autotune(3)

    def moduli(rho):
        g = 1.0 + rho ** 3 * 19.0
        return iso_delta(g, torch.full_like(rho, 0.3), mat0, 3)

    def loss_fn(rho, tol, max_iter, info=None, method='cg'):
        sph, dev = moduli(rho)
        eps = ls_solve(sph, dev, E, green, tol=tol, max_iter=max_iter,
                       info=info, method=method)
        sigma = mean_stress(sph, dev, eps, E, mat0)
        return (0.5 * sigma[5] / E[5] - 3.0) ** 2

    def loss_es(rho, tol, max_iter, info=None, method='cg'):
        sph, dev = moduli(rho)
        sigma = effective_stress(sph, dev, E, green, mat0, tol=tol,
                                 max_iter=max_iter, info=info, method=method)
        return (0.5 * sigma[5] / E[5] - 3.0) ** 2

    def timed(fn, method):
        with torch.no_grad():  # warm-up
            fn(torch.tensor(rho_np, device='cuda'), 0., 3, method=method)
        times = []
        for rep in range(args.adjoint_reps):
            rho = torch.tensor(rho_np, device='cuda', requires_grad=True)
            info = {}
            _sync()
            t0 = time.perf_counter()
            loss = fn(rho, 1e-10, 5000, info, method)
            _sync()
            t1 = time.perf_counter()
            loss.backward()
            _sync()
            t2 = time.perf_counter()
            times.append((t1 - t0, t2 - t1))
        t_fwd, t_bwd = min(r[0] for r in times), min(r[1] for r in times)
        return rho.grad.cpu().numpy(), info, t_fwd, t_bwd, float(loss.item())

    runs, grads = {}, {}
    for method in ('cg', 'basic'):
        g, info, t_fwd, t_bwd, L = timed(loss_fn, method)
        fwd, adj = info['forward'], info['adjoint']
        _log('%s: forward %.4f s (%d it), backward %.4f s (adjoint %d it)'
             % (method, t_fwd, fwd.iterations, t_bwd, adj.iterations))
        grads[method] = g
        runs[method] = {'seconds_forward': round(t_fwd, 4),
                        'seconds_backward': round(t_bwd, 4),
                        'ratio': round(t_bwd / t_fwd, 4),
                        'forward_iterations': fwd.iterations,
                        'adjoint_iterations': adj.iterations,
                        'loss': L}
    # effective_stress: dL/d<sigma> is parallel to E here, so its backward
    # reuses the forward field (no adjoint solve)
    runs_es = {}
    for method in ('cg', 'basic'):
        g, info, t_fwd, t_bwd, L = timed(loss_es, method)
        runs_es[method] = {
            'seconds_forward': round(t_fwd, 4),
            'seconds_backward': round(t_bwd, 5),
            'ratio': round(t_bwd / t_fwd, 4),
            'backward_solves': 0 if info['adjoint'] is None else 1,
            'grad_vs_ls_solve_max_rel': float(
                np.abs(g - grads[method]).max() /
                np.abs(grads[method]).max())}
        _log('effective_stress %s: forward %.4f s, backward %.5f s'
             % (method, t_fwd, t_bwd))
    # directional central difference along a smooth seeded direction
    v = ms.smoothed_density(n, 4.0, seed=7) - 0.5
    h = 1e-6 / np.abs(v).max()
    C0 = reference_stiffness(mat0)
    Eh = _unit(6, 5)

    def loss_at(r_np):
        sph, dev = moduli(torch.tensor(r_np, device='cuda'))
        local = BlockDiagonalOperator3D.from_fields(sph, dev)
        r = LSSolver(green, local).solve_cg(Eh, tol=1e-14, max_iter=400)
        tau, _ = local.means(r.eps)
        return (0.5 * (C0.dot(Eh) + tau)[5] - 3.0) ** 2

    fd = (loss_at(rho_np + h * v) - loss_at(rho_np - h * v)) / (2 * h)
    dd = float((grads['cg'] * v).sum())
    fd_check = {'direction': 'smoothed noise sigma 4 seed 7, minus 0.5',
                'step': h, 'central_difference': fd,
                'adjoint_directional': dd,
                'rel_err': abs(fd - dd) / abs(dd)}
    _log('directional derivative %.10g, central difference %.10g'
         % (dd, fd))
    summary = {
        'value': runs['cg']['ratio'],
        'method': 'cg (ls_solve default)',
        'cg': runs['cg'],
        'basic_scheme': runs['basic'],
        'effective_stress': runs_es,
        'grad_cg_vs_basic_max_rel': float(
            np.abs(grads['cg'] - grads['basic']).max() /
            np.abs(grads['basic']).max()),
        'directional_fd': fd_check,
    }
    if emit:
        _line('adjoint gradient time / forward solve time at %d^3 (1e-10), '
              'conjugate gradients' % n,
              runs['cg']['ratio'], 'ratio', {
                  'workload': '%d^3 SIMP rho smoothed noise sigma 4 seed '
                              '2014, mu=1+19 rho^3 nu 0.3, filtered Gamma0, '
                              'mu0 10.5, E=e_xy, L=(C_1212-3)^2, tol 1e-10'
                              % n,
                  'summary': summary},
              higher=False)
    return summary


def workload_large(args, emit=True):
    """512^3, three phases: matrix mu 1 and randomly oriented 1:1:4
    ellipsoids (semi-axes 8, 8, 32) of mu 10 and 0.1 (nu 0.3), filtered
    Gamma0, midpoint C0 (mu0 = 5.05).  Iterations/s of the six Mandel loads
    as one batch, and one adjoint gradient of L = C_eff[5, 5] (E = e_xy)
    with respect to the three phase moduli, checked against effective_stress
    and (--fd-voxels > 0) a central difference on the mu of phase 1."""
    import torch
    import janus_amd.fft.serial as fft
    import janus_amd.material.elastic.linear.isotropic as material
    from janus_amd import microstructure as ms
    from janus_amd.autograd import (effective_stress, iso_delta, ls_solve,
                                    mean_stress)
    from janus_amd.batch import BatchSolver
    from janus_amd.green import filtered
    from janus_amd.operators import BlockDiagonalOperator3D
    from janus_amd.solver import LSSolver, delta_stiffness
    n = args.grid or 512
    t0 = time.perf_counter()
    phase_d, count, axes = ms.large_three_phase(n)
    frac = torch.bincount(phase_d.flatten().long(), minlength=3).double() \
        / phase_d.numel()
    _log('phase map %d^3 (%d ellipsoids) in %.1f s, fractions %s'
         % (n, count, time.perf_counter() - t0, frac.tolist()))
    phase = phase_d.cpu().numpy()
    del phase_d
    mus = [1.0, 10.0, 0.1]
    mats = [material.create(m, 0.3, 3) for m in mus]
    mat0 = material.create(0.5 * (10.0 + 0.1), 0.3, 3)
    table = np.array([delta_stiffness(m, mat0) for m in mats])
    green = filtered(mat0.green_operator(), (n,) * 3, 1.0,
                     fft.create_real((n,) * 3))
    local = BlockDiagonalOperator3D.from_phases(phase, table)
    solver = LSSolver(green, local)
    if not args.no_autotune:
        _log('autotune %s' % json.dumps(solver.autotune(iters=3)))
    # six loads as one batch: fixed iteration count, iterations/s
    k = args.large_iters
    batch = BatchSolver(iters_per_launch=k, graph=False)
    fields = []
    for j in range(6):
        e = torch.empty(solver.shape, dtype=torch.float64, device='cuda')
        fields.append(e)
        batch.add(solver, _unit(6, j), tol=0., max_iter=k, eps=e)
    t_b = None
    for rep in range(2):  # warm-up, then timed; each restarts from E
        for j, e in enumerate(fields):
            e[...] = torch.from_numpy(_unit(6, j)).cuda()
        _sync()
        t0 = time.perf_counter()
        res = batch.run()
        _sync()
        t_b = time.perf_counter() - t0
    assert all(r.iterations == k for r in res)
    its = 6 * k / t_b
    B = five_pass_bytes((n, n, n))
    _log('6-load batch: %d iterations each in %.3f s (%.2f load-it/s)'
         % (k, t_b, its))
    del batch, fields, res
    torch.cuda.empty_cache()
    # adjoint gradient w.r.t. the phase moduli through the field operator
    ph = torch.from_numpy(phase.astype(np.int64)).cuda()
    E = torch.zeros(6, dtype=torch.float64, device='cuda')
    E[5] = 1.0

    # mu(x) = sum_p mu_p [phase(x) = p]: its backward is three reductions
    # (an indexed gather would scatter-add 134 M values into three bins)
    masks = [(ph == p).double() for p in range(len(mus))]
    del ph

    def moduli(mu_t):
        g = masks[0] * mu_t[0]
        for p in range(1, len(mus)):
            g = g + masks[p] * mu_t[p]
        return iso_delta(g, torch.full_like(g, 0.3), mat0, 3)

    def loss_fn(mu_t, tol, max_iter, info=None):
        sph, dev = moduli(mu_t)
        eps = ls_solve(sph, dev, E, green, tol=tol, max_iter=max_iter,
                       info=info, method=args.method)
        return mean_stress(sph, dev, eps, E, mat0)[5]

    def loss_es(mu_t, tol, max_iter, info=None):
        sph, dev = moduli(mu_t)
        return effective_stress(sph, dev, E, green, mat0, tol=tol,
                                max_iter=max_iter, info=info,
                                method=args.method)[5]

    def timed(fn):
        mu_t = torch.tensor(mus, dtype=torch.float64, device='cuda',
                            requires_grad=True)
        info = {}
        _sync()
        t0 = time.perf_counter()
        L = fn(mu_t, 1e-10, args.max_iter, info)
        _sync()
        t1 = time.perf_counter()
        L.backward()
        _sync()
        t2 = time.perf_counter()
        return mu_t.grad.cpu().numpy(), info, t1 - t0, t2 - t1, float(L.item())

    grad, info, t_f, t_bk, Lv = timed(loss_fn)
    _log('forward %.2f s (%d it), backward %.2f s (%d it), grad %s'
         % (t_f, info['forward'].iterations, t_bk,
            info['adjoint'].iterations, grad))
    torch.cuda.empty_cache()
    grad_es, info_es, t_fe, t_be, _ = timed(loss_es)
    _log('effective_stress: forward %.2f s, backward %.4f s, grad %s'
         % (t_fe, t_be, grad_es))
    fd = None
    if args.fd_voxels > 0:
        hh = 1e-4
        vals = []
        for sgn in (1, -1):
            m = torch.tensor(mus, dtype=torch.float64, device='cuda')
            m[1] += sgn * hh
            with torch.no_grad():
                vals.append(loss_fn(m, 1e-10, args.max_iter).item())
        d = (vals[0] - vals[1]) / (2 * hh)
        fd = {'phase': 1, 'step': hh, 'fd': d, 'adjoint': float(grad[1]),
              'rel_err': abs(d - grad[1]) / abs(grad[1])}
    del masks
    torch.cuda.empty_cache()
    summary = {
        'value': round(its, 3),
        'unit': 'load-iterations/s',
        'iterations_per_load': k,
        'B_alg_per_load_iteration': int(B),
        'GBps_B_alg': round(B * its / 1e9, 1),
        'roofline_frac_B_alg': round(B * its / 1e9 / HBM_PEAK_GBS, 4),
        'phase_fractions': [round(x, 4) for x in frac.tolist()],
        'adjoint': {
            'method': args.method,
            'loss': 'C_eff[5,5] under E=e_xy',
            'value': Lv,
            'dL_dmu_phase': grad.tolist(),
            'seconds_forward': round(t_f, 3),
            'seconds_backward': round(t_bk, 3),
            'ratio': round(t_bk / t_f, 4),
            'forward_iterations': info['forward'].iterations,
            'adjoint_iterations': info['adjoint'].iterations,
            'fd_check': fd},
        'effective_stress': {
            'seconds_forward': round(t_fe, 3),
            'seconds_backward': round(t_be, 4),
            'ratio': round(t_be / t_fe, 4),
            'dL_dmu_phase': grad_es.tolist(),
            'grad_vs_adjoint_max_rel': float(
                np.abs(grad_es - grad).max() / np.abs(grad).max())},
    }
    if emit:
        _line('6-load batch LS iterations/s at %d^3 FP64 (three phases)' % n,
              round(its, 3), 'load-iterations/s', {
                  'workload': '%d^3, %d ellipsoids 1:1:4 (semi-axes %g %g '
                              '%g), mu 1/10/0.1 nu 0.3, filtered Gamma0, mu0 '
                              '5.05, 6 Mandel unit loads as one batch, %d '
                              'iterations' % (n, count, axes[0], axes[1],
                                              axes[2], k),
                  'summary': summary})
    return summary


def workload_fft(args):
    """examples/fftw_python_benchmark_serial.py: r2c of a uniform(-1, 1)
    field (seed 20140121), mean and standard deviation of the time per call
    in ms, at 128^3, 256^3 and 512^3 -- on device tensors (data resident,
    each call bracketed by synchronizations as the script brackets its
    calls with perf_counter) and on numpy arrays (copied in and out through
    PCIe, as the janus API is called).  The script asks for 15000 / 10000 /
    1000 calls; fewer are timed here."""
    import torch
    import janus_amd.fft.serial as fft
    np.random.seed(20140121)
    rows = {}
    for shape, nd, nh in (((128,) * 3, 200, 20), ((256,) * 3, 100, 5),
                          ((512,) * 3, 20, 2)):
        t = fft.create_real(shape)
        r = np.random.uniform(-1., 1., t.ishape)
        rd = torch.from_numpy(r).cuda()
        cd = torch.empty(t.oshape, dtype=torch.float64, device='cuda')
        t.r2c(rd, cd)
        times = []
        for _ in range(nd):
            _sync()
            t0 = time.perf_counter()
            t.r2c(rd, cd)
            _sync()
            times.append(1e3 * (time.perf_counter() - t0))
        back = torch.empty_like(rd)
        t.c2r(cd, back)
        times_inv = []
        for _ in range(nd):
            _sync()
            t0 = time.perf_counter()
            t.c2r(cd, back)
            _sync()
            times_inv.append(1e3 * (time.perf_counter() - t0))
        roundtrip = float((back - rd).abs().max())
        del back
        # GPU time per call with calls queued back to back (no host
        # synchronization between them), against the bytes the three passes
        # move: the last-axis pass reads N and writes Nh complex values, each
        # strided pass reads and writes Nh complex values
        reps = 4 * nd
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        _sync()
        ev0.record()
        for _ in range(reps):
            t.r2c(rd, cd)
        ev1.record()
        _sync()
        ms_queued = ev0.elapsed_time(ev1) / reps
        ncell = int(np.prod(shape))
        nhalf = ncell // shape[-1] * (shape[-1] // 2 + 1)
        pass_bytes = 8 * ncell + 16 * nhalf + (len(shape) - 1) * 32 * nhalf
        c = np.empty(t.oshape)
        t.r2c(r, c)
        host = []
        for _ in range(nh):
            t0 = time.perf_counter()
            t.r2c(r, c)
            host.append(1e3 * (time.perf_counter() - t0))
        nbytes = 8 * (r.size + int(np.prod(t.oshape)))
        rows['%d^3' % shape[0]] = {
            'device_ms_mean': round(float(np.mean(times)), 4),
            'device_ms_std': round(float(np.std(times)), 4),
            'device_GBps_in_plus_out': round(
                nbytes / (np.mean(times) * 1e-3) / 1e9, 1),
            'device_c2r_ms_mean': round(float(np.mean(times_inv)), 4),
            'device_c2r_GBps_in_plus_out': round(
                nbytes / (np.mean(times_inv) * 1e-3) / 1e9, 1),
            'c2r_r2c_roundtrip_max_abs_err': roundtrip,
            'device_ms_queued': round(ms_queued, 4),
            'pass_bytes': pass_bytes,
            'pass_GBps_queued': round(pass_bytes / (ms_queued * 1e-3) / 1e9,
                                      1),
            'pass_frac_of_8TBps_queued': round(
                pass_bytes / (ms_queued * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
            'host_arrays_ms_mean': round(float(np.mean(host)), 2),
            'host_arrays_ms_std': round(float(np.std(host)), 2),
            'calls': [nd, nh]}
        _log('r2c %s: device %.3f ms (c2r %.3f ms; queued %.3f ms, %.0f '
             'GB/s on the three passes), numpy arrays %.1f ms'
             % (shape, np.mean(times), np.mean(times_inv), ms_queued,
                pass_bytes / (ms_queued * 1e-3) / 1e9, np.mean(host)))
        del rd, cd
        torch.cuda.empty_cache()
    copy = stream_copy_gbps()
    _line('r2c ms per call at 256^3 (janus fftw_python_benchmark_serial)',
          rows['256^3']['device_ms_mean'], 'ms',
          {'workload': 'r2c of uniform(-1, 1) fields, seed 20140121, '
                       '128^3 / 256^3 / 512^3', 'shapes': rows,
           'stream_copy_GBps': round(copy, 1)},
          higher=False)


def workload_patterns(args):
    """Memory rate of the strided passes' access pattern with no arithmetic
    (jx_tile_probe): the 6-component padded spectrum of a 256^3 and a 512^3
    grid copied along axis 0 (the Gamma-hat pass) and axis 1 (passes B and
    D) in tiles of 1, 2 and 4 points, read + write bytes over hipEvent time,
    beside the stream copy."""
    import ctypes
    import torch
    from janus_amd import _lib
    lib = _lib.lib()
    stream = _lib.current_stream()
    table = {}
    for n in (256, 512):
        shape = (ctypes.c_int64 * 3)(n, n, n)
        nbytes = ctypes.c_int64()
        _lib.check(lib.jx_tile_probe_bytes(shape, ctypes.byref(nbytes)))
        a = torch.ones(nbytes.value // 8, dtype=torch.float64, device='cuda')
        b = torch.zeros_like(a)
        moved = 2 * 96 * n * n * (n // 2 + 1)
        for axis in (0, 1):
            for pts in (1, 2, 4):
                def run():
                    _lib.check(lib.jx_tile_probe(shape, axis, pts, a.data_ptr(),
                                                 b.data_ptr(), nbytes.value,
                                                 stream))
                run()
                _sync()
                ev0 = torch.cuda.Event(enable_timing=True)
                ev1 = torch.cuda.Event(enable_timing=True)
                reps = 10
                ev0.record()
                for _ in range(reps):
                    run()
                ev1.record()
                _sync()
                ms = ev0.elapsed_time(ev1) / reps
                key = '%d^3 axis %d, %d-point tiles' % (n, axis, pts)
                table[key] = {'ms': round(ms, 4),
                              'GBps': round(moved / (ms * 1e-3) / 1e9, 1)}
                _log('%s: %.4f ms, %.0f GB/s' % (key, ms,
                                                  moved / (ms * 1e-3) / 1e9))
        # every spectrum value copied (the slab padding stays zero)
        table['%d^3 all values copied' % n] = (
            float(b.sum()) == 12. * n * n * (n // 2 + 1))
        del a, b
        torch.cuda.empty_cache()
    copy = stream_copy_gbps()
    _line('strided-pass access pattern without arithmetic, 256^3 axis 0 '
          '2-point tiles', table['256^3 axis 0, 2-point tiles']['GBps'],
          'GB/s', {'workload': 'jx_tile_probe: padded 6-component spectrum, '
                               'read + write bytes over hipEvent time',
                   'table': table, 'stream_copy_GBps': round(copy, 1)})


WORKLOADS = {'plain2d': workload_plain2d, 'stiffness': workload_stiffness,
             'batched': workload_batched, 'adjoint': workload_adjoint,
             'large': workload_large, 'fft': workload_fft,
             'patterns': workload_patterns}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--grid', type=int, default=None,
                    help='grid size (default: 256 for the metric, the '
                         "workload's own size otherwise)")
    ap.add_argument('--kind', default='filtered',
                    choices=['filtered', 'truncated'])
    ap.add_argument('--profile-iters', type=int, default=20)
    ap.add_argument('--no-autotune', action='store_true',
                    help='keep the default plan variants and schedule')
    ap.add_argument('--no-cpu', action='store_true',
                    help='skip the CPU baseline leg')
    ap.add_argument('--cpu-iters', type=int, default=5,
                    help='LS iterations of the CPU baseline (C restatement '
                         'of janus on the same 256^3 workload)')
    ap.add_argument('--no-configs', action='store_true',
                    help='skip the batched, differentiable and 512^3 '
                         'configurations of the default line')
    ap.add_argument('--schedule', default='auto',
                    choices=['auto', 'fused', 'five'],
                    help='iteration schedule (auto: the library default)')
    ap.add_argument('--copy-mb', type=int, default=0,
                    help='after the bench, copy a buffer of this size '
                         '(reference kernel for HBM traffic counters)')
    ap.add_argument('--workload', default='iteration',
                    choices=['iteration'] + sorted(WORKLOADS),
                    help='iteration: the 256^3 metric (default); others: '
                         'the configurations listed in the module docstring')
    ap.add_argument('--max-iter', type=int, default=5000,
                    help='iteration cap of the converged workloads')
    ap.add_argument('--batch-count', type=int, default=32,
                    help='microstructures of the batched workload')
    ap.add_argument('--iters-per-launch', type=int, default=8)
    ap.add_argument('--batch-streams', type=int, default=2,
                    help='batched workload: streams the groups of different '
                         'grids run on at once (enqueued replays)')
    ap.add_argument('--graph', action='store_true',
                    help='batched workload: replay captured hipGraphs '
                         '(default: enqueue the launches, 2.5%% faster here)')
    ap.add_argument('--gmres-restart', type=int, default=20,
                    help='Arnoldi steps per GMRES cycle (stiffness workload)')
    ap.add_argument('--per-group', action='store_true',
                    help='batched workload: also time each grid size alone')
    ap.add_argument('--no-batched-cg', action='store_true',
                    help='batched workload: skip the conjugate-gradient '
                         'comparison')
    ap.add_argument('--adjoint-reps', type=int, default=2)
    ap.add_argument('--fd-voxels', type=int, default=0,
                    help='large workload: > 0 adds a central difference '
                         'on the mu of phase 1 (two more solves)')
    ap.add_argument('--large-iters', type=int, default=60,
                    help='large workload: iterations of the timed 6-load '
                         'batch (its first iteration also runs the unfused '
                         'pass A)')
    ap.add_argument('--method', default='cg', choices=['basic', 'cg'],
                    help='solver of the large workload\'s forward and '
                         'adjoint solves')
    ap.add_argument('--stats-summary', nargs=2, metavar=('STATS_CSV', 'COMMAND'),
                    help='print the kernel table of a rocprofv3 --stats run '
                         '(no GPU work)')
    ap.add_argument('--summary-title', default='256^3 filtered LS iteration',
                    help='title of --stats-summary')
    ap.add_argument('--traffic-summary', nargs=3,
                    metavar=('FETCH_CSV', 'WRITE_CSV', 'COMMAND'),
                    help='print the counter summary of two rocprofv3 --pmc '
                         'runs (no GPU work)')
    args = ap.parse_args()
    if args.traffic_summary:
        print(traffic_summary(*args.traffic_summary,
                              title=args.summary_title), end='')
        return
    if args.stats_summary:
        print(stats_summary(*args.stats_summary, title=args.summary_title),
              end="")
        return
    if args.gpus != 1:
        print('bench.py: janus_amd runs on one GPU; the multi-GPU slab '
              'decomposition is not implemented (--gpus %d refused)'
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from janus_amd import _lib
    torch.cuda.set_device(0)
    if args.workload != 'iteration':
        WORKLOADS[args.workload](args)
        return
    n = args.grid or 256
    solver, phase, table, mat0 = build_workload(n, args.kind)
    tuned = None
    if not args.no_autotune:
        # the library's autotuner: plan variant per pass, then schedule
        tuned = solver.autotune(iters=5)
        print('bench: autotune %s -> plans %s, fused %s'
              % (json.dumps(tuned), solver.green.plans(), solver.fused),
              file=sys.stderr, flush=True)
    if args.schedule != 'auto':
        solver.fused = args.schedule == 'fused'
    E = np.zeros(6)
    E[5] = 1.0
    eps = torch.empty((n, n, n, 6), dtype=torch.float64, device='cuda')
    eps[...] = torch.from_numpy(E).cuda()
    print('bench: %d^3 %s, warmup %d' % (n, args.kind, args.warmup),
          file=sys.stderr, flush=True)
    if args.warmup > 0:
        solver.solve(E, tol=0.0, max_iter=args.warmup, eps=eps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = solver.solve(E, tol=0.0, max_iter=args.steps, eps=eps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert r.iterations == args.steps
    its = args.steps / elapsed
    print('bench: %d iterations in %.3f s (%.1f it/s), last residual %.3e'
          % (args.steps, elapsed, its, r.residuals[-1]), file=sys.stderr,
          flush=True)

    # per-pass durations with hipEvents on the launch stream
    ms = np.zeros(6)
    _lib.check(_lib.lib().jx_solver_profile(
        solver._handle, eps.data_ptr(), E.ctypes.data, args.profile_iters,
        ms.ctypes.data, _lib.current_stream()))
    fused = solver.fused
    plan_kept = solver.green.plans()
    names = PASS_NAMES_FUSED if fused else PASS_NAMES
    nbytes = pass_bytes((n, n, n), fused)
    dom = int(np.argmax(ms[:5]))
    achieved = nbytes[dom] / (ms[dom] * 1e-3) / 1e9
    iter_bytes = sum(nbytes)
    iter_ms = 1e3 / its
    passes = {names[i]: {'ms': round(float(ms[i]), 4),
                              'bytes': int(nbytes[i]),
                              'GBps': round(nbytes[i] / (ms[i] * 1e-3) / 1e9, 1)
                              if ms[i] > 0 and nbytes[i] else None}
              for i in range(6) if not (fused and i == 0)}
    print('bench: passes ' + json.dumps(passes), file=sys.stderr, flush=True)

    copy_gbps = stream_copy_gbps()
    print('bench: stream copy %.1f GB/s' % copy_gbps, file=sys.stderr,
          flush=True)

    cpu = None
    cpu_all = None
    if not args.no_cpu:
        sample = ('%d LS iterations of the same %d^3 workload with the C '
                  'restatement of janus\'s serial path (oracle/janus_serial.c: '
                  'per-component r2c, janus\'s per-frequency Gamma-hat loop, '
                  'c2r; own batched Stockham FFT, not FFTW)'
                  % (args.cpu_iters, n))
        try:
            _log('CPU baseline (%d iterations, 1 thread)' % args.cpu_iters)
            v, t, res = cpu_baseline(n, args.kind, phase, table, mat0,
                                     args.cpu_iters, 1)
            cpu = {'value': round(v, 6), 'unit': 'iterations/s', 'cores': 1,
                   'kind': 'port', 'sample': sample + ', %.1f s' % t,
                   'residuals': [float(x) for x in res]}
            workers = host_workers()
            _log('CPU baseline (%d iterations, %d threads)'
                 % (args.cpu_iters, workers))
            v, t, _ = cpu_baseline(n, args.kind, phase, table, mat0,
                                   args.cpu_iters, workers)
            cpu_all = {'value': round(v, 6), 'unit': 'iterations/s',
                       'cores': workers, 'kind': 'port',
                       'cpus_available': len(os.sched_getaffinity(0)),
                       'cpu_model': cpu_model(),
                       'sample': sample + ' on %d OpenMP threads, %.1f s'
                                 % (workers, t)}
        except RuntimeError as e:
            _log('CPU baseline skipped: %s' % e)

    configs = None
    adjoint_ratio = None
    if not args.no_configs:
        # free the 256^3 solver's fields before the larger configurations
        del solver, eps, r
        import torch
        torch.cuda.empty_cache()
        configs = {}
        for name, fn in (('differentiable_128', workload_adjoint),
                         ('batched_mix', workload_batched),
                         ('large_512', workload_large)):
            _log('configuration %s' % name)
            a = argparse.Namespace(**vars(args))
            a.grid = None
            t0 = time.perf_counter()
            configs[name] = fn(a, emit=False)
            configs[name]['seconds_total'] = round(time.perf_counter() - t0,
                                                   1)
            torch.cuda.empty_cache()
        adjoint_ratio = configs['differentiable_128']['value']

    traffic, traffic_src = read_traffic(
        n, args.kind, names[dom],
        16 if plan_kept[PASS_PLAN[names[dom]]] == 0 else 8)
    line = {
        'metric': 'LS iterations/s at %d^3 FP64 (%% HBM roofline); '
                  'adjoint-grad/forward time' % n,
        'value': round(its, 3),
        'unit': 'iterations/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(iter_ms, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic',
        'config': {
            'workload': '3D %d^3 basic-scheme LS iteration, %s Gamma0, '
                        'spheres r=%g vf~0.2 seed 2014, mu 1/50 nu 0.3/0.25, '
                        'C0 mu0=25.5 nu0=0.3, E=e_xy, one load case'
                        % (n, args.kind, 16.0 * n / 256.0),
            'grid': [n, n, n],
            'green': args.kind,
            'schedule': 'fused E+A' if fused else 'five passes',
            'plans': plan_kept,
            'autotune_ms': tuned,
            'iteration_bytes_alg': int(iter_bytes),
            'iteration_GBps': round(iter_bytes / (iter_ms * 1e-3) / 1e9, 1),
            'iteration_roofline_frac': round(
                iter_bytes / (iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            'B_alg_5pass': int(five_pass_bytes((n, n, n))),
            'B_alg_5pass_roofline_frac': round(
                five_pass_bytes((n, n, n)) / (iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                4),
            'passes': passes,
            'hbm_copy_GBps': round(copy_gbps, 1),
        },
        'roofline': {
            'bound': 'hbm',
            'kernel': names[dom],
            'achieved': round(achieved, 1),
            'peak': HBM_PEAK_GBS,
            'unit': 'GB/s',
            'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': traffic,
            'traffic_source': traffic_src,
        },
        'adjoint_over_forward': adjoint_ratio,
        'cpu_baseline': cpu,
        'cpu_baseline_all_cores': cpu_all,
        'configs': configs,
    }
    print(json.dumps(line), flush=True)
    if args.copy_mb > 0:
        # the library's stream copy of known size (k_stream_copy), outside
        # the timed region: the reference for calibrating traffic counters
        nbytes = args.copy_mb * (1 << 20)
        a = torch.ones(nbytes // 8, dtype=torch.float64, device='cuda')
        b = torch.empty_like(a)
        for _ in range(3):
            _lib.check(_lib.lib().jx_stream_copy(
                a.data_ptr(), b.data_ptr(), nbytes, _lib.current_stream()))
        torch.cuda.synchronize()
        print('bench: stream-copied %d MiB x 3' % args.copy_mb,
              file=sys.stderr, flush=True)


if __name__ == '__main__':
    main()
