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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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


def survey_bytes(n, sym=6, m=1):
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


def cpu_baseline(n, kind, phase, table, mat0, iters=1):
    """The numpy oracle (pocketfft, one thread) on the same workload."""
    from threadpoolctl import threadpool_limits
    from oracle import janus_oracle as jo
    sph = table[phase, 0]
    dev = table[phase, 1]
    E = np.zeros(6)
    E[5] = 1.0
    eps = np.empty((n, n, n, 6))
    eps[...] = E
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for _ in range(iters):
            tau = jo.iso4_apply(sph, dev, eps)
            new = E - jo.green_apply(tau, kind, mat0.g, mat0.nu, 1.0)
            np.linalg.norm(new - eps)
            eps = new
        t = time.perf_counter() - t0
    return iters / t, t


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


def traffic_file(n, kind):
    return os.path.join(ROOT, 'profiles',
                        'ls_iteration_%d_%s_traffic.md' % (n, kind))


def read_traffic(n, kind, kernel):
    """HBM bytes per launch of a kernel from the committed counter summary
    (lines '<tag> <kernel> <bytes>'), or None."""
    path = traffic_file(n, kind)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.strip().strip('`').split()
        if len(f) == 3 and f[0] == TRAFFIC_TAG and f[1] == kernel:
            return int(f[2])
    return None


# kernel symbol prefix -> bench pass name (256^3, sym = 6)
KERNEL_PASS = [('k_rows_update_forward_reg', 'EA_rows_c2r_update_local_r2c'),
               ('k_rows_forward_reg', 'A_rows_r2c_local'),
               ('k_rows_inverse_reg', 'E_rows_c2r_update'),
               ('k_axis_reg<256, 6, 12, -1>', 'B_axis1_fwd'),
               ('k_axis_reg<256, 6, 12, 1>', 'D_axis1_inv'),
               ('k_axis_reg<256, 6, 12, 0>', 'C_axis0_gamma'),
               ('k_stream_copy', 'stream_copy')]


def traffic_summary(fetch_csv, write_csv, cmd):
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
    out = ['# HBM traffic per launch, 256^3 filtered LS iteration', '',
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
        for pre, pname in KERNEL_PASS:
            if pre in name:
                tags.append('    %s %s %d' % (TRAFFIC_TAG, pname, int(fr + wr)))
                break
    out += ['', 'Traffic per launch (bytes) read by bench.py:', ''] + tags
    return '\n'.join(out) + '\n'


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--grid', type=int, default=256)
    ap.add_argument('--kind', default='filtered',
                    choices=['filtered', 'truncated'])
    ap.add_argument('--profile-iters', type=int, default=20)
    ap.add_argument('--no-cpu', action='store_true',
                    help='skip the CPU baseline leg')
    ap.add_argument('--cpu-iters', type=int, default=2)
    ap.add_argument('--schedule', default='auto',
                    choices=['auto', 'fused', 'five'],
                    help='iteration schedule (auto: the library default)')
    ap.add_argument('--copy-mb', type=int, default=0,
                    help='after the bench, copy a buffer of this size '
                         '(reference kernel for HBM traffic counters)')
    ap.add_argument('--traffic-summary', nargs=3,
                    metavar=('FETCH_CSV', 'WRITE_CSV', 'COMMAND'),
                    help='print the counter summary of two rocprofv3 --pmc '
                         'runs (no GPU work)')
    args = ap.parse_args()
    if args.traffic_summary:
        print(traffic_summary(*args.traffic_summary), end='')
        return
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU only; the multi-GPU '
              'slab decomposition is design-only (--gpus %d refused)'
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from janus_amd import _lib
    torch.cuda.set_device(0)
    n = args.grid
    solver, phase, table, mat0 = build_workload(n, args.kind)
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
    if not args.no_cpu:
        print('bench: CPU baseline (%d iteration(s), 1 thread)' % args.cpu_iters,
              file=sys.stderr, flush=True)
        v, t = cpu_baseline(n, args.kind, phase, table, mat0, args.cpu_iters)
        cpu = {'value': round(v, 6), 'unit': 'iterations/s', 'cores': 1,
               'kind': 'port',
               'sample': '%d LS iteration(s) of the same %d^3 workload with '
                         'the numpy oracle (pocketfft), %.1f s'
                         % (args.cpu_iters, n, t)}

    line = {
        'metric': 'LS iterations/s at 256^3 FP64 (% HBM roofline)',
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
            'iteration_bytes_alg': int(iter_bytes),
            'iteration_GBps': round(iter_bytes / (iter_ms * 1e-3) / 1e9, 1),
            'iteration_roofline_frac': round(
                iter_bytes / (iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            'B_alg_5pass': int(survey_bytes((n, n, n))),
            'B_alg_5pass_roofline_frac': round(
                survey_bytes((n, n, n)) / (iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
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
            'traffic': read_traffic(n, args.kind, names[dom]),
        },
        'cpu_baseline': cpu,
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
