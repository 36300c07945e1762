"""Build ``janus_amd/libjanus_amd.so`` in-tree with hipcc for gfx950."""

import concurrent.futures
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
SOURCES = ['jx_api.hip', 'jx_batch.hip', 'jx_rows_fwd.hip', 'jx_rows_inv.hip',
           'jx_rows_fused.hip', 'jx_axis.hip', 'jx_misc.hip', 'jx_cg.hip',
           'jx_gmres.hip']
OUTPUT = os.path.join(HERE, 'libjanus_amd.so')

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
    # no packed FP32 instructions anywhere in this library
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
]


def hipcc():
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    return os.path.join(rocm, 'bin', 'hipcc')


def _run(cmd, verbose):
    if verbose:
        print(' '.join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = '\n'.join(line for line in proc.stdout.splitlines()
                    if 'packed-fp32-ops' not in line)
    return proc.returncode, out


def build_library(verbose=False, output=OUTPUT, defines=(), jobs=None):
    """Compile every HIP source into the shared library (always rebuilds).

    The sources compile in parallel (``jobs`` processes, default: the
    usable CPUs, at most 8) into a temporary folder and are then linked.
    ``defines`` are extra ``NAME=VALUE`` macros (tuning variants)."""
    flags = HIPCC_FLAGS + ['-D' + d for d in defines]
    if jobs is None:
        jobs = max(1, min(8, len(os.sched_getaffinity(0))))
    tmp = output + '.tmp'
    with tempfile.TemporaryDirectory(prefix='jx_build_') as work:
        objs, cmds = [], []
        for src in SOURCES:
            obj = os.path.join(work, os.path.splitext(src)[0] + '.o')
            objs.append(obj)
            cmds.append([hipcc()] + flags + ['-c', '-o', obj,
                                             os.path.join(CSRC, src)])
        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
            results = list(ex.map(lambda c: _run(c, verbose), cmds))
        for (rc, out), src in zip(results, SOURCES):
            if rc != 0:
                raise RuntimeError('hipcc failed on {0}:\n{1}'.format(src, out))
            if verbose and out.strip():
                print(out)
        rc, out = _run([hipcc()] + flags + ['-shared', '-o', tmp] + objs,
                       verbose)
        if rc != 0:
            raise RuntimeError('hipcc link failed:\n' + out)
    os.replace(tmp, output)
    return output


if __name__ == '__main__':
    args = [a for a in sys.argv[1:] if a != '-v']
    out = OUTPUT
    defs = []
    for a in args:
        if a.startswith('-o='):
            out = os.path.abspath(a[3:])
        elif a.startswith('-D'):
            defs.append(a[2:])
    print(build_library(verbose='-v' in sys.argv, output=out, defines=defs))
