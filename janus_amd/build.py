"""Build ``janus_amd/libjanus_amd.so`` in-tree with hipcc for gfx950."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
SOURCES = ['jx_api.hip', 'jx_kernels.hip']
OUTPUT = os.path.join(HERE, 'libjanus_amd.so')

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    # no packed FP32 instructions anywhere in this library
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
]


def hipcc():
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    return os.path.join(rocm, 'bin', 'hipcc')


def build_library(verbose=False, output=OUTPUT, defines=()):
    """Compile every HIP source into the shared library (always rebuilds).

    ``defines`` are extra ``NAME=VALUE`` macros (tuning variants)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = output + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-D' + d for d in defines] + \
        ['-o', tmp] + srcs
    if verbose:
        print(' '.join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = '\n'.join(line for line in proc.stdout.splitlines()
                    if 'packed-fp32-ops' not in line)
    if proc.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + out)
    if verbose and out.strip():
        print(out)
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
