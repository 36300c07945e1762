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


def build_library(verbose=False):
    """Compile every HIP source into the shared library (always rebuilds)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUTPUT + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-o', tmp] + srcs
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
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == '__main__':
    print(build_library(verbose='-v' in sys.argv))
