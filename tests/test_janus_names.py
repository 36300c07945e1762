"""janus's own import lines resolve to janus_amd after
janus_amd.janus_names.install() (square_basic.py:7-12,
square_inclusion_scipy.py:14-17).  Each case runs in a fresh interpreter so
that the registered names do not leak into the other tests."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(code):
    env = dict(os.environ)
    env['PYTHONPATH'] = ROOT + os.pathsep + env.get('PYTHONPATH', '')
    return subprocess.run([sys.executable, '-c', code], cwd=ROOT, env=env,
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=300)


def test_janus_import_lines_resolve():
    p = _run('''
import janus_amd, janus_amd.janus_names as jn
assert not jn.installed()
jn.install()
jn.install()  # idempotent
assert jn.installed()
import janus.green as green
import janus.fft.serial as fft
import janus.material.elastic.linear.isotropic as material
import janus.operators as operators
from janus.operators import isotropic_4
import janus.mandelvoigt
import janus
import janus_amd.green, janus_amd.fft.serial, janus_amd.operators
assert janus is janus_amd
assert green is janus_amd.green and janus.green is green
assert fft is janus_amd.fft.serial
assert operators is janus_amd.operators
assert isotropic_4 is janus_amd.operators.isotropic_4
assert janus.fft.FFTW_ESTIMATE == 64
mat = material.create(1.0, 0.3, 3)
assert abs(mat.g - 1.0) < 1e-15 and abs(mat.nu - 0.3) < 1e-15
t = isotropic_4(1.0, 0.5, 3)
assert t.sph == 1.0 and t.dev == 0.5
for name in ('janus.fft.parallel', 'janus.utils.checkarray'):
    try:
        __import__(name)
    except ModuleNotFoundError:
        pass
    else:
        raise AssertionError(name + ' should not resolve')
print('ok')
''')
    assert p.returncode == 0 and p.stdout.strip().endswith('ok'), p.stdout


def test_foreign_janus_is_not_replaced():
    p = _run('''
import sys, types
sys.modules['janus'] = types.ModuleType('janus')
import janus_amd.janus_names as jn
try:
    jn.install()
except RuntimeError as e:
    assert 'already imported' in str(e)
    print('ok')
''')
    assert p.returncode == 0 and p.stdout.strip().endswith('ok'), p.stdout
