"""Device tensors of the wrong dtype are refused before any launch.

The kernels read and write 8-byte values; a float32 tensor of the right
shape handed to them would be overrun.  janus's typed memoryviews refuse
other dtypes with ValueError (utils/checkarray.pyx:3-53 raises ValueError
for every other argument error), and so does the facade.
"""

import numpy as np
import pytest

import janus_amd.fft.serial as fft
import janus_amd.material.elastic.linear.isotropic as material
from janus_amd.batch import BatchSolver
from janus_amd.green import filtered
from janus_amd.operators import BlockDiagonalOperator3D, isotropic_4
from janus_amd.solver import LSSolver

pytestmark = pytest.mark.gpu

SHAPE = (8, 6, 10)


def _ops():
    mat0 = material.create(2.0, 0.3, 3)
    green = filtered(mat0.green_operator(), SHAPE, 1.0,
                     fft.create_real(SHAPE))
    phase = np.zeros(SHAPE, dtype=np.uint8)
    local = BlockDiagonalOperator3D.from_phases(phase, [(1.0, 0.5)])
    return green, local


@pytest.mark.parametrize('dtype', ['float32', 'float16', 'int64'])
def test_wrong_dtype_refused(dtype):
    import torch
    dt = getattr(torch, dtype)
    green, local = _ops()
    good = torch.zeros(SHAPE + (6,), dtype=torch.float64, device='cuda')
    bad = torch.zeros(SHAPE + (6,), dtype=dt, device='cuda')
    with pytest.raises(ValueError, match='float64'):
        green.apply(bad)
    with pytest.raises(ValueError, match='float64'):
        green.apply(good, bad)
    with pytest.raises(ValueError, match='float64'):
        local.apply(bad)
    with pytest.raises(ValueError, match='float64'):
        local.apply(good, bad)
    with pytest.raises(ValueError, match='float64'):
        BlockDiagonalOperator3D(np.full(SHAPE, isotropic_4(1.0, 0.5, 3),
                                        dtype=object)).apply(good, bad)

    t = fft.create_real(SHAPE)
    r = torch.zeros(SHAPE, dtype=torch.float64, device='cuda')
    c = torch.zeros(t.oshape, dtype=torch.float64, device='cuda')
    with pytest.raises(ValueError, match='float64'):
        t.r2c(r.to(dt))
    with pytest.raises(ValueError, match='float64'):
        t.r2c(r, c.to(dt))
    with pytest.raises(ValueError, match='float64'):
        t.c2r(c.to(dt))
    with pytest.raises(ValueError, match='float64'):
        t.c2r(c, r.to(dt))

    S = LSSolver(green, local)
    E = np.array([0., 0., 0., 0., 0., 1.])
    with pytest.raises(ValueError, match='float64'):
        S.solve(E, eps=bad)
    with pytest.raises(ValueError, match='float64'):
        S.solve(E, rhs=bad)
    with pytest.raises(ValueError, match='float64'):
        S.solve_cg(E, rhs=bad)
    with pytest.raises(ValueError, match='float64'):
        S.solve_gmres(E, rhs=bad)
    B = BatchSolver()
    with pytest.raises(ValueError, match='float64'):
        B.add(S, E, rhs=bad)
    # the refused calls launched nothing and left the operators usable
    y = green.apply(good)
    assert float(y.abs().max()) == 0.0


def test_misaligned_device_arrays_refused():
    """A float64 view starting 8 bytes into an allocation is refused by the
    C ABI (the kernels use 16-byte accesses) before any launch."""
    import torch
    t = fft.create_real((16, 12, 10))
    n = 16 * 12 * 10
    base = torch.zeros(n + 1, dtype=torch.float64, device='cuda')
    r = base[1:].view(16, 12, 10)
    assert r.data_ptr() % 16 == 8
    with pytest.raises(ValueError, match='aligned'):
        t.r2c(r)
    green, local = _ops()
    nf = 8 * 6 * 10 * 6
    buf = torch.zeros(nf + 1, dtype=torch.float64, device='cuda')
    x = buf[1:].view(SHAPE + (6,))
    with pytest.raises(ValueError, match='aligned'):
        green.apply(x, torch.zeros(SHAPE + (6,), dtype=torch.float64,
                                   device='cuda'))
    with pytest.raises(ValueError, match='aligned'):
        LSSolver(green, local).solve(np.array([0., 0., 0., 0., 0., 1.]),
                                     eps=x)
    torch.cuda.synchronize()
    assert float(buf.abs().max()) == 0.0
