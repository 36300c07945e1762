"""Write tests/golden/green_fixture_samples.npz from janus's 18 npz fixtures.

janus ships ``tests/data/{truncated,filtered}_green_operator_*.npz`` (about
42 MB): each holds ``x`` (a unit polarization on a 10x10(x10) box) and ``y``
(Gamma0[x] from an independent code, scripts/raw2npz.py:15-37), consumed by
tests/test_discrete_green_operator.py:178-199.  This script keeps, per
fixture:

* the grid shape, the kind, the box origin/size and the polarized component
  read from the data (the 2D filtered files have xx/yy swapped in their
  names, so the file name is not trusted);
* a check that ``x`` is exactly that unit box (so ``x`` can be rebuilt);
* ``y`` at 1024 seeded random voxels plus up to 1024 voxels of a 3-voxel halo
  around the box (all components);
* whole-field checksums of ``y``: per-component sum, sum of |y|, sum of y^2.

Run in a container where the janus sources are mounted:
    python tests/golden/make_green_fixture_samples.py /root/reference/tests/data
"""

import glob
import os
import re
import sys

import numpy as np

NAME_RE = re.compile(r'(truncated|filtered)_green_operator_([0-9x]+)_unit_tau_'
                     r'([xyz]{2})_([0-9x]+)\+([0-9]+)\+([0-9]+)(?:\+([0-9]+))?\.npz')


def main(data_dir, out_path):
    rng = np.random.default_rng(20150312)
    record = {}
    names = []
    for path in sorted(glob.glob(os.path.join(data_dir, '*_green_operator_*.npz'))):
        base = os.path.basename(path)
        m = NAME_RE.match(base)
        assert m, base
        kind = m.group(1)
        with np.load(path, allow_pickle=False) as d:
            x = d['x']
            y = d['y']
        shape = x.shape[:-1]
        dim = len(shape)
        nz = np.argwhere(x != 0)
        comps = np.unique(nz[:, -1])
        assert comps.size == 1, base
        comp = int(comps[0])
        lo = nz[:, :-1].min(axis=0)
        hi = nz[:, :-1].max(axis=0) + 1
        rebuilt = np.zeros_like(x)
        rebuilt[tuple(slice(a, b) for a, b in zip(lo, hi)) + (comp,)] = 1.0
        assert np.array_equal(rebuilt, x), base
        # sampled voxels: random + halo around the box
        rnd = np.stack([rng.integers(0, n, 1024) for n in shape], axis=-1)
        halo_axes = [np.arange(max(a - 3, 0), min(b + 3, n))
                     for a, b, n in zip(lo, hi, shape)]
        halo = np.stack(np.meshgrid(*halo_axes, indexing='ij'), axis=-1)
        halo = halo.reshape(-1, dim)
        if halo.shape[0] > 1024:
            halo = halo[rng.choice(halo.shape[0], 1024, replace=False)]
        idx = np.concatenate([rnd, halo]).astype(np.int32)
        vals = y[tuple(idx.T)]
        key = '{0}_{1}_c{2}'.format(kind, 'x'.join(map(str, shape)), comp)
        names.append(key)
        record[key + '/shape'] = np.array(shape, dtype=np.int64)
        record[key + '/kind'] = np.array(0 if kind == 'truncated' else 1)
        record[key + '/comp'] = np.array(comp)
        record[key + '/lo'] = lo.astype(np.int64)
        record[key + '/hi'] = hi.astype(np.int64)
        record[key + '/idx'] = idx
        record[key + '/y'] = vals
        record[key + '/sum'] = y.reshape(-1, y.shape[-1]).sum(axis=0)
        record[key + '/sum_abs'] = np.array(np.abs(y).sum())
        record[key + '/sum_sq'] = np.array(np.square(y).sum())
        record[key + '/max_abs'] = np.array(np.abs(y).max())
        record[key + '/source'] = np.array(base)
    record['names'] = np.array(names)
    np.savez_compressed(out_path, **record)
    print('wrote', out_path, 'with', len(names), 'fixtures')


if __name__ == '__main__':
    data_dir = sys.argv[1] if len(sys.argv) > 1 else '/root/reference/tests/data'
    here = os.path.dirname(os.path.abspath(__file__))
    main(data_dir, os.path.join(here, 'green_fixture_samples.npz'))
