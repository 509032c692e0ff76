"""Write the oracle's optimisation histories on the synthetic non-cuboid domain as golden fixtures.

    python tests/golden/make_value_noise_history.py

value_noise_history.json: the value-noise design domain (top_gpu_on_matlab_amd/synthetic.py: seed 0,
4 octaves, base period 32, 40 % solid, largest 26-connected component, bottom 4 layers clamped, unit
load on the centred 8x8 top patch) on a 64 x 32 x 32 box, read through the .TopVoxel file branch,
V0 = 0.4, ft = 2, rMin = 2, 4 iterations, FP64, for the non-dyadic (spans 4, 2) and the dyadic
(spans 2) hierarchy.  Each history comes with ulpPerturbed, the oracle's own history with the load
vector perturbed at the unit-roundoff level: on this domain the PCG residual stagnates near tol
for tens of iterations, so rounding alone moves the iteration counts by a few and the compliance by
~1e-8, and another implementation is judged against that spread.  mask_sha256 is the SHA-256 of the
packed mask bytes (numpy packbits of the [ny, nx, nz] C-order mask) of the small box and of the
full 256 x 128 x 128 workload, pinning the domain generator.
"""
import hashlib
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import mesh as M  # noqa: E402
from oracle import topopt as T  # noqa: E402
from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise_mask  # noqa: E402
from top_gpu_on_matlab_amd.topvoxel import read_topvoxel, write_topvoxel  # noqa: E402

CASE = dict(shape_xyz=[64, 32, 32], seed=0, octaves=4, period=32, solid_fraction=0.4, V0=0.4, rMin=2.0,
            nLoop=4, ft=2)


def mask_sha256(mask):
    return hashlib.sha256(np.packbits(np.ascontiguousarray(mask, dtype=bool)).tobytes()).hexdigest()


def oracle_model(path, non_dyadic, perturb_seed=None):
    d = read_topvoxel(path)
    m = M.build_file_model(d['nelx'], d['nely'], d['nelz'], d['voxels'], d['fixing'], d['loads'][0],
                           non_dyadic=non_dyadic)
    if perturb_seed is not None:
        m.F = m.F * (1 + 4e-16 * np.random.default_rng(perturb_seed).standard_normal(m.F.shape))
    return m


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    nx, ny, nz = CASE['shape_xyz']
    mask = value_noise_mask(nx, ny, nz, seed=CASE['seed'], octaves=CASE['octaves'], period=CASE['period'],
                            solid_fraction=CASE['solid_fraction'])
    full = value_noise_mask(256, 128, 128, seed=CASE['seed'], octaves=CASE['octaves'], period=CASE['period'],
                            solid_fraction=CASE['solid_fraction'])
    fixing, loads = bottom_clamp_top_patch(mask)
    path = os.path.join(tempfile.mkdtemp(), 'value_noise.TopVoxel')
    write_topvoxel(path, mask, fixing, loads)
    out = {'case': CASE, 'mask_sha256': mask_sha256(mask), 'solid_voxels': int(mask.sum()),
           'full_256x128x128': {'mask_sha256': mask_sha256(full), 'solid_voxels': int(full.sum())},
           'tol': 1e-3, 'runs': {}}
    for nd in (1, 0):
        t = time.time()
        r = T.run_top(oracle_model(path, nd), V0=CASE['V0'], rMin=CASE['rMin'], nLoop=CASE['nLoop'], ft=CASE['ft'])
        p = T.run_top(oracle_model(path, nd, perturb_seed=0), V0=CASE['V0'], rMin=CASE['rMin'], nLoop=CASE['nLoop'],
                      ft=CASE['ft'])
        keys = ('cHist', 'volHist', 'sharpHist', 'lssIts')
        out['runs'][str(nd)] = dict({k: r[k] for k in keys}, cSolid=r['cSolid'], resTail=r['resTail'],
                                    ulpPerturbed=dict({k: p[k] for k in keys}, cSolid=p['cSolid']),
                                    oracle_seconds=time.time() - t)
        print('non_dyadic %d: its %s / perturbed %s, c %s, %.0fs' % (nd, r['lssIts'], p['lssIts'], r['cHist'],
                                                                      time.time() - t), flush=True)
    with open(os.path.join(here, 'value_noise_history.json'), 'w') as f:
        json.dump(out, f, indent=1)


if __name__ == '__main__':
    main()
