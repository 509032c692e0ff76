"""Write the oracle's 50-iteration history of the super-element MBB beam as a golden fixture.

    python tests/golden/make_c2_history.py

c2_history.json: TOP_GPU(true(10,60,10), 'optCase',2, 'V0',0.2, 'ft',2) in super-element mode (4 x 4 x 4
sub-voxels, the convolution density filter with rMin = 4 sqrt(3) sub-voxels), Emin = 1e-6, FP64, 50
iterations, computed by the oracle (oracle/topopt.py run_top): compliance, volume, sharpness,
constraint and change histories, PCG iteration counts, the final xPhys checksum, and ulpPerturbed: the
oracle's own history with the load vector perturbed at the unit-roundoff level, and perturbed: three more
runs with the loads perturbed by 3e-13 relative (random signs), the size of the difference another
implementation's rounding makes in the first solve's compliance (the device's is 6e-13); the history is
chaotic, so later deviations scale with that first difference, and these runs give the spread that such
a start produces.
"""
import json
import math
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

sys.path.insert(0, os.path.join(ROOT, 'tests', 'golden'))
from make_c1_history import history  # noqa: E402

import numpy as np  # noqa: E402

from oracle import mesh as M  # noqa: E402
from oracle import topopt as T  # noqa: E402


def ulp_spread_super(shape, V0, rMin, nLoop, ft, seed=0, delta=4e-16, signs=False):
    m = M.build_cuboid_model(*shape, opt_case=2, super_element=True)
    rng = np.random.default_rng(seed)
    m.F = m.F * (1 + delta * (rng.choice([-1.0, 1.0], size=m.F.shape) if signs else rng.standard_normal(m.F.shape)))
    out = T.run_top(m, V0=V0, rMin=rMin, nLoop=nLoop, ft=ft)
    return {'seed': seed, 'cHist': out['cHist'], 'volHist': out['volHist'], 'sharpHist': out['sharpHist'],
            'lssIts': out['lssIts']}


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    shape = (10, 60, 10)
    h = history(shape, 2, 0.2, math.sqrt(3), 50, 2, super_element=True)
    h['ulpPerturbed'] = ulp_spread_super(shape, 0.2, math.sqrt(3), 50, 2)
    h['perturbed'] = {'delta': 3e-13, 'runs': [ulp_spread_super(shape, 0.2, math.sqrt(3), 50, 2, seed=s2, delta=3e-13,
                                                                 signs=True) for s2 in (1, 2, 3)]}
    with open(os.path.join(here, 'c2_history.json'), 'w') as f:
        json.dump(h, f, indent=1)
    print('c2: %d iterations, c = %.10e, %.0fs' % (len(h['cHist']), h['cHist'][-1], h['oracle_seconds']))


if __name__ == '__main__':
    main()
