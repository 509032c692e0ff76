"""Write the oracle's optimisation histories for the small reference cases as golden fixtures.

    python tests/golden/make_c1_history.py

c1_history.json: TOP_GPU(true(24,48,24), 'optCase',1, 'V0',0.12) -- 50 iterations, ft = 1,
rMin = sqrt(3), Emin = 1e-6 (the code's value, TOP_GPU.m:390), FP64 -- the compliance,
volume, sharpness, constraint and change histories, the PCG iteration counts and the final
xPhys checksum, computed by the oracle (oracle/topopt.py run_top).  resTail holds the last two
relative PCG residuals of every solve: an implementation whose rounding differs may stop one
iteration earlier or later only where one of them is within rounding distance of tol.  ocMargin
holds, per iteration, the smallest |fval + mean(xnew - x)| the OC bisection branched on: where it
is within rounding distance of zero another implementation may branch the other way, which moves
the Lagrange multiplier by up to the bisection tolerance (1e-6 relative).  ulpPerturbed is the
oracle's own history with the load vector perturbed at the unit-roundoff level: the spread that
rounding alone produces, against which another implementation's deviation is judged.
"""
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import mesh as M  # noqa: E402
from oracle import topopt as T  # noqa: E402


def history(shape, opt_case, V0, rMin, nLoop, ft, super_element=False):
    m = M.build_cuboid_model(*shape, opt_case=opt_case, super_element=super_element)
    t = time.time()
    out = T.run_top(m, V0=V0, rMin=rMin, nLoop=nLoop, ft=ft)
    xp = np.asarray(out['xPhys']).reshape(-1, order='F')
    return {
        'case': dict(shape=list(shape), optCase=opt_case, V0=V0, rMin=rMin, nLoop=nLoop, ft=ft,
                     super_element=super_element, Emin=1e-6, precision='fp64'),
        'cSolid': out['cSolid'], 'cHist': out['cHist'], 'volHist': out['volHist'], 'sharpHist': out['sharpHist'],
        'consHist': out['consHist'], 'changeHist': out['changeHist'], 'lssIts': out['lssIts'],
        'resTail': out['resTail'], 'tol': 1e-3, 'ocMargin': out['ocMargin'],
        'xPhys_sum': float(xp.sum()), 'xPhys_sha256_f32': hashlib.sha256(xp.astype(np.float32).tobytes()).hexdigest(),
        'oracle_seconds': time.time() - t,
    }


def ulp_spread(shape, opt_case, V0, rMin, nLoop, ft, seed=0):
    """The same run with the load vector perturbed at the unit-roundoff level (relative 4e-16): how
    far rounding alone moves the oracle's own history (CG and the OC bisection amplify it)."""
    m = M.build_cuboid_model(*shape, opt_case=opt_case)
    m.F = m.F * (1 + 4e-16 * np.random.default_rng(seed).standard_normal(m.F.shape))
    out = T.run_top(m, V0=V0, rMin=rMin, nLoop=nLoop, ft=ft)
    return {'seed': seed, 'cHist': out['cHist'], 'volHist': out['volHist'], 'sharpHist': out['sharpHist'],
            'lssIts': out['lssIts']}


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    h = history((24, 48, 24), 1, 0.12, math.sqrt(3), 50, 1)
    h['ulpPerturbed'] = ulp_spread((24, 48, 24), 1, 0.12, math.sqrt(3), 50, 1)
    with open(os.path.join(here, 'c1_history.json'), 'w') as f:
        json.dump(h, f, indent=1)
    print('c1: %d iterations, c = %.10e, %.0fs' % (len(h['cHist']), h['cHist'][-1], h['oracle_seconds']))


if __name__ == '__main__':
    main()
