"""The reference workloads beyond C1 and C5 on the device: the non-cuboid value-noise domain (file
input, both hierarchy rules), the super-element MBB beam at its full size, and the mixed-precision
cantilever against the FP64 solve of the same systems."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from oracle import mesh as M
from oracle import topopt as T

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize('non_dyadic', [1, 0])
def test_value_noise_domain_history_against_golden(tmp_path, non_dyadic):
    """Non-cuboid domain through the .TopVoxel file branch: the 256 x 128 x 128 value-noise workload
    (seed 0, 4 octaves, base period 32, 40% solid, largest 26-connected component, bottom 4 layers
    clamped, unit load on the centred 8x8 top patch) with the same noise field on its 64 x 32 x 32
    corner box (thresholded at that box's own 60th percentile), so the features keep their full-size
    width (~16 voxels); V0 = 0.4, ft = 2, rMin = 2, 4 iterations, non-dyadic (span 4, 2) and dyadic
    (spans 2) hierarchies, against the oracle's history in tests/golden/value_noise_history.json.

    On this domain the PCG residual stagnates near tol for tens of iterations: the oracle with its
    load vector perturbed at the unit-roundoff level (the golden's ulpPerturbed run) moves its own
    iteration counts by up to 7 and its compliance history by up to ~4e-5.  The device must stay
    within ten times that spread (compliance, volume, sharpness; for the solid design's compliance,
    ten times the larger spread of the two first solves) and, in every iteration, within three
    times the largest count spread of the history (at least 2) of the oracle's PCG counts.  The operators on this domain are pinned tightly
    by the solver tests' value-noise case."""
    from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise_mask
    from top_gpu_on_matlab_amd.top_gpu import TOP_GPU
    from top_gpu_on_matlab_amd.topvoxel import load_topvoxel_model, write_topvoxel
    g = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'value_noise_history.json')))
    c = g['case']
    nx, ny, nz = c['shape_xyz']
    mask = value_noise_mask(nx, ny, nz, seed=c['seed'], octaves=c['octaves'], period=c['period'],
                            solid_fraction=c['solid_fraction'])
    assert hashlib.sha256(np.packbits(mask).tobytes()).hexdigest() == g['mask_sha256']
    fixing, loads = bottom_clamp_top_patch(mask)
    path = os.path.join(tmp_path, 'value_noise.TopVoxel')
    write_topvoxel(path, mask, fixing, loads)
    gm = load_topvoxel_model(path, non_dyadic=non_dyadic)
    got = TOP_GPU(path, V0=c['V0'], rMin=c['rMin'], nLoop=c['nLoop'], ft=c['ft'], model=gm)
    ref = g['runs'][str(non_dyadic)]
    up = ref['ulpPerturbed']
    print('\n[value-noise non_dyadic=%d] levels %d, its %s (oracle %s, perturbed %s)'
          % (non_dyadic, gm.num_levels, got['lssIts'], ref['lssIts'], up['lssIts']), flush=True)
    assert len(got['cHist']) == len(ref['cHist'])
    # cSolid and cHist[0] are both first solves from zero in the same stagnating regime, each with one
    # perturbed sample of its spread (2.2e-8 and 2.5e-7 for the non-dyadic hierarchy): the larger one
    # estimates the spread of either
    sp = max(abs(up['cSolid'] - ref['cSolid']) / ref['cSolid'], abs(up['cHist'][0] - ref['cHist'][0]) / ref['cHist'][0])
    print('  cSolid: oracle first-solve spread %.2e, device %.2e'
          % (sp, abs(got['cSolid'] - ref['cSolid']) / ref['cSolid']), flush=True)
    assert abs(got['cSolid'] - ref['cSolid']) < max(10 * sp, 1e-10) * ref['cSolid']
    for k in ('cHist', 'volHist', 'sharpHist'):
        r = np.array(ref[k])
        spread = (np.abs(np.array(up[k]) - r) / np.abs(r)).max()
        dev = (np.abs(np.array(got[k]) - r) / np.abs(r)).max()
        print('  %s: oracle ulp-spread %.2e, device %.2e' % (k, spread, dev), flush=True)
        assert dev < max(10 * spread, 1e-10), (k, got[k], ref[k])
    its_spread = max(abs(p - b) for b, p in zip(ref['lssIts'], up['lssIts']))
    for a, b in zip(got['lssIts'], ref['lssIts']):
        assert abs(a - b) <= max(2, 3 * its_spread), (got['lssIts'], ref['lssIts'], up['lssIts'])


def test_super_element_mbb_full_size_matches_oracle():
    """MBB beam, 10 x 60 x 10 voxels, super-element mode (4x4x4 sub-voxels), ft = 2 (the convolution
    filter in this mode), V0 = 0.2, rMin = sqrt(3), 3 iterations against the oracle run live."""
    from top_gpu_on_matlab_amd.top_gpu import TOP_GPU
    shape = (10, 60, 10)
    om = M.build_cuboid_model(*shape, opt_case=2, super_element=True)
    ref = T.run_top(om, V0=0.2, rMin=math.sqrt(3), nLoop=3, ft=2)
    got = TOP_GPU(np.ones(shape, bool), optCase=2, V0=0.2, rMin=math.sqrt(3), nLoop=3, ft=2, super_element=1)
    assert got['lssIts'] == ref['lssIts']
    assert abs(got['cSolid'] - ref['cSolid']) < 1e-10 * ref['cSolid']
    for k in ('cHist', 'volHist', 'sharpHist'):
        assert _rel(got[k], ref[k]) < 1e-10, (k, got[k], ref[k])


def test_super_element_mbb_fifty_iterations_against_golden():
    """The full super-element MBB configuration: 10 x 60 x 10, ft = 2, V0 = 0.2, 50 iterations, against
    the oracle's committed history (tests/golden/c2_history.json).

    The history is chaotic with respect to rounding: every warm-started PCG solve, filter and OC
    bisection passes last-bit differences on, growing for ~12 iterations, after which both runs settle on
    the same design and the differences die out.  The oracle itself, with its loads perturbed by 3e-13
    (the golden's `perturbed` runs), moves by up to 2e-9; the device, whose rounding differs in every
    operation of every iteration, by up to ~1e-7 near iteration 13, decaying to a few 1e-10 by the end
    (measured 3.5e-10 over the last 20).  So the history is checked where rounding cannot have been
    amplified and where it has died out: iteration 1 (compliance, volume, sharpness, PCG count) within
    1e-10, the last 20 iterations within 2e-9, and every iteration within 1e-6 (a wrong operator, filter
    or update moves it by orders of magnitude more)."""
    from top_gpu_on_matlab_amd.top_gpu import TOP_GPU
    g = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'c2_history.json')))
    c = g['case']
    got = TOP_GPU(np.ones(c['shape'], bool), optCase=c['optCase'], V0=c['V0'], rMin=c['rMin'], nLoop=c['nLoop'],
                  ft=c['ft'], super_element=1)
    n = len(g['cHist'])
    assert len(got['cHist']) == n
    assert abs(got['cSolid'] - g['cSolid']) < 1e-10 * g['cSolid']
    assert got['lssIts'][0] == g['lssIts'][0]
    for key in ('cHist', 'volHist', 'sharpHist'):
        ref = np.array(g[key])
        spread = np.max([np.abs(np.array(r[key]) - ref) / np.abs(ref) for r in g['perturbed']['runs']], axis=0)
        dev = np.abs(np.array(got[key]) - ref) / np.abs(ref)
        print('%s: oracle spread (3e-13 start) %s; device %s' % (key, ['%.1e' % v for v in spread],
                                                                  ['%.1e' % v for v in dev]), flush=True)
        assert dev[0] < 1e-10, key
        assert dev[-20:].max() < 2e-9, (key, dev[-20:])
        assert dev.max() < 1e-6, (key, dev)


def test_mixed_precision_cantilever_96x192x96_against_fp64():
    """The mixed-precision configuration at full length: 96 x 192 x 96 cantilever, ft = 2, rMin = 3,
    V0 = 0.12, 50 optimisation iterations driven by the FP64 solver; every iteration's system is solved
    by the mixed solver too (FP32 V-cycle, FP64 PCG), warm-started from its own previous solution, and by
    the mixed solver with the finest element products as BF16 operands on the matrix cores, once with
    each operand as one BF16 value and once split into two (hi + lo).  Per iteration, the FP32 V-cycle and
    the split BF16 products: relative compliance difference within 1e-6 (measured ~1e-11 and up to
    4.8e-7) and the same PCG count, or one more or fewer; the single-BF16 solver's differences (up to
    ~1e-5, up to 13 more PCG iterations) are reported."""
    import torch
    from top_gpu_on_matlab_amd import _lib
    from top_gpu_on_matlab_amd import top_gpu as G
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver
    shape = (96, 192, 96)
    sv = {p: Solver(cuboid_model(*shape, 1, precision=p)) for p in ('fp64', 'mixed')}
    sv['bf16'] = Solver(cuboid_model(*shape, 1, precision='mixed'))
    sv['bf16'].set_tuning(kfine_mfma='bf16')
    sv['bf16x3'] = Solver(cuboid_model(*shape, 1, precision='mixed'))
    sv['bf16x3'].set_tuning(kfine_mfma='bf16x3')
    model = sv['fp64'].model
    Ne = model.num_elements
    F = torch.from_numpy(model.get_table(1, 'F')[:, 0].copy()).cuda()
    filt = G.PDEFilter(model, 3.0, tol=1e-4)
    x = torch.full((Ne,), 0.12, dtype=torch.float64, device='cuda')
    xPhys = x.clone()
    U = {p: None for p in sv}
    ones = torch.ones_like(x)
    errs, its_hist, errs16, errs163 = [], [], [], []
    for it in range(50):
        E = G.simp(xPhys, 1e-6, 1.0, 3.0)
        c, its = {}, {}
        for p, s in sv.items():
            s.set_modulus(E)
            s.assemble()
            U[p], its[p], _ = s.pcg(F, y0=U[p])
            c[p] = float(torch.dot(F, U[p]))
        errs.append(abs(c['mixed'] - c['fp64']) / c['fp64'])
        errs16.append(abs(c['bf16'] - c['fp64']) / c['fp64'])
        errs163.append(abs(c['bf16x3'] - c['fp64']) / c['fp64'])
        its_hist.append((its['fp64'], its['mixed'], its['bf16'], its['bf16x3']))
        ce = G.unit_compliance(sv['fp64'], U['fp64'])
        dc = filt(G.sensitivity(xPhys, ce, 1.0, 1e-6, 1.0, 3.0))
        dv = filt(ones)
        fval = G.reduce(_lib.TG_RED_SUM, xPhys) / Ne - 0.12
        x, _ = G.oc_update(x, dc, dv, fval, 0.2)
        xPhys = filt(x)
    print('\n[96x192x96 mixed vs fp64, 50 iterations] rel compliance difference per iteration: FP32 V-cycle %s; '
          'BF16 element products %s; split BF16 %s; PCG its (fp64, mixed, bf16, bf16x3) %s'
          % (['%.1e' % e for e in errs], ['%.1e' % e for e in errs16], ['%.1e' % e for e in errs163], its_hist),
          flush=True)
    assert max(errs) < 1e-6
    assert all(abs(a - b) <= 1 for a, b, _, _ in its_hist), its_hist
    # split BF16 (measured: within 4.8e-7, the FP64 counts or one more)
    assert max(errs163) < 1e-6
    assert all(abs(a - d) <= 1 for a, _, _, d in its_hist), its_hist


@pytest.mark.parametrize('non_dyadic', [1, 0])
def test_value_noise_pcg_with_smaller_jacobi_weight_matches_oracle(tmp_path, non_dyadic):
    """The V-cycle smoother weight as a parameter (the reference's global weightFactorJacobi_): on the
    64 x 32 x 32 value-noise box, omega = 0.4 instead of 0.6 (0.6 exceeds 2 / lambda_max(D^-1 K)
    there) cuts the first solve from 105 to 24 PCG iterations (54 to 18 on the dyadic hierarchy).

    The device follows the oracle with the same weight: the same PCG count, or one fewer or more where
    the oracle's relative residual at the earlier of the two stops lies within 5% of tol.  The
    non-dyadic solve is such a near-tie: the oracle's residual after 23 iterations is 1.005e-3 on one
    CPU and 1.020e-3 on another (its own rounding moves it by 1.5%), and rounding-level differences in
    the coarsest factorisation decide it either way.  Run for the same number of iterations, the
    compliance must agree within 20x the oracle's own spread under a unit-roundoff perturbation of the
    load."""
    import torch
    from oracle import solver as S
    from top_gpu_on_matlab_amd.solver import Solver
    from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise_mask
    from top_gpu_on_matlab_amd.topvoxel import load_topvoxel_model, read_topvoxel, write_topvoxel
    mask = value_noise_mask(64, 32, 32, seed=0, octaves=4, period=32)
    path = os.path.join(tmp_path, 'value_noise.TopVoxel')
    write_topvoxel(path, mask, *bottom_clamp_top_patch(mask))
    d = read_topvoxel(path)
    om = M.build_file_model(d['nelx'], d['nely'], d['nelz'], d['voxels'], d['fixing'], d['loads'][0],
                            non_dyadic=non_dyadic)
    osv = S.Solver(om, weight_factor_jacobi=0.4)
    osv.assemble(np.ones(om.levels[0].numElements))
    b = om.F[:, 0]
    _, its_ref, hist = osv.pcg(b, 1e-3, 800)
    gm = load_topvoxel_model(path, non_dyadic=non_dyadic)
    gsv = Solver(gm)
    gsv.set_modulus(torch.ones(gm.num_elements, dtype=torch.float64, device='cuda'))
    gsv.assemble()
    its06 = gsv.pcg(b)[1]
    gsv.set_jacobi_weight(0.4)
    y, its, _ = gsv.pcg(b)
    near_tie = abs(hist[min(its, its_ref) - 1] / 1e-3 - 1) < 0.05
    # the oracle run for exactly `its` iterations (tol 0), and its unit-roundoff spread there
    bp = b * (1 + 2.0 ** -52 * np.random.default_rng(11).standard_normal(b.shape))
    y_same = osv.pcg(b, 0.0, its)[0]
    y_per = osv.pcg(bp, 0.0, its)[0]
    c_ref = float(b @ y_same)
    spread = abs(float(b @ y_per) - c_ref) / c_ref
    dev = abs(float(b @ y.cpu().numpy()) - c_ref) / c_ref
    print('\n[value-noise non_dyadic=%d] PCG its: omega 0.6 -> %d, omega 0.4 -> %d (oracle %d, residual tail %s); '
          'compliance after %d iterations %.1e (oracle spread %.1e)'
          % (non_dyadic, its06, its, its_ref, ['%.4e' % v for v in hist[-2:]], its, dev, spread), flush=True)
    assert its == its_ref or (abs(its - its_ref) == 1 and near_tie)
    assert dev < max(20 * spread, 1e-12)
    assert 2 * its < its06


def test_value_noise_full_size_256x128x128_converges():
    """The full 256 x 128 x 128 value-noise configuration (1,676,295 voxels, 5.67M DOFs, .TopVoxel file
    branch, V0 = 0.4, ft = 2, rMin = 2) for its specified 30 iterations.  With the reference's weight
    0.6 every PCG solve runs to maxIT = 800 at a relative residual near 0.1 (0.6 exceeds
    2 / lambda_max(D^-1 K) on this domain); with weightFactorJacobi = 0.4 every solve converges (< 300 PCG
    iterations), the volume stays at V0 and the compliance decreases overall."""
    import tempfile
    from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise_mask
    from top_gpu_on_matlab_amd.top_gpu import TOP_GPU
    from top_gpu_on_matlab_amd.topvoxel import load_topvoxel_model, write_topvoxel
    g = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'value_noise_history.json')))
    mask = value_noise_mask(256, 128, 128, seed=0, octaves=4, period=32)
    assert hashlib.sha256(np.packbits(mask).tobytes()).hexdigest() == g['full_256x128x128']['mask_sha256']
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, 'value_noise.TopVoxel')
        write_topvoxel(path, mask, *bottom_clamp_top_patch(mask))
        gm = load_topvoxel_model(path)
        got = TOP_GPU(path, V0=0.4, rMin=2.0, nLoop=30, ft=2, model=gm, weightFactorJacobi=0.4)
    print('\n[value-noise 256x128x128, omega 0.4, 30 iterations] its %s, c %s'
          % (got['lssIts'], ['%.6e' % v for v in got['cHist']]), flush=True)
    assert len(got['cHist']) == 30
    assert max(got['lssIts']) < 300
    assert all(abs(v - 0.4) < 1e-4 for v in got['volHist'])
    assert all(np.isfinite(got['cHist'])) and got['cHist'][2] < got['cHist'][1] < got['cHist'][0]
    assert got['cHist'][-1] < got['cHist'][0]
