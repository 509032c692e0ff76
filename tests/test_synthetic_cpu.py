"""CPU tests of the synthetic non-cuboid domain (value noise, largest 26-connected component,
clamped bottom, loaded top patch) and its .TopVoxel round trip against the oracle's file branch."""
import os

import numpy as np
import pytest
import scipy.ndimage

from oracle import mesh as M

pytest.importorskip('top_gpu_on_matlab_amd._lib')


def test_value_noise_domain_properties():
    from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise, value_noise_mask
    v1 = value_noise(40, 20, 24, seed=0, octaves=4, period=8)
    assert np.array_equal(v1, value_noise(40, 20, 24, seed=0, octaves=4, period=8))   # seeded, deterministic
    assert not np.array_equal(v1, value_noise(40, 20, 24, seed=1, octaves=4, period=8))
    assert v1.shape == (20, 40, 24) and 0 <= v1.min() and v1.max() < 2.0             # sum of 2^-o [0,1) octaves
    mask = value_noise_mask(40, 20, 24, seed=0, octaves=4, period=8, solid_fraction=0.4)
    frac = mask.mean()
    assert 0.25 < frac <= 0.4 + 1e-9                                                 # component of the top 40 %
    _, n = scipy.ndimage.label(mask, structure=np.ones((3, 3, 3), bool))
    assert n == 1
    m6 = value_noise_mask(40, 20, 24, seed=0, octaves=4, period=8, solid_fraction=0.4, connectivity=6)
    assert m6.sum() <= mask.sum() and scipy.ndimage.label(m6)[1] == 1
    fixing, loads = bottom_clamp_top_patch(mask)
    assert len(fixing) > 0 and len(loads) > 0
    assert abs(loads[:, 3].sum() + 1.0) < 1e-12 and np.all(loads[:, 1:3] == 0)
    assert not set(fixing[:, 0]) & set(loads[:, 0])
    # every clamped node lies in the bottom 4 voxel layers and on the mesh boundary
    NY, NX = 21, 41
    z = (fixing[:, 0].astype(np.int64) - 1) // (NY * NX)
    assert z.max() <= 4


@pytest.mark.parametrize('non_dyadic', [1, 0])
def test_value_noise_domain_model_matches_oracle(tmp_path, non_dyadic):
    from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise_mask
    from top_gpu_on_matlab_amd.topvoxel import load_topvoxel_model, read_topvoxel, write_topvoxel
    mask = value_noise_mask(48, 24, 24, seed=0, octaves=4, period=8, connectivity=6)
    fixing, loads = bottom_clamp_top_patch(mask)
    path = os.path.join(tmp_path, 'c3.TopVoxel')
    write_topvoxel(path, mask, fixing, loads)
    d = read_topvoxel(path)
    gm = load_topvoxel_model(path, non_dyadic=non_dyadic)
    om = M.build_file_model(d['nelx'], d['nely'], d['nelz'], d['voxels'], d['fixing'], d['loads'][0],
                            non_dyadic=non_dyadic)
    assert gm.num_levels == om.numLevels
    for lvl in range(1, om.numLevels + 1):
        assert np.array_equal(gm.get_map(lvl, 'eNodMat'), om.levels[lvl - 1].eNodMat)
    assert np.array_equal(gm.get_table(1, 'F'), om.F)
    assert np.array_equal(gm.get_table(1, 'fixedDOFs').astype(bool), om.fixedDOFs)


def test_value_noise_masks_match_committed_hashes():
    """The generator is pinned by the SHA-256 of the packed masks: the 64 x 32 x 32 parity box and the
    full 256 x 128 x 128 workload (~1.68M solid voxels)."""
    import hashlib
    import json
    from top_gpu_on_matlab_amd.synthetic import value_noise_mask
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    g = json.load(open(os.path.join(root, 'tests', 'golden', 'value_noise_history.json')))
    c = g['case']
    for shape, want in ((c['shape_xyz'], g), ((256, 128, 128), g['full_256x128x128'])):
        m = value_noise_mask(*shape, seed=c['seed'], octaves=c['octaves'], period=c['period'],
                             solid_fraction=c['solid_fraction'])
        assert int(m.sum()) == want['solid_voxels']
        assert hashlib.sha256(np.packbits(m).tobytes()).hexdigest() == want['mask_sha256']


def test_value_noise_domain_overdamps_jacobi_at_the_reference_weight(tmp_path):
    """Why PCG stalls on ragged domains with the reference's smoother weight (weightFactorJacobi_ =
    0.6, TOP_GPU.m:397): a Rayleigh quotient of D^-1 K (a lower bound on its largest eigenvalue) on
    the 48 x 24 x 24 value-noise domain (period 32, as the 256 x 128 x 128 workload) exceeds 2 / 0.6,
    so I - 0.6 D^-1 K has an eigenvalue below -1: the smoother amplifies that mode and the V-cycle is
    no longer a positive definite preconditioner.  With omega = 0.4 the oracle's PCG converges in a
    fraction of the iterations."""
    from oracle import solver as S
    from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise_mask
    from top_gpu_on_matlab_amd.topvoxel import read_topvoxel, write_topvoxel
    mask = value_noise_mask(48, 24, 24, seed=0, octaves=4, period=32)
    path = os.path.join(tmp_path, 'value_noise.TopVoxel')
    write_topvoxel(path, mask, *bottom_clamp_top_patch(mask))
    d = read_topvoxel(path)
    om = M.build_file_model(d['nelx'], d['nely'], d['nelz'], d['voxels'], d['fixing'], d['loads'][0])
    its = {}
    for w in (0.6, 0.4):
        sv = S.Solver(om, weight_factor_jacobi=w)
        sv.assemble(np.ones(om.levels[0].numElements))
        its[w] = sv.pcg(om.F[:, 0], 1e-3, 800)[1]
    dg = om.levels[0].diagK
    v = np.random.default_rng(0).standard_normal(dg.size)
    for _ in range(40):
        w = sv.kbyu(v, 1) / dg
        v = w / np.linalg.norm(w)
    rq = float(v @ sv.kbyu(v, 1)) / float(v @ (dg * v))
    assert 0.6 * rq > 2.0 > 0.4 * rq, rq
    assert its[0.4] * 2 < its[0.6], its


def test_unfiltered_random_density_overdamps_jacobi_at_the_reference_weight():
    """Why the first optimisation iteration of the 128 x 256 x 128 benchmark runs both solvers to maxIT:
    its densities are the unfiltered start x0 = 0.12 + U(-0.05, 0.05) (seeded), so neighbouring moduli
    differ by up to 14x, and the largest eigenvalue of D^-1 K on the finest level rises from 3.09 (any
    uniform field) to 3.9, past 2 / 0.6: the damped Jacobi smoother with the reference's weight amplifies
    those modes.  The oracle's PCG on the same recipe at 16 x 32 x 16 needs several times the iterations
    of omega = 0.5, which stays below 2 / lambda_max.  After the first density filter the field is smooth
    and the weight 0.6 converges again."""
    from oracle import solver as S
    m = M.build_cuboid_model(16, 32, 16, opt_case=1)
    Ne = m.levels[0].numElements
    x0 = 0.12 + np.random.default_rng(12345).uniform(-0.05, 0.05, Ne)
    E = 1e-6 + x0 ** 3 * (1 - 1e-6)

    def lam_max(sv):
        dg = sv.levels[0].diagK
        v = np.random.default_rng(0).standard_normal(dg.size)
        for _ in range(60):
            w = sv.kbyu(v, 1) / dg
            v = w / np.linalg.norm(w)
        return float(v @ sv.kbyu(v, 1)) / float(v @ (dg * v))   # Rayleigh quotient: a lower bound

    its = {}
    for w in (0.6, 0.5):
        sv = S.Solver(m, weight_factor_jacobi=w)
        sv.assemble(E)
        its[w] = sv.pcg(m.F[:, 0], 1e-3, 800)[1]
    lam_rand = lam_max(sv)
    su = S.Solver(m)
    su.assemble(np.full(Ne, 1e-6 + 0.12 ** 3))
    lam_unif = lam_max(su)
    assert lam_unif * 0.6 < 2.0 < lam_rand * 0.6, (lam_unif, lam_rand)
    assert lam_rand * 0.5 < 2.0
    assert its[0.5] * 3 < its[0.6], its
