"""GPU parity of the device solver (libtopgpu, through the C ABI) against the oracle."""
import os
import sys

import numpy as np
import pytest

from oracle import mesh as M
from oracle import solver as S
from tests.conftest import holey_mask, random_modulus, smooth_modulus

pytestmark = pytest.mark.gpu

CASES = {
    'box': dict(shape=(8, 16, 8), case=1, vox=None, se=False),
    'holes': dict(shape=(12, 20, 12), case=1, vox=holey_mask(12, 20, 12, seed=3), se=False),
    'mbb_super': dict(shape=(10, 20, 10), case=2, vox=None, se=True),
    # the value-noise design domain (file input, boundary-node clamps, top-patch load) on a 48 x 24 x 24
    # box, non-dyadic hierarchy
    'value_noise': dict(shape=(24, 48, 24), case=None, vox=None, se=False, noise=dict(period=32, non_dyadic=1)),
}


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope='module', params=list(CASES))
def pair(request, tmp_path_factory):
    import torch
    from top_gpu_on_matlab_amd.model import Model
    from top_gpu_on_matlab_amd.solver import Solver
    c = CASES[request.param]
    ny, nx, nz = c['shape']
    print('\n[case %s shape %s]' % (request.param, c['shape']), flush=True)
    if c.get('noise'):
        from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise_mask
        from top_gpu_on_matlab_amd.topvoxel import load_topvoxel_model, read_topvoxel, write_topvoxel
        mask = value_noise_mask(nx, ny, nz, seed=0, octaves=4, period=c['noise']['period'])
        path = os.path.join(str(tmp_path_factory.mktemp('noise')), 'value_noise.TopVoxel')
        write_topvoxel(path, mask, *bottom_clamp_top_patch(mask))
        d = read_topvoxel(path)
        om = M.build_file_model(d['nelx'], d['nely'], d['nelz'], d['voxels'], d['fixing'], d['loads'][0],
                                non_dyadic=c['noise']['non_dyadic'])
        gm = load_topvoxel_model(path, non_dyadic=c['noise']['non_dyadic'])
    else:
        vox = c['vox'] if c['vox'] is not None else np.ones(c['shape'], bool)
        om = M.build_cuboid_model(ny, nx, nz, opt_case=c['case'], voxels=vox, super_element=c['se'])
        gm = Model(vox, super_element=c['se'])
        gm.builtin_case(c['case'])
        gm.finalize()
    Ne = om.levels[0].numElements
    E = random_modulus(Ne * (64 if c['se'] else 1), seed=2)
    if c['se']:
        E = E.reshape(64, Ne, order='F')
    osv = S.Solver(om)
    osv.assemble(E)
    gsv = Solver(gm)
    gsv.set_modulus(np.asarray(E).reshape(-1, order='F'))
    gsv.assemble()
    torch.cuda.synchronize()
    return request.param, om, osv, gm, gsv, E


def test_kbyu_all_levels(pair):
    name, om, osv, gm, gsv, E = pair
    rng = np.random.default_rng(1)
    for level in range(1, om.numLevels + 1):
        n = 3 * om.levels[level - 1].numNodes
        u = rng.standard_normal(n)
        ref = osv.kbyu(u, level)
        got = gsv.kbyu(u, level).cpu().numpy()
        assert _rel(got, ref) < 1e-13, (name, level, _rel(got, ref))


def test_kbyu_element_matrix_forms_agree(pair):
    """The FP64 per-element-matrix K*u (coarse levels, and the finest level of a super-element model)
    from the packed upper triangles (the default) and from the full matrices split over 24 and 4
    threads per element: each within 1e-13 of the oracle, the two splits bitwise equal (each output
    is one thread's sum in the same order)."""
    name, om, osv, gm, gsv, E = pair
    rng = np.random.default_rng(5)
    try:
        for level in range(1, om.numLevels + 1):
            n = 3 * om.levels[level - 1].numNodes
            u = rng.standard_normal(n)
            ref = osv.kbyu(u, level)
            got = {}
            for parts in (1, 24, 4):
                gsv.set_tuning(kelem_parts=parts)
                got[parts] = gsv.kbyu(u, level).cpu().numpy()
                assert _rel(got[parts], ref) < 1e-13, (name, level, parts, _rel(got[parts], ref))
            assert np.array_equal(got[24], got[4]), (name, level)
    finally:
        gsv.set_tuning(kelem_parts=1)


@pytest.mark.parametrize('nu', [0.3, 0.27])
def test_finest_kbyu_kernels_agree(nu):
    """The LDS-tiled Walsh-basis K*u (compile-time unit matrix for nu = 0.3, runtime matrix
    otherwise) against the node-centric gather kernel and the oracle, on a perforated grid whose
    z extent spans several tile chunks (chunk sizes 1, 3 and 16 node planes)."""
    import torch
    from top_gpu_on_matlab_amd.model import Model
    from top_gpu_on_matlab_amd.solver import Solver
    shape = (20, 33, 18)                      # 2 x 3 tiles of 15 x 15 nodes in (y, x)
    vox = holey_mask(*shape, seed=7, frac=0.7)
    om = M.build_cuboid_model(*shape, opt_case=1, voxels=vox, nu=nu)
    gm = Model(vox, poisson_ratio=nu)
    gm.builtin_case(1)
    gm.finalize()
    E = random_modulus(om.levels[0].numElements, seed=5)
    osv = S.Solver(om)
    osv.assemble(E)
    gsv = Solver(gm)
    gsv.set_modulus(E)
    gsv.assemble()
    u = np.random.default_rng(6).standard_normal(3 * om.levels[0].numNodes)
    ref = osv.kbyu(u, 1)
    gsv.set_tuning(kfine_kernel='gather')
    gat = gsv.kbyu(u).cpu().numpy()
    assert _rel(gat, ref) < 1e-13
    gsv.set_tuning(kfine_kernel='tiled')
    for zc in (1, 3, 16):
        gsv.set_tuning(kfine_zchunk=zc)
        til = gsv.kbyu(u).cpu().numpy()
        assert _rel(til, ref) < 1e-13, (nu, zc, _rel(til, ref))
        assert _rel(til, gat) < 1e-13
    torch.cuda.synchronize()


def test_transfer_operators(pair):
    name, om, osv, gm, gsv, E = pair
    rng = np.random.default_rng(2)
    for ii in range(2, om.numLevels + 1):
        rf = rng.standard_normal(3 * om.levels[ii - 2].numNodes)
        xc = rng.standard_normal(3 * om.levels[ii - 1].numNodes)
        assert _rel(gsv.restrict(rf, ii).cpu().numpy(), osv.restrict(rf, ii)) < 1e-13
        assert _rel(gsv.interpolate(xc, ii).cpu().numpy(), osv.interpolate(xc, ii)) < 1e-13


def test_galerkin_ks_and_diagonals(pair):
    name, om, osv, gm, gsv, E = pair
    for level in range(2, om.numLevels + 1):
        ks = gsv.ks(level).cpu().numpy()                       # [e][row][col]
        ref = om.levels[level - 1].Ks.transpose(2, 0, 1)
        assert _rel(ks, ref) < 1e-13, (name, level)
    for level in range(1, om.numLevels):
        d = gsv.diag(level).cpu().numpy()
        assert _rel(d, om.levels[level - 1].diagK) < 1e-13, (name, level)


def test_coarse_solve(pair):
    name, om, osv, gm, gsv, E = pair
    r = np.random.default_rng(3).standard_normal(3 * om.levels[-1].numNodes)
    assert _rel(gsv.coarse_solve(r).cpu().numpy(), osv.coarse_solve(r)) < 1e-9


def test_coarse_solve_graph_replay_matches_direct_launches(pair):
    """The step-by-step coarsest solve launched kernel by kernel and replayed as a captured HIP graph
    run the same kernels on the same buffers: bitwise-identical results, on repeated replays."""
    name, om, osv, gm, gsv, E = pair
    r = np.random.default_rng(4).standard_normal(3 * om.levels[-1].numNodes)
    try:
        gsv.set_tuning(coarse_persistent=False)
        direct = gsv.coarse_solve(r).clone()
        gsv.set_tuning(coarse_graph=True)
        for k in range(3):
            assert gsv.coarse_solve(r).equal(direct), (name, k)
            assert gsv.coarse_solve(2.0 * r).equal(2.0 * direct), (name, k)   # scaling by 2 is exact
        gsv.set_tuning(coarse_graph=False)
        assert gsv.coarse_solve(r).equal(direct)
    finally:
        gsv.set_tuning(coarse_graph=False, coarse_persistent=True)


def test_coarse_solve_persistent_matches_step_by_step(pair):
    """The coarsest solve as one persistent launch (stages passing the recurrence vectors to each other,
    the block-diagonal product folded into the forward recurrence) against the step-by-step launches on
    the same factors: the same recurrences summed in another order, so equal to rounding (1e-12
    relative); and repeated persistent launches, whose vectors are refilled with the "not written"
    sentinel for every solve, give bitwise the same result every time, also after step-by-step
    launches in between."""
    name, om, osv, gm, gsv, E = pair
    r = np.random.default_rng(5).standard_normal(3 * om.levels[-1].numNodes)
    try:
        gsv.set_tuning(coarse_persistent=False)
        step = gsv.coarse_solve(r).clone()
        gsv.set_tuning(coarse_persistent=True)
        pers = gsv.coarse_solve(r).clone()
        print('[%s] persistent vs step by step: %.2e' % (name, _rel(pers.cpu().numpy(), step.cpu().numpy())), flush=True)
        assert _rel(pers.cpu().numpy(), step.cpu().numpy()) < 1e-12, name
        assert _rel(pers.cpu().numpy(), osv.coarse_solve(r)) < 1e-9, name
        for k in range(6):
            assert gsv.coarse_solve(r).equal(pers), (name, k)
            assert gsv.coarse_solve(2.0 * r).equal(2.0 * pers), (name, k)
        gsv.set_tuning(coarse_persistent=False)
        assert gsv.coarse_solve(r).equal(step), name
        gsv.set_tuning(coarse_persistent=True)
        assert gsv.coarse_solve(r).equal(pers), name
    finally:
        gsv.set_tuning(coarse_persistent=True)


def test_vcycle(pair):
    name, om, osv, gm, gsv, E = pair
    r = np.random.default_rng(4).standard_normal(3 * om.levels[0].numNodes)
    assert _rel(gsv.vcycle(r).cpu().numpy(), osv.vcycle(r)) < 1e-9


def test_vcycle_coarse_solve_forms_agree(pair):
    """The V-cycle with its coarsest solve as the persistent launch (the right-hand side packed into
    plane order by the restriction's last gather, the solution written to the full grid by the
    launch itself) and step by step (separate pack and unpack launches): equal to rounding, and
    switching back to the persistent form reproduces its result bitwise."""
    name, om, osv, gm, gsv, E = pair
    r = np.random.default_rng(6).standard_normal(3 * om.levels[0].numNodes)
    try:
        pers = gsv.vcycle(r).clone()
        gsv.set_tuning(coarse_persistent=False)
        step = gsv.vcycle(r).clone()
        gsv.set_tuning(coarse_persistent=True)
        again = gsv.vcycle(r).clone()
    finally:
        gsv.set_tuning(coarse_persistent=True)
    print('[%s] V-cycle, persistent vs step-by-step coarsest solve: %.2e'
          % (name, _rel(pers.cpu().numpy(), step.cpu().numpy())), flush=True)
    assert _rel(pers.cpu().numpy(), step.cpu().numpy()) < 1e-10, name
    assert pers.equal(again), name


def test_pcg_parity_and_determinism(pair):
    """PCG on a smooth density field against the oracle.  The device and the oracle differ only by
    rounding (summation order in element products, transfers and the coarse solve), which CG
    amplifies; the bound is calibrated on the oracle itself: a relative perturbation of b at the
    unit-roundoff level (2^-52) moves the oracle's compliance b'y and iterate by some spread, and
    the device must agree with the oracle to within 20x that spread with the same iteration count --
    or, where the perturbed oracle's own count moves (a residual that stagnates near tol, as on the
    value-noise domain, where the spread reaches ~1e-7 in compliance), within 3x that move.  Where
    the oracle is well conditioned (spread below 5e-10) the compliance must agree to 1e-8."""
    name, om, osv, gm, gsv, E = pair
    Es = smooth_modulus(om.levels[0], per=64 if om.super_element else 1)
    osv = S.Solver(om)
    osv.assemble(Es)
    gsv.set_modulus(np.asarray(Es).reshape(-1, order='F'))
    gsv.assemble()
    b = om.F[:, 0]
    y_ref, its_ref, hist = osv.pcg(b, tol=1e-3, maxIT=800)
    bp = b * (1 + 2.0 ** -52 * np.random.default_rng(11).standard_normal(b.shape))
    y_per, its_per, _ = osv.pcg(bp, tol=1e-3, maxIT=800)
    c_ref = float(b @ y_ref)
    c_spread = abs(float(b @ y_per) - c_ref) / abs(c_ref)
    y_spread = _rel(y_per, y_ref)
    y, its, res = gsv.pcg(b, tol=1e-3, max_it=800)
    y2, its2, res2 = gsv.pcg(b, tol=1e-3, max_it=800)
    c_dev = abs(float(b @ y.cpu().numpy()) - c_ref) / abs(c_ref)
    print('\n[%s] its %d; oracle ulp-spread: compliance %.1e iterate %.1e; device: compliance %.1e iterate %.1e'
          % (name, its_ref, c_spread, y_spread, c_dev, _rel(y.cpu().numpy(), y_ref)), flush=True)
    its_slack = 3 * abs(its_per - its_ref)
    assert abs(its - its_ref) <= its_slack, (name, its, its_ref, its_per, hist[-3:], res)
    assert c_dev < max(20 * c_spread, 1e-12)
    assert c_dev < 1e-8 or c_spread > 5e-10
    assert _rel(y.cpu().numpy(), y_ref) < max(20 * y_spread, 1e-10)
    assert its2 == its and bool((y == y2).all()), 'PCG is not bitwise reproducible'
    # warm start from the previous solution (TOP_GPU.m:186)
    y3, its3, _ = gsv.pcg(b, y0=y, tol=1e-3, max_it=800)
    y3r, its3r, _ = osv.pcg(b, tol=1e-3, maxIT=800, y0=y.cpu().numpy())
    assert abs(its3 - its3r) <= its_slack
    assert _rel(y3.cpu().numpy(), y3r) < max(20 * y_spread, 1e-10)
    gsv.set_modulus(np.asarray(E).reshape(-1, order='F'))      # restore the fixture's state
    gsv.assemble()


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_adapted_vcycle_and_pcg_against_oracle(precision):
    """The Adapted V-cycle (TOP_GPU.m:787-788, 808-809): every level restricts its own residual and
    smooths with omega D^-1 r alone, without K*x.  On a perforated 12 x 20 x 12 cantilever, against the
    oracle's Adapted V-cycle: within 1e-9 in FP64 and 2e-5 for the FP32 V-cycle of the mixed mode.  PCG
    preconditioned with it takes the oracle's iteration count (one more or fewer at most) and reaches
    its compliance to 1e-8 (FP64) or 1e-6 (mixed)."""
    import torch
    from top_gpu_on_matlab_amd.model import Model
    from top_gpu_on_matlab_amd.solver import Solver
    shape = (12, 20, 12)
    vox = holey_mask(*shape, seed=3)
    om = M.build_cuboid_model(*shape, opt_case=1, voxels=vox)
    gm = Model(vox, precision=precision, vcycle_type='Adapted')
    gm.builtin_case(1)
    gm.finalize()
    Es = smooth_modulus(om.levels[0])
    osv = S.Solver(om, vcycle_type='Adapted')
    osv.assemble(Es)
    gsv = Solver(gm)
    gsv.set_modulus(np.asarray(Es))
    gsv.assemble()
    r = np.random.default_rng(12).standard_normal(3 * om.levels[0].numNodes)
    err = _rel(gsv.vcycle(r).cpu().numpy(), osv.vcycle(r))
    standard = S.Solver(om)
    standard.assemble(Es)
    print('\n[adapted %s] V-cycle vs oracle %.1e (the Standard cycle differs by %.1e)'
          % (precision, err, _rel(standard.vcycle(r), osv.vcycle(r))), flush=True)
    assert err < (1e-9 if precision == 'fp64' else 2e-5)
    b = om.F[:, 0]
    y_ref, its_ref, _ = osv.pcg(b, tol=1e-3, maxIT=800)
    y, its, _ = gsv.pcg(b, tol=1e-3, max_it=800)
    c_ref = float(b @ y_ref)
    c_dev = abs(float(b @ y.cpu().numpy()) - c_ref) / abs(c_ref)
    print('[adapted %s] PCG its %d (oracle %d), compliance %.1e' % (precision, its, its_ref, c_dev), flush=True)
    assert abs(its - its_ref) <= 1
    assert c_dev < (1e-8 if precision == 'fp64' else 1e-6)
    torch.cuda.synchronize()


@pytest.mark.parametrize('shape', [(28, 12, 12), (12, 12, 28), (16, 16, 8)], ids=['y_longest', 'z_longest', 'y_x_tie'])
def test_coarse_plane_axes(shape):
    """The coarsest solve orders its node planes along the coarsest grid's longest axis (y first on a
    tie, then x); the other cases run along x.  Perforated cantilevers whose coarsest grid is longest in
    y, longest in z, and tied between y and x: the coarsest solve, persistent and step by step, within
    1e-9 of the oracle and 1e-10 of each other; the V-cycle (whose restriction packs the persistent
    solve's right-hand side in plane order) within 1e-9 in FP64 and 2e-5 with the mixed mode's FP32
    cycle.  FP64 PCG takes the oracle's iteration count (one more or fewer at most) and reaches its
    compliance within 20x the oracle's own spread under a 2^-52 perturbation of b (the z-longest case
    spreads by ~5e-8, the others by ~1e-15).  Mixed-mode PCG converges within twice the oracle's count and
    1e-5 of its compliance: its FP32 V-cycle is a slightly inexact preconditioner, which costs CG more
    iterations on the ill-conditioned z-longest case (51 against 36; 22 against 22 on y-longest)."""
    import torch
    from top_gpu_on_matlab_amd.model import Model
    from top_gpu_on_matlab_amd.solver import Solver
    vox = holey_mask(*shape, seed=5)
    om = M.build_cuboid_model(*shape, opt_case=1, voxels=vox)
    Es = smooth_modulus(om.levels[0])
    osv = S.Solver(om)
    osv.assemble(Es)
    sv = {}
    for p in ('fp64', 'mixed'):
        gm = Model(vox, precision=p)
        gm.builtin_case(1)
        gm.finalize()
        s = Solver(gm)
        s.set_modulus(np.asarray(Es))
        s.assemble()
        sv[p] = s
    g = sv['fp64']
    rng = np.random.default_rng(14)
    rc = rng.standard_normal(3 * om.levels[-1].numNodes)
    ref = osv.coarse_solve(rc)
    pers = g.coarse_solve(rc).cpu().numpy()
    try:
        g.set_tuning(coarse_persistent=False)
        step = g.coarse_solve(rc).cpu().numpy()
    finally:
        g.set_tuning(coarse_persistent=True)
    print('\n[%s] coarsest solve vs oracle: persistent %.1e, step %.1e' % (shape, _rel(pers, ref), _rel(step, ref)),
          flush=True)
    assert _rel(pers, ref) < 1e-9 and _rel(step, ref) < 1e-9 and _rel(pers, step) < 1e-10
    r = rng.standard_normal(3 * om.levels[0].numNodes)
    zref = osv.vcycle(r)
    assert _rel(g.vcycle(r).cpu().numpy(), zref) < 1e-9
    assert _rel(sv['mixed'].vcycle(r).cpu().numpy(), zref) < 2e-5
    b = om.F[:, 0]
    y_ref, its_ref, _ = osv.pcg(b, tol=1e-3, maxIT=800)
    c_ref = float(b @ y_ref)
    bp = b * (1 + 2.0 ** -52 * np.random.default_rng(11).standard_normal(b.shape))
    c_spread = abs(float(b @ osv.pcg(bp, tol=1e-3, maxIT=800)[0]) - c_ref) / abs(c_ref)
    for p, s in sv.items():
        y, its, _ = s.pcg(b, tol=1e-3, max_it=800)
        c_dev = abs(float(b @ y.cpu().numpy()) - c_ref) / abs(c_ref)
        print('[%s %s] PCG its %d (oracle %d), compliance %.1e (oracle spread %.1e)'
              % (shape, p, its, its_ref, c_dev, c_spread), flush=True)
        if p == 'fp64':
            assert abs(its - its_ref) <= 1
            assert c_dev < max(20 * c_spread, 1e-12)
        else:
            assert its <= 2 * its_ref
            assert c_dev < 1e-5
    torch.cuda.synchronize()


def test_coarse_solve_planes_wider_than_1024_dofs():
    """A 144 x 144 x 144 cantilever: its coarsest grid has 18^3 elements, so a node plane holds 361
    nodes, 1083 DOFs, more than the 1024 the persistent solve keeps in registers.  The solve then runs
    step by step, with the solve operators' rows padded to 2048 and processed in 1024-wide chunks (every
    other case has planes of at most 867 DOFs).  The oracle is too slow at this size, so the checks are:
    the residual of the coarsest system, K_c x = r to 1e-10 (K_c x from the coarsest-level K*u, checked
    against the oracle at small sizes); the captured-graph replay equal to the direct launches bit for
    bit; PCG with the V-cycle around it reaching its tolerance, confirmed by the finest K*u."""
    import torch
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver
    s = Solver(cuboid_model(144, 144, 144, 1, precision='fp64'))
    L = s.num_levels
    coarsest = s.levels[-1]
    assert L == 3 and (coarsest['resY'], coarsest['resX'], coarsest['resZ']) == (18, 18, 18)
    # a smooth density field, the kind an optimisation produces (an unfiltered random one makes the
    # damped-Jacobi weight 0.6 too large, whatever the coarsest solve)
    e = np.arange(144 ** 3)
    ey, ex, ez = e % 144, (e // 144) % 144, e // 144 ** 2
    x = 0.12 + 0.05 * np.sin(ex / 7.0) * np.cos(ey / 5.0) * np.sin(ez / 6.0 + 1)
    s.set_modulus(1e-6 + x ** 3 * (1 - 1e-6))
    s.assemble()
    g = torch.Generator().manual_seed(3)
    r = torch.randn(s.ndof(L), generator=g, dtype=torch.float64).cuda()
    x = s.coarse_solve(r).clone()
    res = float((s.kbyu(x, L) - r).norm() / r.norm())
    try:
        s.set_tuning(coarse_graph=True)
        xg = s.coarse_solve(r).clone()
    finally:
        s.set_tuning(coarse_graph=False)
    b = torch.randn(s.ndof(1), generator=g, dtype=torch.float64).cuda()
    y, its, resnorm = s.pcg(b, tol=1e-6, max_it=400)
    true_res = float((s.kbyu(y, 1) - b).norm() / b.norm())
    print('\n[144^3, planes of 1083 DOFs] coarsest residual %.1e; PCG %d its, residual %.1e (reported %.1e)'
          % (res, its, true_res, resnorm), flush=True)
    assert res < 1e-10
    assert xg.equal(x)
    assert its < 400 and true_res < 1.5e-6


def test_reassembly_is_bitwise_reproducible():
    """Assembling the same moduli again must give bitwise the same hierarchy: coarse element
    matrices, Jacobi diagonals, coarsest factorisation (checked through its solve) and V-cycle.
    96 x 192 x 96 cantilever (coarsest level 12675 DOFs in 25 planes), twelve reassemblies."""
    import torch
    from top_gpu_on_matlab_amd import top_gpu as G
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver
    s = Solver(cuboid_model(96, 192, 96, 1, precision='fp64'))
    rng = np.random.default_rng(0)
    x = torch.from_numpy(0.12 + rng.uniform(-0.05, 0.05, s.model.num_elements)).cuda()
    E = G.simp(x, 1e-6, 1.0, 3.0)
    s.set_modulus(E)
    s.assemble()
    L = s.num_levels
    r = torch.from_numpy(rng.standard_normal(s.ndof(1))).cuda()
    rc = torch.from_numpy(rng.standard_normal(s.ndof(L))).cuda()
    ks0 = [s.ks(l).clone() for l in range(2, L + 1)]
    dg0 = [s.diag(l).clone() for l in range(1, L)]
    c0 = s.coarse_solve(rc).clone()
    z0 = s.vcycle(r).clone()
    for k in range(12):
        s.set_modulus(E)
        s.assemble()
        assert all(torch.equal(a, s.ks(l)) for a, l in zip(ks0, range(2, L + 1))), k
        assert all(torch.equal(a, s.diag(l)) for a, l in zip(dg0, range(1, L))), k
        c = s.coarse_solve(rc)
        assert torch.equal(c, c0), (k, float((c - c0).abs().max() / c0.abs().max()))
        assert torch.equal(s.vcycle(r), z0), k


def test_coarse_solve_persistent_alternating_right_hand_sides():
    """The persistent coarsest solve at a realistic size (96 x 192 x 96 cantilever: 12675 DOFs in 25
    planes of 507), where its workgroups run stages ahead of each other: right-hand sides alternate
    from launch to launch (a stale or early-read vector from a neighbouring launch or stage would show),
    each result within 1e-12 of the step-by-step solve of the same system and bitwise equal to the
    persistent solve of the same right-hand side before."""
    import torch
    from top_gpu_on_matlab_amd import top_gpu as G
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver
    s = Solver(cuboid_model(96, 192, 96, 1, precision='fp64'))
    rng = np.random.default_rng(7)
    x = torch.from_numpy(0.12 + rng.uniform(-0.05, 0.05, s.model.num_elements)).cuda()
    s.set_modulus(G.simp(x, 1e-6, 1.0, 3.0))
    s.assemble()
    L = s.num_levels
    rs = [torch.from_numpy(rng.standard_normal(s.ndof(L))).cuda() for _ in range(3)]
    try:
        s.set_tuning(coarse_persistent=False)
        step = [s.coarse_solve(r).clone() for r in rs]
        s.set_tuning(coarse_persistent=True)
        pers = [s.coarse_solve(r).clone() for r in rs]
        for a, b in zip(pers, step):
            assert _rel(a.cpu().numpy(), b.cpu().numpy()) < 1e-12
        for k in range(12):
            i = (5 * k + 1) % 3
            assert torch.equal(s.coarse_solve(rs[i]), pers[i]), k
    finally:
        s.set_tuning(coarse_persistent=True)


def test_coarse_solve_persistent_give_up_is_an_error():
    """A stage of the persistent coarsest solve that waits too long for its input gives up, every other
    wait of the launch then ends at once, the launch completes, and the solve reports TG_ERR_HIP instead
    of returning a result.  Forced from the host with a poll limit of 1 on the 96 x 192 x 96 cantilever
    (25 planes: later stages wait several microseconds); afterwards the default limit gives the same
    solution as before, bitwise."""
    import torch
    from top_gpu_on_matlab_amd import _lib
    from top_gpu_on_matlab_amd import top_gpu as G
    from top_gpu_on_matlab_amd._lib import TopGpuError
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver
    s = Solver(cuboid_model(96, 192, 96, 1, precision='fp64'))
    rng = np.random.default_rng(9)
    x = torch.from_numpy(0.12 + rng.uniform(-0.05, 0.05, s.model.num_elements)).cuda()
    s.set_modulus(G.simp(x, 1e-6, 1.0, 3.0))
    s.assemble()
    r = torch.from_numpy(rng.standard_normal(s.ndof(s.num_levels))).cuda()
    ref = s.coarse_solve(r).clone()
    failed = 0
    try:
        s.set_tuning(coarse_polls=1)
        for _ in range(5):
            try:
                s.coarse_solve(r)
            except TopGpuError as e:
                assert e.code == _lib.TG_ERR_HIP and 'timed out' in str(e), e
                failed += 1
                break
    finally:
        s.set_tuning(coarse_polls=1 << 17)
    assert failed == 1
    assert torch.equal(s.coarse_solve(r), ref)


@pytest.mark.parametrize('super_element', [False, True])
def test_mixed_vcycle_matches_fp64(super_element):
    """The FP32 V-cycle (FP32 smoother, transfers and per-element matrices, the latter packed
    symmetric; FP64 coarsest solve) against the FP64 V-cycle on the same system: they differ by FP32
    rounding only, and the FP32 one stays symmetric to that level."""
    import torch
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver
    shape, case = ((10, 20, 10), 2) if super_element else ((24, 48, 24), 1)
    sv = {}
    for p in ('fp64', 'mixed'):
        s = Solver(cuboid_model(*shape, case, precision=p, super_element=super_element))
        n = s.model.num_elements * (64 if super_element else 1)
        s.set_modulus(torch.from_numpy(0.1 + np.random.default_rng(5).uniform(0, 1, n)).cuda())
        s.assemble()
        sv[p] = s
    rng = np.random.default_rng(7)
    r1, r2 = (torch.from_numpy(rng.standard_normal(sv['fp64'].ndof(1))).cuda() for _ in range(2))
    z64, z32 = sv['fp64'].vcycle(r1), sv['mixed'].vcycle(r1)
    assert float((z32 - z64).abs().max() / z64.abs().max()) < 2e-5
    a, b = float(torch.dot(sv['mixed'].vcycle(r1), r2)), float(torch.dot(r1, sv['mixed'].vcycle(r2)))
    assert abs(a - b) < 2e-5 * abs(a)


@pytest.mark.parametrize('nu', [0.3, 0.27])
def test_mixed_vcycle_matrix_core_element_products(nu):
    """The mixed V-cycle's finest element products on the matrix cores (dense 24 x 24: MFMA f32 32x32x2 with
    FP32 operands, and MFMA f32 32x32x16 bf16 with BF16 operands and FP32 accumulation, once with each
    operand as one BF16 value and once split into two) against the Walsh form on the VALU (the default),
    on a perforated grid spanning several tiles and z chunks (chunk sizes 1, 3, 16): the FP32 products
    differ from the FP64 V-cycle by FP32 rounding only; the BF16 ones by the operands' 8-bit mantissas
    (element products to ~2^-9), the split BF16 ones by 16 bits (~2^-17)."""
    import torch
    from top_gpu_on_matlab_amd.model import Model
    from top_gpu_on_matlab_amd.solver import Solver
    shape = (20, 33, 18)
    vox = holey_mask(*shape, seed=7, frac=0.7)
    sv = {}
    for p in ('fp64', 'mixed'):
        gm = Model(vox, poisson_ratio=nu, precision=p)
        gm.builtin_case(1)
        gm.finalize()
        s = Solver(gm)
        s.set_modulus(random_modulus(gm.num_elements, seed=5))
        s.assemble()
        sv[p] = s
    r = torch.from_numpy(np.random.default_rng(8).standard_normal(sv['fp64'].ndof(1))).cuda()
    z64 = sv['fp64'].vcycle(r)
    sv['mixed'].set_tuning(kfine_mfma='walsh')
    zw = sv['mixed'].vcycle(r)
    dw = float((zw - z64).abs().max() / z64.abs().max())
    # bf16x3 must beat plain BF16 (measured 1.8e-4 and more at 128 x 256 x 128) by an order of magnitude
    for mode, bar in (('mfma32', 2e-5), ('bf16', 2e-2), ('bf16x3', 5e-5)):
        sv['mixed'].set_tuning(kfine_mfma=mode)
        for zc in (1, 3, 16):
            sv['mixed'].set_tuning(kfine_zchunk=zc)
            zm = sv['mixed'].vcycle(r)
            dm = float((zm - z64).abs().max() / z64.abs().max())
            dmw = float((zm - zw).abs().max() / zw.abs().max())
            print('nu %.2f %s chunk %d: vs fp64 %.1e (walsh vs fp64 %.1e), vs walsh %.1e' % (nu, mode, zc, dm, dw, dmw),
                  flush=True)
            assert dm < bar and dmw < bar, (mode, zc, dm, dw, dmw)
            assert torch.equal(sv['mixed'].vcycle(r), zm)  # deterministic
    sv['mixed'].set_tuning(kfine_mfma='walsh', kfine_zchunk=12)
    torch.cuda.synchronize()
