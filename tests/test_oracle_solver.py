"""Pin the oracle's solver restatement by checks independent of its own formulas:
direct sparse FE assembly on the voxel grid, adjointness, symmetry, exact reproduction of
linear fields, exact Galerkin products, and convergence to a sparse direct solve."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import constants as C
from oracle import mesh as M
from oracle import solver as S
from oracle import topopt as T
from tests.conftest import holey_mask, random_modulus

VY, VX, VZ = C.VERTEX_OFFSETS[:, 0], C.VERTEX_OFFSETS[:, 1], C.VERTEX_OFFSETS[:, 2]


def direct_global_K(model, E):
    """Assemble K_bc directly on the voxel grid: solid voxels in column-major order, nodes of the
    padded grid numbered by ascending full index among nodes touched by a solid voxel; fixed DOFs
    get zero rows/columns and a unit diagonal."""
    lv = model.levels[0]
    ny, nx, nz = lv.resY, lv.resX, lv.resZ
    NY, NX = ny + 1, nx + 1
    solid = np.zeros(ny * nx * nz, bool)
    solid[lv.eleMapBack - 1] = True
    vox = np.flatnonzero(solid)
    ey, ex, ez = vox % ny, (vox // ny) % nx, vox // (ny * nx)
    full = (ey[:, None] + VY) + NY * ((ex[:, None] + VX) + NX * (ez[:, None] + VZ))
    act = np.unique(full)
    num = np.searchsorted(act, full)                              # 0-based node ids
    dofs = (3 * num[:, :, None] + np.arange(3)).reshape(len(vox), 24)
    Ke = lv.Ke
    rows = np.repeat(dofs, 24, axis=1).reshape(-1)
    cols = np.tile(dofs, (1, 24)).reshape(-1)
    vals = (E[:, None, None] * Ke[None]).reshape(-1)
    n = 3 * len(act)
    K = sp.csr_matrix((vals, (rows, cols)), shape=(n, n)).tolil()
    fixed = np.flatnonzero(model.fixedDOFs)
    K[fixed, :] = 0
    K[:, fixed] = 0
    for d in fixed:
        K[d, d] = 1.0
    return K.tocsr()


@pytest.fixture(scope='module', params=['box', 'holes'])
def setup(request):
    if request.param == 'box':
        m = M.build_cuboid_model(8, 16, 8, opt_case=1)
    else:
        m = M.build_cuboid_model(12, 20, 12, opt_case=1, voxels=holey_mask(12, 20, 12, seed=3))
    E = random_modulus(m.levels[0].numElements)
    sv = S.Solver(m)
    sv.assemble(E)
    return m, E, sv


def test_kbyu_equals_direct_assembly(setup):
    m, E, sv = setup
    K = direct_global_K(m, E)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(K.shape[0])
    y = sv.kbyu(u, 1)
    assert np.abs(y - K @ u).max() <= 1e-13 * np.abs(K @ u).max()


def test_gather_scatter_brute_force_and_adjoint():
    rng = np.random.default_rng(1)
    Nn, Ne = 17, 9
    eNod = rng.integers(1, Nn + 1, size=(Ne, 8)).astype(np.int32)
    eNod[0, 3] = Nn + 4                                           # out of range: skipped by scatter
    u = rng.standard_normal((Nn, 3))
    uMat = S.gathering(u, np.clip(eNod, 1, Nn))
    for e in range(Ne):
        for n in range(8):
            for c in range(3):
                assert uMat[e, 3 * n + c] == u[min(eNod[e, n], Nn) - 1, c]
    V = rng.standard_normal((Ne, 24))
    Y = S.scattering(np.zeros((Nn, 3)), eNod, V)
    Yb = np.zeros((Nn, 3))
    for e in range(Ne):
        for j in range(8):
            if 1 <= eNod[e, j] <= Nn:
                Yb[eNod[e, j] - 1] += V[e, 3 * j:3 * j + 3]
    assert np.abs(Y - Yb).max() < 1e-14
    e2 = np.clip(eNod, 1, Nn)
    assert abs(np.sum(S.gathering(u, e2) * V) - np.sum(u * S.scattering(np.zeros((Nn, 3)), e2, V))) < 1e-12
    keys = rng.integers(0, 7, size=40).astype(np.int32)           # includes 0: skipped
    Tm = rng.standard_normal((40, 3))
    X = S.scatter_accum3(keys, Tm, np.zeros((6, 3)))
    Xb = np.zeros((6, 3))
    for i, k in enumerate(keys):
        if 1 <= k <= 6:
            Xb[k - 1] += Tm[i]
    assert np.abs(X - Xb).max() < 1e-14


def test_transfer_adjoint_and_linear_reproduction(setup):
    m, E, sv = setup
    rng = np.random.default_rng(7)
    for ii in range(2, m.numLevels + 1):
        fine, coarse = m.levels[ii - 2], m.levels[ii - 1]
        r = rng.standard_normal(3 * fine.numNodes)
        x = rng.standard_normal(3 * coarse.numNodes)
        lhs = sv.restrict(r, ii) @ x
        rhs = r @ sv.interpolate(x, ii)
        assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)
        # a linear field on the coarse nodes is reproduced exactly on the fine nodes
        def coords(lv, scale):
            nf = lv.nodMapBack.astype(np.int64) - 1 if ii - 1 == 1 and lv is fine else None
            return nf
        cf = _node_coords(m, ii)
        ff = _node_coords(m, ii - 1)
        f = lambda p: np.stack([1 + p[:, 0] - 2 * p[:, 1], 0.5 * p[:, 2], p[:, 0] + p[:, 1] + p[:, 2]], 1)
        xf = sv.interpolate(f(cf).reshape(-1), ii).reshape(-1, 3)
        assert np.abs(xf - f(ff)).max() < 1e-12


def _node_coords(model, level):
    """Physical (finest-grid units) coordinates of a level's solid nodes."""
    lv = model.levels[level - 1]
    span = 1
    for k in range(1, level):
        span *= model.levels[k].spanWidth
    n = lv.nodMapBack.astype(np.int64) - 1 if level == 1 else None
    NY, NX = lv.resY + 1, lv.resX + 1
    if level == 1:
        n = lv.nodMapBack.astype(np.int64) - 1
        return np.stack([n % NY, (n // NY) % NX, n // (NY * NX)], 1).astype(float)
    # coarse nodMapBack holds finest-grid node numbers
    f = model.levels[0]
    NYf, NXf = f.resY + 1, f.resX + 1
    n = lv.nodMapBack.astype(np.int64) - 1
    return np.stack([n % NYf, (n // NYf) % NXf, n // (NYf * NXf)], 1).astype(float)


def test_galerkin_level2_is_exact_patch_product():
    m = M.build_cuboid_model(8, 16, 8, opt_case=1)
    E = random_modulus(m.levels[0].numElements, seed=9)
    sv = S.Solver(m)
    sv.assemble(E)
    lv2 = m.levels[1]
    # independent: per coarse cell, assemble the 5x5x5-node fine patch, interpolate its corners
    K = direct_global_K(m, E)
    fixed = m.fixedDOFs
    for e in [0, 5, lv2.numElements - 1]:
        Pb = _patch_prolongation(m, e)                           # [3*Nn_fine x 24]
        # restrict K_bc to the patch's elements only: rebuild from the children directly
        Kp = _patch_matrix(m, E, e)
        assert np.abs(Pb.T @ Kp @ Pb - lv2.Ks[:, :, e]).max() < 1e-12 * np.abs(lv2.Ks[:, :, e]).max()


def _patch_prolongation(m, e):
    lv1, lv2 = m.levels[0], m.levels[1]
    s = lv2.spanWidth
    ce = lv2.eleMapBack[e] - 1
    cy, cx, cz = ce % lv2.resY, (ce // lv2.resY) % lv2.resX, ce // (lv2.resY * lv2.resX)
    NY, NX = lv1.resY + 1, lv1.resX + 1
    P = np.zeros((3 * lv1.numNodes, 24))
    for a in range(s + 1):
        for b in range(s + 1):
            for c in range(s + 1):
                n = lv1.nodMapForward[(s * cy + a) + NY * ((s * cx + b) + NX * (s * cz + c))] - 1
                for j in range(8):
                    w = ((a / s) if VY[j] else 1 - a / s) * ((b / s) if VX[j] else 1 - b / s) * \
                        ((c / s) if VZ[j] else 1 - c / s)
                    for d in range(3):
                        P[3 * n + d, 3 * j + d] = w
    return P


def _patch_matrix(m, E, e):
    """Sum of the (BC-modified) element matrices of coarse cell e's children, globally indexed."""
    lv1 = m.levels[0]
    n = 3 * lv1.numNodes
    K = np.zeros((n, n))
    for f in m.levels[1].elementUpwardMap[e]:
        if f <= 0:
            continue
        u = m.mapUniqueKes[f - 1]
        if u > 0:
            Kf = (m.uniqueKesFree[:, u - 1] * E[f - 1] + m.uniqueKesFixed[:, u - 1]).reshape(24, 24, order='F')
        else:
            Kf = lv1.Ke * E[f - 1]
        d = (3 * (lv1.eNodMat[f - 1].astype(np.int64) - 1)[:, None] + np.arange(3)).reshape(-1)
        K[np.ix_(d, d)] += Kf
    return K


def test_galerkin_span2_child_order():
    """Level >= 3 pairs SubEleNodMat(2) rows with elementUpwardMap columns as written: children
    3/4 and 7/8 use each other's interpolation blocks (TOP_GPU.m:1941-1948 vs 1329-1336)."""
    m = M.build_cuboid_model(8, 16, 8, opt_case=1)
    sv = S.Solver(m)
    sv.assemble(random_modulus(m.levels[0].numElements, seed=4))
    lv2, lv3 = m.levels[1], m.levels[2]
    P = S.build_pch(2)
    e = 0
    up = lv3.elementUpwardMap[e]
    as_written = sum(P[:, :, s].T @ lv2.Ks[:, :, up[s] - 1] @ P[:, :, s] for s in range(8))
    # the geometric pairing: child at column-major position p uses the hex-vertex-ordered block
    perm = [0, 1, 3, 2, 4, 5, 7, 6]
    geometric = sum(P[:, :, perm[s]].T @ lv2.Ks[:, :, up[s] - 1] @ P[:, :, perm[s]] for s in range(8))
    assert np.abs(lv3.Ks[:, :, e] - as_written).max() < 1e-12 * np.abs(as_written).max()
    assert np.abs(lv3.Ks[:, :, e] - geometric).max() > 1e-6 * np.abs(geometric).max()
    # still symmetric positive semi-definite with the 6 rigid modes (interior-free cell)
    Ks = lv3.Ks[:, :, e]
    assert np.abs(Ks - Ks.T).max() < 1e-12 * np.abs(Ks).max()


def test_coarse_solve_and_vcycle_symmetry(setup):
    m, E, sv = setup
    Kc = S.coarsest_matrix(m.levels[-1])
    rng = np.random.default_rng(11)
    r = rng.standard_normal(Kc.shape[0])
    x = sv.coarse_solve(r)
    assert np.abs(Kc @ x - r).max() < 1e-10 * np.abs(r).max()
    n = 3 * m.levels[0].numNodes
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    lhs, rhs = sv.vcycle(a) @ b, a @ sv.vcycle(b)
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


@pytest.mark.parametrize('vtype', ['Standard', 'Adapted'])
def test_vcycle_equals_its_matrix_form(vtype):
    """Solving_Vcycle (TOP_GPU.m:783-815) as matrices, on a 4 x 8 x 4 cantilever with three levels.  With
    S_l = omega D_l^-1, the Standard cycle is M_l = S_l + (I - S_l A_l)(S_l + P_l M_(l+1) R_l (I - A_l S_l)),
    the Adapted one M_l = 2 S_l + P_l M_(l+1) R_l, and M_L = A_L^-1.  A_1 is assembled directly on the
    voxel grid; A_2 is the oracle's level-2 operator, D_l the diagonals (each pinned above), R_l and P_l
    the oracle's transfers applied to unit vectors, A_L^-1 the inverse of the coarsest matrix.  The
    oracle's level loop must agree to 1e-12: a dropped smoothing term, a wrong weight, or a residual
    taken from the wrong level does not."""
    m = M.build_cuboid_model(4, 8, 4, opt_case=1)
    assert m.numLevels == 3
    E = random_modulus(m.levels[0].numElements, seed=4)
    sv = S.Solver(m, vcycle_type=vtype)
    sv.assemble(E)
    n = [3 * lv.numNodes for lv in m.levels]
    w = sv.omega
    A1 = direct_global_K(m, E).toarray()
    A2 = np.stack([sv.kbyu(np.eye(n[1])[j], 2) for j in range(n[1])], axis=1)
    R = [np.stack([sv.restrict(np.eye(n[l])[j], l + 2) for j in range(n[l])], axis=1) for l in range(2)]
    P = [np.stack([sv.interpolate(np.eye(n[l + 1])[j], l + 2) for j in range(n[l + 1])], axis=1) for l in range(2)]
    S_ = [np.diag(w / m.levels[l].diagK) for l in range(2)]
    M3 = np.linalg.inv(S.coarsest_matrix(m.levels[-1]).toarray())

    def cycle(l, A, Mc):
        I = np.eye(n[l])
        if vtype == 'Adapted':
            return 2 * S_[l] + P[l] @ Mc @ R[l]
        return S_[l] + (I - S_[l] @ A) @ (S_[l] + P[l] @ Mc @ R[l] @ (I - A @ S_[l]))
    M1 = cycle(0, A1, cycle(1, A2, M3))
    r = np.random.default_rng(13).standard_normal(n[0])
    ref = M1 @ r
    got = sv.vcycle(r)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    # the two cycle types are different operators on this mesh
    other = S.Solver(m, vcycle_type='Adapted' if vtype == 'Standard' else 'Standard')
    other.assemble(E)
    assert np.abs(other.vcycle(r) - ref).max() > 1e-3 * np.abs(ref).max()


def test_pcg_converges_to_direct_solve(setup):
    m, _, _ = setup
    E = 0.2 + 0.8 * np.random.default_rng(12).random(m.levels[0].numElements)
    sv = S.Solver(m)
    sv.assemble(E)
    K = direct_global_K(m, E)
    b = m.F[:, 0]
    y, its, hist = sv.pcg(b, tol=1e-12, maxIT=400)
    ref = spla.spsolve(K.tocsc(), b)
    assert its < 400
    assert np.abs(y - ref).max() < 1e-8 * np.abs(ref).max()
    # compliance identity: F'u = sum_e E_e u_e' Ke u_e (fixed DOFs carry ~0 displacement)
    ce = T.unit_compliance(m.levels[0], y, m.levels[0].Ke)
    assert abs(b @ y - E @ ce) < 1e-9 * abs(b @ y)


def test_pde_filter_constants_and_symmetry():
    m = M.build_cuboid_model(8, 16, 8, opt_case=1)
    f = T.PDEFilter(m.levels[0], np.sqrt(3))
    f.tol = 1e-13
    Ne = m.levels[0].numElements
    assert np.abs(f(np.ones(Ne)) - 1).max() < 1e-10
    rng = np.random.default_rng(2)
    a, b = rng.random(Ne), rng.random(Ne)
    assert abs(f(a) @ b - a @ f(b)) < 1e-10 * abs(a @ f(b))


def test_sensitivity_finite_difference():
    """dc of the unfiltered compliance against central differences on a tiny cantilever."""
    m = M.build_cuboid_model(8, 8, 8, opt_case=1, voxels=np.ones((8, 8, 8), bool))
    Ne = m.levels[0].numElements
    rng = np.random.default_rng(3)
    x = 0.3 + 0.4 * rng.random(Ne)
    Emin, p = 1e-6, 3.0
    F = m.F[:, 0]

    def compliance(xx):
        sv = S.Solver(m)
        sv.assemble(Emin + xx ** p * (1 - Emin))
        u, _, _ = sv.pcg(F, tol=1e-13, maxIT=500)
        return F @ u, u

    c0, u = compliance(x)
    ce = T.unit_compliance(m.levels[0], u, m.levels[0].Ke)
    dc = -p * (1 - Emin) * x ** (p - 1) * ce
    h = 1e-6
    for e in rng.choice(Ne, 4, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (compliance(xp)[0] - compliance(xm)[0]) / (2 * h)
        assert abs(fd - dc[e]) < 1e-5 * abs(dc[e])


def test_vectorised_assembly_matches_literal_loop():
    m = M.build_cuboid_model(8, 16, 8, opt_case=1)
    E = random_modulus(m.levels[0].numElements, seed=13)
    a, b = S.Solver(m), S.Solver(m)
    a.assemble(E, fast=False)
    ks_lit = [lv.Ks.copy() for lv in m.levels[1:]]
    b.assemble(E, fast=True)
    for k, lv in zip(ks_lit, m.levels[1:]):
        assert np.abs(lv.Ks - k).max() < 1e-13 * np.abs(k).max()


@pytest.mark.parametrize('se', [False, True])
def test_passive_elements_held_at_one(se):
    """Passive elements (TOP_GPU.m:239, 247, 270, 273): xPhys is 1 there after every iteration (in
    super-element mode only sub-voxel row 4, as xnew(Num_, passiveElements) writes), the constraint
    still holds on the full design, and the history differs from the run without them."""
    shape, case = ((8, 16, 8), 1) if not se else ((8, 8, 8), 2)
    m = M.build_cuboid_model(*shape, opt_case=case, super_element=se)
    pas = np.array([0, 9, 130, 300])
    a = T.run_top(m, V0=0.12, nLoop=2, ft=2, passive=pas)
    b = T.run_top(m, V0=0.12, nLoop=2, ft=2)
    xp = np.asarray(a['xPhys'])
    if se:
        assert np.all(xp[3, pas] == 1.0) and np.all(xp[:3, pas] < 1.0)
    else:
        assert np.all(xp[pas] == 1.0)
        assert np.all(np.asarray(b['xPhys'])[pas] < 1.0)
    assert np.abs(np.array(a['cHist']) - np.array(b['cHist'])).max() > 1e-6 * abs(b['cHist'][-1])


def test_load_cases_weighted_equally():
    """Two load cases are solved one after the other and their unit compliances averaged
    (objWeightingList_ = ones / 2, TOP_GPU.m:186, 461, 599): the fully solid compliance and the first
    iteration's compliance are the means of the two single-case runs' (the same design, and each case's
    PCG warm-started from its own solid solution), and the PCG counts are theirs in turn.  A sum
    instead of the mean doubles them; a case left out gives one run's values."""
    nely, nelx, nelz = 6, 12, 6
    solid = np.arange(1, nely * nelx * nelz + 1)
    NY, NX = nely + 1, nelx + 1
    node = lambda y, x, z: 1 + y + NY * (x + NX * z)
    fix = np.array([[node(y, 0, z), 1, 1, 1] for z in range(nelz + 1) for y in range(nely + 1)], float)
    A = np.array([[node(0, nelx, z), 0.0, 0.0, -1.0] for z in range(nelz + 1)])
    B = np.array([[node(y, nelx, nelz // 2), 0.0, 1.0, 0.0] for y in range(nely + 1)])
    runs = {name: T.run_top(M.build_file_model(nelx, nely, nelz, solid, fix, loads), nLoop=1, ft=1)
            for name, loads in (('A', A), ('B', B), ('AB', [A, B]))}
    a, b, ab = runs['A'], runs['B'], runs['AB']
    assert abs(ab['cSolid'] - 0.5 * (a['cSolid'] + b['cSolid'])) < 1e-12 * ab['cSolid']
    assert abs(ab['cHist'][0] - 0.5 * (a['cHist'][0] + b['cHist'][0])) < 1e-12 * ab['cHist'][0]
    assert ab['lssIts'] == [a['lssIts'][0], b['lssIts'][0]]
    assert abs(a['cSolid'] - b['cSolid']) > 1e-3 * ab['cSolid']      # the cases differ
