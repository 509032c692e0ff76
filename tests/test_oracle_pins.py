"""Independent pins for the oracle functions that the standard-mode checks leave open: the
super-element K*u and Galerkin product, the super-element energies, the convolution density filter
with its index maps, and the Jacobi diagonals.

Each check computes the same quantity by a route that shares no code with the oracle function it
pins: a global matrix assembled element by element on the voxel grid with the boundary conditions
applied to the assembled matrix (not through the BC element tables), a trilinear prolongation built
from node coordinates, a direct sparse solve, central differences, and shift-and-add correlation."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import constants as C
from oracle import mesh as M
from oracle import solver as S
from oracle import topopt as T
from tests.conftest import holey_mask, random_modulus
from tests.test_oracle_solver import direct_global_K

VY, VX, VZ = C.VERTEX_OFFSETS[:, 0], C.VERTEX_OFFSETS[:, 1], C.VERTEX_OFFSETS[:, 2]


# ---------------------------------------------------------------------------------------------
# independent constructions
# ---------------------------------------------------------------------------------------------

def _apply_bc_global(K, fixed):
    """Zero the fixed rows and columns of an assembled matrix and put 1 on their diagonal."""
    K = K.tolil()
    f = np.flatnonzero(fixed)
    K[f, :] = 0
    K[:, f] = 0
    for d in f:
        K[d, d] = 1.0
    return K.tocsr()


def _element_dofs(eNodMat):
    """Global DOF numbers (0-based) of each element, node-major, component fastest."""
    n = np.asarray(eNodMat, dtype=np.int64) - 1
    return (3 * n[:, :, None] + np.arange(3)).reshape(n.shape[0], 24)


def _assemble(eNodMat, mats, ndof):
    """sum_e P_e^T K_e P_e for element matrices mats [24 x 24 x Ne] (element DOF order 3n + c)."""
    dofs = _element_dofs(eNodMat)
    rows = np.repeat(dofs, 24, axis=1).reshape(-1)          # row index i for entry (i, j), j fastest
    cols = np.tile(dofs, (1, 24)).reshape(-1)
    vals = np.moveaxis(mats, 2, 0).reshape(-1)              # [Ne, i, j] row-major -> (i, j) with j fastest
    return sp.csr_matrix((vals, (rows, cols)), shape=(ndof, ndof))


def direct_global_K_super(model, Esub):
    """Global super-element stiffness: element e contributes sum_q Esub[q, e] I_q; the fixed DOFs are
    then made identity rows/columns of the assembled matrix."""
    lv = model.levels[0]
    Is = model.I_mat.reshape(24, 24, 64, order='F')
    mats = np.einsum('ijq,qe->ije', Is, Esub)
    K = _assemble(lv.eNodMat, mats, 3 * lv.numNodes)
    return _apply_bc_global(K, model.fixedDOFs)


def _own_coords(model, level):
    """Node coordinates (y, x, z) of a level's solid nodes on that level's own grid."""
    lv = model.levels[level - 1]
    NY, NX = lv.resY + 1, lv.resX + 1
    n = np.flatnonzero(lv.nodMapForward)                    # full own-grid index, in solid-node order
    order = lv.nodMapForward[n].astype(np.int64) - 1
    c = np.empty((lv.numNodes, 3), dtype=np.int64)
    c[order] = np.stack([n % NY, (n // NY) % NX, n // (NY * NX)], 1)
    return c


def geometric_prolongation(model, ii):
    """Trilinear interpolation from coarse level ii to level ii - 1 (1-based), built from the node
    coordinates: each fine node takes the trilinear weights of the coarse cell containing it."""
    fine, coarse = model.levels[ii - 2], model.levels[ii - 1]
    s = coarse.spanWidth
    cf = _own_coords(model, ii - 1)
    NYc, NXc = coarse.resY + 1, coarse.resX + 1
    rows, cols, vals = [], [], []
    for j, (y, x, z) in enumerate(cf):
        cell = [min(v // s, r - 1) for v, r in zip((y, x, z), (coarse.resY, coarse.resX, coarse.resZ))]
        t = [(v - s * c) / s for v, c in zip((y, x, z), cell)]
        for dy in (0, 1):
            for dx in (0, 1):
                for dz in (0, 1):
                    w = (t[0] if dy else 1 - t[0]) * (t[1] if dx else 1 - t[1]) * (t[2] if dz else 1 - t[2])
                    if w == 0:
                        continue
                    full = (cell[0] + dy) + NYc * ((cell[1] + dx) + NXc * (cell[2] + dz))
                    k = int(coarse.nodMapForward[full]) - 1
                    assert k >= 0, 'interpolation from a node outside the coarse mesh'
                    for d in range(3):
                        rows.append(3 * j + d)
                        cols.append(3 * k + d)
                        vals.append(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * fine.numNodes, 3 * coarse.numNodes))


def _super_model(shape, opt_case, mask=None):
    return M.build_cuboid_model(*shape, opt_case=opt_case, super_element=True, voxels=mask)


def _super_modulus(model, seed):
    Ne = model.levels[0].numElements
    x = np.random.default_rng(seed).random((64, Ne))
    return 1e-6 + x ** 3 * (1 - 1e-6)


@pytest.fixture(scope='module', params=['cantilever', 'mbb'])
def super_setup(request):
    case = 1 if request.param == 'cantilever' else 2
    m = _super_model((8, 16, 8), case)
    E = _super_modulus(m, 21 + case)
    sv = S.Solver(m)
    sv.assemble(E)
    return m, E, sv


# ---------------------------------------------------------------------------------------------
# super-element K*u (TOP_GPU.m:721-741)
# ---------------------------------------------------------------------------------------------

def test_super_kbyu_equals_direct_assembly(super_setup):
    m, E, sv = super_setup
    assert m.eleWithFixedDOFs.size > 0
    K = direct_global_K_super(m, E)
    u = np.random.default_rng(5).standard_normal(K.shape[0])
    y = sv.kbyu(u, 1)
    ref = K @ u
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_super_kbyu_boundary_override_is_needed(super_setup):
    """The check above can tell the boundary-element override apart: without it (K_mat for every
    element) the fixed rows are not identity rows and the result differs."""
    m, E, sv = super_setup
    lv = m.levels[0]
    Is = m.I_mat.reshape(24, 24, 64, order='F')
    plain = _assemble(lv.eNodMat, np.einsum('ijq,qe->ije', Is, E), 3 * lv.numNodes)
    u = np.random.default_rng(6).standard_normal(plain.shape[0])
    assert np.abs(plain @ u - sv.kbyu(u, 1)).max() > 1e-3 * np.abs(sv.kbyu(u, 1)).max()


# ---------------------------------------------------------------------------------------------
# Galerkin level 2 (standard and super-element): assembled K_2 = P^T K_1 P
# ---------------------------------------------------------------------------------------------

def _assembled_level(model, ii):
    lv = model.levels[ii - 1]
    return _assemble(lv.eNodMat, lv.Ks, 3 * lv.numNodes)


@pytest.mark.parametrize('mode', ['standard-box', 'standard-holes', 'super-cantilever', 'super-mbb'])
def test_level2_galerkin_is_global_ptkp(mode):
    if mode.startswith('super'):
        m = _super_model((8, 16, 8), 1 if mode.endswith('cantilever') else 2)
        E = _super_modulus(m, 31)
        K1 = direct_global_K_super(m, E)
    else:
        mask = holey_mask(12, 20, 12, seed=5) if mode.endswith('holes') else None
        m = M.build_cuboid_model(*(mask.shape if mask is not None else (8, 16, 8)), opt_case=1, voxels=mask)
        E = random_modulus(m.levels[0].numElements, seed=32)
        K1 = direct_global_K(m, E)
    sv = S.Solver(m)
    sv.assemble(E, fast=False)
    assert m.levels[1].spanWidth == 4
    P = geometric_prolongation(m, 2)
    ref = (P.T @ K1 @ P).toarray()
    got = _assembled_level(m, 2).toarray()
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_geometric_prolongation_is_the_oracle_interpolation():
    """The prolongation used above and the oracle's transfer operator agree (so the Galerkin pin
    and the transfer pins refer to the same P)."""
    m = M.build_cuboid_model(12, 20, 12, opt_case=1, voxels=holey_mask(12, 20, 12, seed=5))
    sv = S.Solver(m)
    rng = np.random.default_rng(8)
    for ii in range(2, m.numLevels + 1):
        P = geometric_prolongation(m, ii)
        x = rng.standard_normal(3 * m.levels[ii - 1].numNodes)
        assert np.abs(P @ x - sv.interpolate(x, ii)).max() < 1e-13 * np.abs(x).max()


# ---------------------------------------------------------------------------------------------
# super-element energies (TOP_GPU.m:2622-2671)
# ---------------------------------------------------------------------------------------------

def test_super_energies_identity(super_setup):
    m, E, sv = super_setup
    U = np.random.default_rng(9).standard_normal(3 * m.levels[0].numNodes)
    ce = T.unit_compliance_super(m.levels[0], U, sv.K_mat)
    dq = T.pre_compute_der(m.levels[0], U, m.I_mat)
    assert dq.shape == E.shape
    assert np.abs(np.sum(E * dq, axis=0) - ce).max() < 1e-12 * np.abs(ce).max()
    # each sub-voxel energy is a quadratic form of I_q alone: u_e' I_q u_e from the element DOFs
    Is = m.I_mat.reshape(24, 24, 64, order='F')
    dofs = _element_dofs(m.levels[0].eNodMat)
    for e, q in ((0, 0), (7, 33), (m.levels[0].numElements - 1, 63)):
        ue = U[dofs[e]]
        assert abs(dq[q, e] - ue @ Is[:, :, q] @ ue) < 1e-13 * abs(ue @ Is[:, :, q] @ ue) + 1e-300


def test_super_sensitivity_central_differences():
    """dc/dx_(q,e) = -p (E0 - Emin) x^(p-1) u_e' I_q u_e (the pre_compute_der branch of TOP_GPU.m:189-198)
    against central differences of c = F' K(x)^-1 F with K assembled and solved directly."""
    m = _super_model((8, 8, 8), 1)
    Ne = m.levels[0].numElements
    rng = np.random.default_rng(14)
    x = 0.3 + 0.4 * rng.random((64, Ne))
    Emin, p = 1e-6, 3.0
    F = m.F[:, 0]

    def comp(xx):
        K = direct_global_K_super(m, Emin + xx ** p * (1 - Emin))
        u = spla.spsolve(K.tocsc(), F)
        return float(F @ u), u

    c0, U = comp(x)
    dq = T.pre_compute_der(m.levels[0], U, m.I_mat)
    dc = -p * (1 - Emin) * x ** (p - 1) * dq
    h = 1e-4
    for q, e in ((0, 3), (5, Ne // 2), (63, Ne - 9), (21, 100)):
        xp, xm = x.copy(), x.copy()
        xp[q, e] += h
        xm[q, e] -= h
        fd = (comp(xp)[0] - comp(xm)[0]) / (2 * h)
        assert abs(fd - dc[q, e]) < 1e-5 * abs(dc[q, e]), (q, e, fd, dc[q, e])


# ---------------------------------------------------------------------------------------------
# convolution density filter and gen_index_mapping (TOP_GPU.m:112-121, 220-227, 262-265, 2673-2697)
# ---------------------------------------------------------------------------------------------

def test_gen_index_mapping_brute_force():
    X, Y, Z, n = 3, 2, 2, 4
    o2n, n2o = T.gen_index_mapping(X, Y, Z, n)
    nY, nX = n * Y, n * X
    e = 0
    for ez in range(Z):
        for ex in range(X):
            for ey in range(Y):                       # elements column-major over (nely, nelx, nelz)
                q = 0
                for lz in range(n):
                    for lx in range(n):
                        for ly in range(n):           # sub-voxels q = ly + n (lx + n lz)
                            gy, gx, gz = n * ey + ly, n * ex + lx, n * ez + lz
                            old = gy + nY * (gx + nX * gz)
                            assert o2n[old] == n ** 3 * e + q + 1
                            assert n2o[n ** 3 * e + q] == old + 1
                            q += 1
                e += 1
    assert np.array_equal(np.sort(o2n), np.arange(1, o2n.size + 1))


def _shift_correlate(A, h):
    """out[i] = sum_d h[d] A[i + d] with zeros outside A (correlation), by shift-and-add."""
    c = (h.shape[0] - 1) // 2
    out = np.zeros_like(A)
    P = np.pad(A, c)
    ny, nx, nz = A.shape
    for a in range(h.shape[0]):
        for b in range(h.shape[1]):
            for k in range(h.shape[2]):
                if h[a, b, k] != 0:
                    out += h[a, b, k] * P[a:a + ny, b:b + nx, k:k + nz]
    return out


def test_convolution_filter_brute_force_and_adjoint():
    nelx, nely, nelz, rMin = 3, 2, 2, math.sqrt(3)
    cf = T.ConvolutionFilter(nelx, nely, nelz, rMin)
    r = 4 * rMin
    c = math.ceil(r)
    d = np.arange(-c + 1, c)
    dy, dx, dz = np.meshgrid(d, d, d, indexing='ij')
    h = np.maximum(0.0, r - np.sqrt(dy ** 2 + dx ** 2 + dz ** 2))
    assert np.abs(cf.h - h).max() == 0                # the cone is symmetric in its three axes
    Ne = nelx * nely * nelz
    rng = np.random.default_rng(4)
    xv = rng.random((64, Ne))
    # element-major [64 x Ne] -> sub-voxel grid, by the definition of the sub-voxel order
    grid = np.zeros((4 * nely, 4 * nelx, 4 * nelz))
    for e in range(Ne):
        ey, ex, ez = e % nely, (e // nely) % nelx, e // (nely * nelx)
        for q in range(64):
            ly, lx, lz = q % 4, (q // 4) % 4, q // 16
            grid[4 * ey + ly, 4 * ex + lx, 4 * ez + lz] = xv[q, e]
    Hs = _shift_correlate(np.ones_like(grid), h)
    ref = _shift_correlate(grid, h) / Hs
    got = cf.forward(xv)
    for e in range(Ne):
        ey, ex, ez = e % nely, (e // nely) % nelx, e // (nely * nelx)
        for q in range(64):
            ly, lx, lz = q % 4, (q // 4) % 4, q // 16
            assert abs(got[q, e] - ref[4 * ey + ly, 4 * ex + lx, 4 * ez + lz]) < 1e-13
    # constants are preserved, and <F x, y> = <x, F^T y> between forward and adjoint
    assert np.abs(cf.forward(np.ones((64, Ne))) - 1).max() < 1e-13
    yv = rng.random((64, Ne))
    lhs, rhs = np.sum(cf.forward(xv) * yv), np.sum(xv * cf.adjoint(yv))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # a wrongly oriented map (y and x swapped inside an element) is caught by the brute force
    xs = xv.reshape(4, 4, 4, Ne).transpose(1, 0, 2, 3).reshape(64, Ne)
    assert np.abs(cf.forward(xs) - got).max() > 1e-3


# ---------------------------------------------------------------------------------------------
# Jacobi diagonals (TOP_GPU.m:1149-1239)
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize('mode', ['standard-box', 'standard-holes', 'super-cantilever', 'super-mbb'])
def test_jacobi_diagonals_are_assembled_diagonals(mode):
    if mode.startswith('super'):
        m = _super_model((8, 16, 8), 1 if mode.endswith('cantilever') else 2)
        E = _super_modulus(m, 41)
        K1 = direct_global_K_super(m, E)
    else:
        mask = holey_mask(12, 20, 12, seed=6) if mode.endswith('holes') else None
        m = M.build_cuboid_model(*(mask.shape if mask is not None else (8, 16, 8)), opt_case=1, voxels=mask)
        E = random_modulus(m.levels[0].numElements, seed=42)
        K1 = direct_global_K(m, E)
    sv = S.Solver(m)
    sv.assemble(E)
    d1 = K1.diagonal()
    assert np.abs(m.levels[0].diagK - d1).max() < 1e-13 * np.abs(d1).max()
    assert np.all(m.levels[0].diagK[m.fixedDOFs.astype(bool)] == 1.0)
    for ii in range(2, m.numLevels):
        dk = _assembled_level(m, ii).diagonal()
        assert np.abs(m.levels[ii - 1].diagK - dk).max() < 1e-13 * np.abs(dk).max()
    # level 2's diagonal is also that of P^T K_1 P (no reordering quirk at span 4)
    P = geometric_prolongation(m, 2)
    d2 = (P.T @ K1 @ P).diagonal()
    assert np.abs(m.levels[1].diagK - d2).max() < 1e-12 * np.abs(d2).max()
