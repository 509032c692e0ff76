"""Pin the oracle's constant operators against the reference's own tables (golden vectors
extracted as data by tests/golden/make_ref_tables.py) and against closed-form properties."""
import numpy as np

from oracle import constants as C


def test_ke_matches_reference_closed_form(golden):
    # FEA_VoxelBasedElementStiffnessMatrix (TOP_GPU.m:1465-1503) vs 2x2x2 Gauss quadrature
    Ke = C.element_stiffness(0.3, 1.0)
    ref = np.array(golden['Ke_nu0.3_cell1'])
    assert np.abs(Ke - ref).max() < 1e-15


def test_ke_properties():
    Ke = C.element_stiffness(0.3, 1.0)
    assert np.allclose(Ke, Ke.T, atol=1e-16)
    w = np.linalg.eigvalsh(Ke)
    assert np.sum(np.abs(w) < 1e-12) == 6            # 3 translations + 3 rotations
    assert w.min() > -1e-14
    # rigid body modes in the null space (node coordinates from the vertex signs)
    X = np.stack([C.SX, C.SY, C.SZ], axis=1) / 2
    for t in np.eye(3):
        u = np.tile(t, 8)
        assert np.abs(Ke @ u).max() < 1e-14
    for axis in np.eye(3):
        u = np.cross(axis[None, :], X).reshape(-1)
        assert np.abs(Ke @ u).max() < 1e-14
    # uniform strain energy: u = G x gives 1/2 u'Ke u = 1/2 eps:D:eps * volume (patch test)
    G = np.array([[0.1, 0.02, 0.0], [0.02, -0.05, 0.03], [0.0, 0.03, 0.07]])
    u = (X @ G.T).reshape(-1)
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], 2 * G[1, 2], 2 * G[0, 2], 2 * G[0, 1]])
    assert abs(u @ Ke @ u - eps @ C.isotropic_D() @ eps) < 1e-14
    # linear in cell size
    assert np.abs(C.element_stiffness(0.3, 0.25) - 0.25 * Ke).max() < 1e-16


def test_restriction_interpolation_tables(golden):
    for s in (2, 4):
        RI = C.restriction_interpolation_operator(s)
        assert np.array_equal(RI, np.array(golden['RI_span%d' % s]))
        assert np.allclose(RI.sum(axis=1), 1.0, atol=0)          # partition of unity
        # reproduces linear fields: coarse vertex values f(v) -> f at fine nodes
        n = s + 1
        a, b, c = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij')
        pts = np.stack([v.reshape(-1, order='F') for v in (a, b, c)], axis=1) / s
        f = lambda p: 1.5 + 2 * p[:, 0] - 3 * p[:, 1] + 0.5 * p[:, 2]
        assert np.abs(RI @ f(C.VERTEX_OFFSETS.astype(float)) - f(pts)).max() < 1e-14
    assert np.array_equal(C.restriction_interpolation_operator(2, in_dof=True),
                          np.kron(np.array(golden['RI_span2']), np.eye(3)))


def test_sub_element_tables(golden):
    for s in (2, 4):
        assert np.array_equal(C.sub_element_node_table(s), np.array(golden['SubEleNodMat_span%d' % s]))
    assert np.array_equal(C.RE_ORDERING, np.array(golden['reOrdering']))


def test_super_element_matrices():
    I = C.super_element_matrices(4, 1.0)
    assert I.shape == (576, 64)
    Ks = I.reshape(24, 24, 64, order='F')
    assert np.allclose(Ks, Ks.transpose(1, 0, 2), atol=1e-16)
    # the midpoint rule is not exact: sum_q I_q differs from Ke by ~3% (TOP_GPU.m:2589-2608)
    d = np.abs(Ks.sum(axis=2) - C.element_stiffness()).max() / np.abs(C.element_stiffness()).max()
    assert 0.01 < d < 0.06
    # each I_q annihilates rigid translations
    for t in np.eye(3):
        assert np.abs(np.einsum('ijq,j->iq', Ks, np.tile(t, 8))).max() < 1e-14
    # sub-voxel q = y + 4(x + 4z) with y flipped: q=0 samples eta = +0.75
    assert np.allclose(I[:, 0], C.hex8_BDB_at(-0.75, 0.75, -0.75).reshape(-1, order='F') * 8 / 64)


def test_pde_filter_kernel():
    K = C.pde_filter_kernel(np.sqrt(3))
    assert np.allclose(K, K.T)
    # constants: the Laplacian part annihilates them, the mass part integrates to the volume 1
    K0 = C.pde_filter_kernel(0.0)
    assert abs(K0.sum() - 1.0) < 1e-15
    Kr = C.pde_filter_kernel(2 * np.sqrt(3)) - K0          # r = 1 -> pure KEF0
    assert np.abs(Kr @ np.ones(8)).max() < 1e-15
