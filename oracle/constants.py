"""Constant operators of the TOP_GPU hot path (oracle; test infrastructure only).

Every table here is computed from its mathematical definition rather than transcribed;
tests/test_oracle_constants.py pins each one against the reference's own tables
(golden vectors extracted as data by tests/golden/make_ref_tables.py) and against
closed-form properties.

Local conventions (TOP_GPU.m:2027-2040, 2083-2090):
  * a hexahedral element's 8 vertices, in array indices (row = y, col = x, page = z) with
    the element's base voxel at (iy, ix, iz):
        n1=(iy+1,ix,iz) n2=(iy+1,ix+1,iz) n3=(iy,ix+1,iz) n4=(iy,ix,iz), n5..n8 = same at iz+1;
  * physical y points towards decreasing row index, so in natural coordinates the
    vertices carry the signs sx=[-1 1 1 -1 -1 1 1 -1], sy=[-1 -1 1 1 -1 -1 1 1],
    sz=[-1 -1 -1 -1 1 1 1 1] (TOP_GPU.m:2545-2547);
  * element DOFs are node-interleaved: dof 3(n-1)+c, c in {x,y,z}.
"""
import numpy as np

# vertex signs of the 8-node hexahedron in natural coordinates (TOP_GPU.m:2545-2547)
SX = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=np.float64)
SY = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=np.float64)
SZ = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=np.float64)

# element row offsets (in nodes) of the eNodMat vertex pattern inside a local node block:
# vertex j sits at (dy, dx, dz) relative to the block's base voxel corner.
VERTEX_OFFSETS = np.array([
    [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 0],
    [1, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, 1]], dtype=np.int64)

# Node-interleaved DOF reordering of [x-dofs(8) y-dofs(8) z-dofs(8)] (TOP_GPU.m:1051).
RE_ORDERING = np.array([1 + n + 8 * c for n in range(8) for c in range(3)], dtype=np.int64)


def isotropic_D(E=1.0, nu=0.3):
    """Isotropic elasticity matrix with engineering shear strains (TOP_GPU.m:2578-2584)."""
    coef = E / ((1 + nu) * (1 - 2 * nu))
    D = np.zeros((6, 6))
    D[:3, :3] = nu
    np.fill_diagonal(D[:3, :3], 1 - nu)
    D[3, 3] = D[4, 4] = D[5, 5] = (1 - 2 * nu) / 2
    return coef * D


def hex8_BDB_at(xi, eta, zeta, E=1.0, nu=0.3):
    """B^T D B * detJ of a unit cube Hex8 at one natural point (TOP_GPU.m:2543-2588).

    Strain order (ex, ey, ez, gyz, gxz, gxy); detJ = 1/8, invJ = 2 I for a unit cell.
    """
    dN_dxi = 0.125 * SX * (1 + SY * eta) * (1 + SZ * zeta)
    dN_deta = 0.125 * SY * (1 + SX * xi) * (1 + SZ * zeta)
    dN_dzet = 0.125 * SZ * (1 + SX * xi) * (1 + SY * eta)
    dx, dy, dz = 2.0 * dN_dxi, 2.0 * dN_deta, 2.0 * dN_dzet
    B = np.zeros((6, 24))
    for a in range(8):
        iu, iv, iw = 3 * a, 3 * a + 1, 3 * a + 2
        B[0, iu] = dx[a]
        B[1, iv] = dy[a]
        B[2, iw] = dz[a]
        B[3, iv] = dz[a]; B[3, iw] = dy[a]
        B[4, iu] = dz[a]; B[4, iw] = dx[a]
        B[5, iu] = dy[a]; B[5, iv] = dx[a]
    return (B.T @ isotropic_D(E, nu) @ B) * 0.125


def element_stiffness(nu=0.3, cell_size=1.0):
    """Hex8 element stiffness Ke for E = 1 (FEA_VoxelBasedElementStiffnessMatrix, TOP_GPU.m:1465-1503).

    The reference writes Ke in closed form for a unit cube and scales it by cellSize_
    (TOP_GPU.m:1502).  The integrand B^T D B of a trilinear cube is of degree <= 2 in each
    coordinate, so 2x2x2 Gauss quadrature integrates it exactly; this restatement uses that.
    """
    g = 1.0 / np.sqrt(3.0)
    K = np.zeros((24, 24))
    for xi in (-g, g):
        for eta in (-g, g):
            for zeta in (-g, g):
                K += hex8_BDB_at(xi, eta, zeta, 1.0, nu)
    return K * cell_size


def restriction_interpolation_operator(span, in_dof=False):
    """Trilinear coarse-to-fine table RI [(span+1)^3 x 8]
    (Solving_Operator4MultiGridRestrictionAndInterpolation, TOP_GPU.m:2165-2328).

    Row i is the fine node (a, b, c) of the (span+1)^3 block in column-major (y, x, z)
    order; column j is coarse vertex j in eNodMat order; entry = trilinear weight.
    With in_dof the table is kron(RI, I3) (TOP_GPU.m:2324-2326).
    """
    n = span + 1
    a, b, c = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij')
    a, b, c = (v.reshape(-1, order='F') for v in (a, b, c))
    t = np.stack([a, b, c], axis=1) / span                     # fractional position in [0,1]
    RI = np.ones((n ** 3, 8))
    for j in range(8):
        for d in range(3):
            off = VERTEX_OFFSETS[j, d]
            RI[:, j] *= t[:, d] if off == 1 else (1.0 - t[:, d])
    if in_dof:
        RI = np.kron(RI, np.eye(3))
    return RI


def sub_element_node_table(span):
    """Local node indices (1-based, in the (span+1)^3 block) of the span^3 child elements
    (Solving_SubEleNodMat, TOP_GPU.m:1938-2017).

    Span 4 lists children in column-major (y, x, z) order, the same order as the columns of
    elementUpwardMap (TOP_GPU.m:1329-1336).  Span 2 lists each z-layer in hex-vertex order
    (0,0),(1,0),(1,1),(0,1) over (y, x) (TOP_GPU.m:1941-1948), so its rows 3/4 and 7/8 are
    swapped relative to elementUpwardMap; the Galerkin assembly pairs them as written.
    """
    n = span + 1
    if span == 4:
        order = [(ey, ex, ez) for ez in range(4) for ex in range(4) for ey in range(4)]
    elif span == 2:
        order = [(ey, ex, ez) for ez in range(2) for (ey, ex) in ((0, 0), (1, 0), (1, 1), (0, 1))]
    else:
        raise ValueError('Wrong input of span width!')
    rows = []
    for (ey, ex, ez) in order:
        rows.append([1 + (ey + dy) + n * (ex + dx) + n * n * (ez + dz) for dy, dx, dz in VERTEX_OFFSETS])
    return np.array(rows, dtype=np.int64)


def super_element_matrices(num=4, cell_size=1.0, nu=0.3):
    """I_mat [576 x num^3]: per-sub-voxel stiffness of a num^3 super-element
    (computeK_ij, TOP_GPU.m:2589-2608).

    Sub-voxel q = y + num*(x + num*z) (0-based) samples B^T D B at its centre with the
    y coordinate flipped (yy = span(num+1-y)), scaled by 8*cellSize/num^3.
    """
    span = -1.0 + (np.arange(1, num + 1) - 0.5) / num * 2.0
    K = np.zeros((576, num ** 3))
    eid = 0
    for z in range(num):
        for x in range(num):
            for y in range(num):
                K[:, eid] = hex8_BDB_at(span[x], span[num - 1 - y], span[z], 1.0, nu).reshape(-1, order='F')
                eid += 1
    K = K / (num ** 3)
    return K * 8 * cell_size


def pde_filter_kernel(filter_radius, ele_size=1.0):
    """Element matrix of the nodal Helmholtz filter, iKEF = r^2*KEF0 + KEF1
    (TopOpti_SetupPDEfilter_matrixFree, TOP_GPU.m:889-920).

    KEF0 = dShape' * dShape with natural-coordinate derivatives at the 8 Gauss points
    (unit weights); KEF1 = N' diag(w/8) N; r = filterRadius*eleSize/2/sqrt(3).
    """
    g = 1.0 / np.sqrt(3.0)
    s = SX * g
    t = SY * g
    p = SZ * g
    N = np.zeros((8, 8))
    dS = np.zeros((24, 8))
    for q in range(8):
        N[q] = 0.125 * (1 + SX * s[q]) * (1 + SY * t[q]) * (1 + SZ * p[q])
        dS[3 * q] = 0.125 * SX * (1 + SY * t[q]) * (1 + SZ * p[q])
        dS[3 * q + 1] = 0.125 * SY * (1 + SX * s[q]) * (1 + SZ * p[q])
        dS[3 * q + 2] = 0.125 * SZ * (1 + SX * s[q]) * (1 + SY * t[q])
    wgt = np.ones(8) * (1.0 / 8.0)
    KEF0 = dS.T @ dS
    KEF1 = N.T @ np.diag(wgt) @ N
    iRmin = (filter_radius * ele_size) / 2 / np.sqrt(3)
    return iRmin ** 2 * KEF0 + KEF1
