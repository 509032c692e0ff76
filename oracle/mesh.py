"""Voxel discretisation, multigrid mesh hierarchy and boundary conditions (oracle).

Literal, vectorised numpy restatement of the MATLAB host setup of TOP_GPU.m.  All index
arrays keep MATLAB conventions: int32, 1-based, column-major flattening, 0 = "none".
Test infrastructure only (see oracle/__init__.py).
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import constants as C


def matlab_round(x):
    """MATLAB round(): halves away from zero."""
    return math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)


@dataclass
class Level:
    """One level of meshHierarchy_ (Data_CartesianMeshStruct, TOP_GPU.m:2121-2163)."""
    resX: int = 0
    resY: int = 0
    resZ: int = 0
    spanWidth: int = 0
    eleSize: np.ndarray = None
    numElements: int = 0
    numNodes: int = 0
    numDOFs: int = 0
    eNodMat: np.ndarray = None               # [Ne x 8] int32, 1-based
    eleMapBack: np.ndarray = None            # [Ne] int32, 1-based full-grid element index
    eleMapForward: np.ndarray = None         # [nx*ny*nz] int32, 0 = void
    nodMapBack: np.ndarray = None            # [Nn] int32 (finest-grid node index on levels >= 2)
    nodMapForward: np.ndarray = None         # [(nx+1)(ny+1)(nz+1)] int32
    numNod2ElesVec: np.ndarray = None        # level 1 only
    nodesOnBoundary: np.ndarray = None       # int32, 1-based
    elementsOnBoundary: np.ndarray = None    # level 1 only
    elementUpwardMap: np.ndarray = None      # [Ne x span^3] int32 (levels >= 2)
    transferMat: np.ndarray = None           # [(span+1)^3 x Ne] int32 (levels >= 2)
    transferMatCoeffi: np.ndarray = None     # [intermediateNumNodes] float64
    solidNodeMapCoarser2Finer: np.ndarray = None
    intermediateNumNodes: int = 0
    multiGridOperatorRI: np.ndarray = None
    Ke: np.ndarray = None
    # solver state filled by the assembly
    Ks: np.ndarray = None                    # [24 x 24 x Ne] (level 1: Ke)
    diagK: np.ndarray = None
    eleModulus: np.ndarray = None


def _full_element_nodes(nx, ny, nz):
    """eNodVec + vertex offsets for every element of an nx*ny*nz box (TOP_GPU.m:2080-2088)."""
    nodenrs = np.arange(1, (nx + 1) * (ny + 1) * (nz + 1) + 1, dtype=np.int64).reshape(
        (ny + 1, nx + 1, nz + 1), order='F')
    eNodVec = (nodenrs[:-1, :-1, :-1] + 1).reshape(-1, order='F')
    tmp = np.array([0, ny + 1, ny, -1], dtype=np.int64)
    tmp = np.concatenate([tmp, (ny + 1) * (nx + 1) + tmp])
    return eNodVec[:, None] + tmp[None, :]                    # [nx*ny*nz x 8]


def _recover_half(half):
    """Common_RecoverHalfeNodMat (TOP_GPU.m:1436-1442): columns [3 4 7 8] -> all 8."""
    out = np.zeros((half.shape[0], 8), dtype=half.dtype)
    out[:, [2, 3, 6, 7]] = half
    out[:, [1, 0, 5, 4]] = half + 1
    return out


def _discretize(lv, eleMapBack):
    """Element/node numbering of one level from its solid-element list (TOP_GPU.m:2080-2099, 1388-1403).

    Returns (eNodMat compressed [Ne x 8] int32, full-grid eNodMatHalfTemp [nAll x 4]).
    """
    nx, ny, nz = lv.resX, lv.resY, lv.resZ
    full = _full_element_nodes(nx, ny, nz)
    half = full[:, [2, 3, 6, 7]]
    eNod = full[eleMapBack.astype(np.int64) - 1]
    nodMapBack = np.unique(eNod)
    lv.nodMapBack = nodMapBack.astype(np.int32)
    lv.numNodes = int(nodMapBack.size)
    lv.numDOFs = 3 * lv.numNodes
    fwd = np.zeros((nx + 1) * (ny + 1) * (nz + 1), dtype=np.int32)
    fwd[nodMapBack - 1] = np.arange(1, lv.numNodes + 1, dtype=np.int32)
    lv.nodMapForward = fwd
    eNodMat = fwd[eNod - 1].astype(np.int32)
    return eNodMat, half


def voxel_discretization(voxels, coarsest_resolution_control=20000):
    """FEA_VoxelBasedDiscretization (TOP_GPU.m:2019-2119).

    voxels: bool [nely, nelx, nelz].  Returns (finest Level, numLevels, eNodMatHalfTemp, padded voxels).
    """
    vox = np.asarray(voxels, dtype=bool)
    nely, nelx, nelz = vox.shape
    numVoxels = int(np.count_nonzero(vox))
    numLevels = 0
    while numVoxels >= coarsest_resolution_control:
        numLevels += 1
        numVoxels = matlab_round(numVoxels / 8)
    numLevels = max(3, numLevels)
    f = 2 ** numLevels
    adjX = math.ceil(nelx / f) * f
    adjY = math.ceil(nely / f) * f
    adjZ = math.ceil(nelz / f) * f
    numLevels += 1
    padded = np.zeros((adjY, adjX, adjZ), dtype=bool)
    padded[:nely, :nelx, :nelz] = vox

    lv = Level(resX=adjX, resY=adjY, resZ=adjZ)
    lv.eleSize = np.array([adjX, adjY, adjZ], dtype=np.float64) / np.array([adjX, adjY, adjZ])
    flat = padded.reshape(-1, order='F')
    lv.eleMapBack = (np.flatnonzero(flat) + 1).astype(np.int32)
    lv.numElements = int(lv.eleMapBack.size)
    lv.eleMapForward = np.zeros(adjX * adjY * adjZ, dtype=np.int32)
    lv.eleMapForward[lv.eleMapBack - 1] = np.arange(1, lv.numElements + 1, dtype=np.int32)
    eNodMat, half = _discretize(lv, lv.eleMapBack)
    # 6. boundary info (TOP_GPU.m:2101-2116)
    cnt = np.bincount(eNodMat.reshape(-1).astype(np.int64) - 1, minlength=lv.numNodes).astype(np.float64)
    lv.numNod2ElesVec = cnt
    lv.nodesOnBoundary = (np.flatnonzero(cnt < 8) + 1).astype(np.int32)
    onb = np.zeros(lv.numNodes, dtype=np.int32)
    onb[lv.nodesOnBoundary - 1] = 1
    tmp = onb[eNodMat.astype(np.int64) - 1].sum(axis=1)
    lv.elementsOnBoundary = (np.flatnonzero(tmp > 0) + 1).astype(np.int32)
    lv.eNodMat = eNodMat
    return lv, numLevels, half, padded


def build_mesh_hierarchy(finest, numLevels, half_finest, non_dyadic=1):
    """Solving_BuildingMeshHierarchy (TOP_GPU.m:1276-1434).

    Returns (list of Levels, numLevels after the non-dyadic merge, non_dyadic flag).
    """
    levels = [finest]
    resX1, resY1, resZ1 = finest.resX, finest.resY, finest.resZ
    nodeVolume = np.arange(1, (resX1 + 1) * (resY1 + 1) * (resZ1 + 1) + 1, dtype=np.int32).reshape(
        (resY1 + 1, resX1 + 1, resZ1 + 1), order='F')
    if non_dyadic == 1 and numLevels >= 4:
        numLevels -= 1
    else:
        non_dyadic = 0
    half = half_finest
    for ii in range(2, numLevels + 1):
        prev = levels[-1]
        span = 4 if (ii == 2 and non_dyadic == 1) else 2
        nx, ny, nz = prev.resX // span, prev.resY // span, prev.resZ // span
        lv = Level(resX=nx, resY=ny, resZ=nz, spanWidth=span)
        lv.eleSize = prev.eleSize * span
        # 3.1 raw per-coarse-element maps (TOP_GPU.m:1309-1340), vectorised over (gg, kk, jj)
        vol = prev.eleMapForward.reshape((span * ny, span * nx, span * nz), order='F')
        s6 = vol.reshape((span, ny, span, nx, span, nz), order='F').transpose(1, 3, 5, 0, 2, 4)
        up_all = s6.reshape((nx * ny * nz, span ** 3), order='F').astype(np.int32)
        tmpIdx = np.arange(1, span ** 3 * nx * ny * nz + 1, dtype=np.int64).reshape(
            (span * ny, span * nx, span * nz), order='F')
        t6 = tmpIdx.reshape((span, ny, span, nx, span, nz), order='F').transpose(1, 3, 5, 0, 2, 4)
        upTemp_all = t6.reshape((nx * ny * nz, span ** 3), order='F')
        # fine-grid node numbers of each coarse cell's (span+1)^3 block
        gy, gx, gz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing='ij')
        gy, gx, gz = (v.reshape(-1, order='F') for v in (gy, gx, gz))
        a, b, c = np.meshgrid(np.arange(span + 1), np.arange(span + 1), np.arange(span + 1), indexing='ij')
        a, b, c = (v.reshape(-1, order='F') for v in (a, b, c))
        NY, NX = span * ny + 1, span * nx + 1
        fy = span * gy[None, :] + a[:, None]
        fx = span * gx[None, :] + b[:, None]
        fz = span * gz[None, :] + c[:, None]
        transferTemp_all = 1 + fy + NY * (fx + NX * fz)          # [(span+1)^3 x nx*ny*nz]
        # 3.2 (TOP_GPU.m:1354-1371)
        unempty = np.flatnonzero(up_all.astype(np.int64).sum(axis=1) > 0)
        upTemp = upTemp_all[unempty]
        elemsIncVoid = upTemp.reshape(-1, order='F')
        nodesIncVoid = np.unique(_recover_half(half[elemsIncVoid - 1]))
        lv.intermediateNumNodes = int(nodesIncVoid.size)
        transferTemp = transferTemp_all[:, unempty]
        temp = np.zeros(NY * NX * (span * nz + 1), dtype=np.int32)
        temp[nodesIncVoid - 1] = np.arange(1, lv.intermediateNumNodes + 1, dtype=np.int32)
        lv.transferMat = temp[transferTemp - 1]
        coeff = np.zeros(lv.intermediateNumNodes)
        for kk in range((span + 1) ** 3):
            coeff[lv.transferMat[kk].astype(np.int64) - 1] += 1
        lv.transferMatCoeffi = coeff
        nodesLast = np.unique(_recover_half(half[prev.eleMapBack.astype(np.int64) - 1]))
        common = np.intersect1d(nodesIncVoid, nodesLast)
        lv.solidNodeMapCoarser2Finer = (np.searchsorted(nodesIncVoid, common) + 1).astype(np.int32)
        # 3.3 solid coarse elements (TOP_GPU.m:1381-1386)
        lv.eleMapForward = np.zeros(nx * ny * nz, dtype=np.int32)
        lv.eleMapBack = (unempty + 1).astype(np.int32)
        lv.numElements = int(unempty.size)
        lv.eleMapForward[unempty] = np.arange(1, lv.numElements + 1, dtype=np.int32)
        lv.elementUpwardMap = up_all[unempty]
        # 4. discretize (TOP_GPU.m:1388-1407)
        eNodMat, half = _discretize(lv, lv.eleMapBack)
        kk = ii if non_dyadic == 1 else ii - 1
        sub = nodeVolume[::2 ** kk, ::2 ** kk, ::2 ** kk].reshape(-1, order='F')
        lv.nodMapBack = sub[lv.nodMapBack.astype(np.int64) - 1].astype(np.int32)
        # 5./6. (TOP_GPU.m:1410-1419)
        lv.multiGridOperatorRI = C.restriction_interpolation_operator(span)
        cnt = np.bincount(eNodMat.reshape(-1).astype(np.int64) - 1, minlength=lv.numNodes)
        lv.nodesOnBoundary = (np.flatnonzero(cnt < 8) + 1).astype(np.int32)
        lv.eNodMat = eNodMat
        levels.append(lv)
    return levels, numLevels, non_dyadic


# ---------------------------------------------------------------------------------------
# Built-in boundary-condition cases (CreateVoxelFEAmodel, TOP_GPU.m:497-535)
# ---------------------------------------------------------------------------------------

def builtin_case(finest, nely, nelx, nelz, opt_case):
    """Fixations and loads of the built-in cuboid cases, mapped to solid node numbers.

    Returns (fixingCond [nFix x 4] float64, loadingCond [nLoad x 4] float64).
    """
    shape = (finest.resY + 1, finest.resX + 1, finest.resZ + 1)

    def nodes(mask):
        full = np.flatnonzero(mask.reshape(-1, order='F')) + 1
        return finest.nodMapForward[full - 1].astype(np.float64)

    vol = np.zeros(shape, dtype=bool)
    if opt_case == 1:          # cantilever (TOP_GPU.m:507-516)
        vol[0:nely + 1, 0, 0:nelz + 1] = True
        fix = nodes(vol)
        fixing = np.column_stack([fix, np.ones((fix.size, 3))])
        vol = np.zeros(shape, dtype=bool)
        vol[0:nely + 1, nelx, 0:matlab_round(nelz / 6 + 1)] = True
        ld = nodes(vol)
        loading = np.column_stack([ld, np.tile(np.array([0.0, 0.0, -1.0]) / ld.size, (ld.size, 1))])
    elif opt_case == 2:        # MBB (TOP_GPU.m:517-529)
        vol[0, 0, 0] = vol[0, nelx, 0] = vol[nely, 0, 0] = vol[nely, nelx, 0] = True
        fix = nodes(vol)
        fixing = np.column_stack([fix, np.ones((fix.size, 3))])
        vol = np.zeros(shape, dtype=bool)
        vol[nely // 2, nelx // 2, nelz] = True
        ld = nodes(vol)
        loading = np.column_stack([ld, np.tile(np.array([0.0, 0.0, -1.0]), (ld.size, 1))])
    else:
        raise ValueError('unknown optCase %r' % opt_case)
    return fixing, loading


def apply_boundary_condition(finest, fixing, loadings):
    """FEA_ApplyBoundaryCondition (TOP_GPU.m:1505-1546).

    Returns (fixingCond after load removal, list of loads, F [3Nn x nLoad] dense,
    fixedDOFs bool [3Nn]).
    """
    loads = []
    for iLoad in loadings:
        keep = ~np.isin(fixing[:, 0], iLoad[:, 0])
        vals = fixing[keep, 0]
        # setdiff returns the sorted unique values with first-occurrence indices
        _, first = np.unique(vals, return_index=True)
        fixing = fixing[keep][first]
        _, first = np.unique(fixing[:, 0], return_index=True)
        fixing = fixing[first]
        _, first = np.unique(iLoad[:, 0], return_index=True)
        loads.append(iLoad[first])
    Nn = finest.numNodes
    F = np.zeros((3 * Nn, len(loads)))
    for ii, iLoad in enumerate(loads):
        arr = np.zeros((Nn, 3))
        arr[iLoad[:, 0].astype(np.int64) - 1, :] = iLoad[:, 1:]
        F[:, ii] = arr.reshape(-1)                      # reshape(iFarr', numDOFs, 1): interleaved
    nodes = fixing[:, 0].astype(np.int64)
    dofs = (3 * nodes[:, None] - np.array([2, 1, 0])[None, :]).reshape(-1)
    state = fixing[:, 1:].reshape(-1)
    fixed = np.zeros(3 * Nn, dtype=bool)
    fixed[dofs[state == 1] - 1] = True
    return fixing, loads, F, fixed


def _apply_bc_on_element(free, fixed, fixed_dofs_local, n_adj):
    """Solving_ApplyBConEleStiffMat_B (TOP_GPU.m:1729-1741); free may be [24x24] or [24x24xq]."""
    d = np.asarray(fixed_dofs_local, dtype=np.int64) - 1
    free[d, ...] = 0.0
    free[:, d, ...] = 0.0
    for k in d:
        fixed[k, k] = 1.0 / n_adj


def setup_ke_with_fixed_dofs(finest, fixing, Ke_or_Imat, modulus=1.0, super_element=False):
    """Solving_SetupKeWithFixedDOFs (TOP_GPU.m:1548-1727).

    Returns (mapUniqueKes [Ne] int32, uniqueKesFree, uniqueKesFixed [576 x nU],
    eleWithFixedDOFs [nU] int32 1-based).  uniqueKesFree is [576 x nU] in standard mode and
    [576 x 64 x nU] in super-element mode (columns of I_mat with the BC applied).
    """
    Ne, Nn = finest.numElements, finest.numNodes
    eNod = finest.eNodMat.astype(np.int64)
    onB = np.zeros(Ne, dtype=bool)
    onB[finest.elementsOnBoundary - 1] = True
    bpos = np.zeros(Nn, dtype=np.int64)
    bpos[finest.nodesOnBoundary - 1] = np.arange(1, finest.nodesOnBoundary.size + 1)
    lists = [[] for _ in range(finest.nodesOnBoundary.size)]
    for e in np.flatnonzero(onB):
        for n in eNod[e]:
            k = bpos[n - 1]
            if k:
                lists[k - 1].append(e + 1)
    common = np.intersect1d(finest.nodesOnBoundary, fixing[:, 0].astype(np.int64))
    fixedNodeIndices = np.searchsorted(finest.nodesOnBoundary, common)
    fixedNodesStruct = [lists[i] for i in fixedNodeIndices]
    allE = np.unique(np.array([e for l in fixedNodesStruct for e in l], dtype=np.int64))
    nU = allE.size
    mapU = np.zeros(Ne, dtype=np.int32)
    mapU[allE - 1] = np.arange(1, nU + 1, dtype=np.int32)
    if super_element:
        Imat = np.asarray(Ke_or_Imat)
        free = np.repeat(Imat.reshape(24, 24, -1, order='F')[:, :, :, None], nU, axis=3)
    else:
        Ke = np.asarray(Ke_or_Imat)
        free = np.repeat(Ke[:, :, None], nU, axis=2)
    fixedK = np.zeros((24, 24, nU))
    for ii, arr in enumerate(fixedNodesStruct):
        iNode = int(fixing[ii, 0])             # pairing as written (TOP_GPU.m:1597-1598)
        state = fixing[ii, 1:]
        nAdj = len(arr)
        for e in arr:
            loc = np.flatnonzero(eNod[e - 1] == iNode)
            if loc.size != 1:
                raise ValueError('fixation/boundary-node pairing mismatch (TOP_GPU.m:1597-1604)')
            dofs = 3 * (loc[0] + 1) - np.array([2, 1, 0])
            dofs = dofs[state == 1]
            u = mapU[e - 1] - 1
            if super_element:
                f = free[:, :, :, u]
                _apply_bc_on_element(f, fixedK[:, :, u], dofs, nAdj)
            else:
                _apply_bc_on_element(free[:, :, u], fixedK[:, :, u], dofs, nAdj)
    fixedK = fixedK * modulus
    if super_element:
        freeCols = free.reshape(576, -1, nU, order='F')
    else:
        freeCols = free.reshape(576, nU, order='F')
    fixedCols = fixedK.reshape(576, nU, order='F')
    eleWithFixed = (np.flatnonzero(mapU > 0) + 1).astype(np.int32)
    return mapU, freeCols, fixedCols, eleWithFixed


@dataclass
class Model:
    """Everything CreateVoxelFEAmodel + FEA_ApplyBoundaryCondition + FEA_SetupVoxelBased produce."""
    levels: List[Level]
    numLevels: int
    nonDyadic: int
    nelx: int
    nely: int
    nelz: int
    fixing: np.ndarray
    loads: list
    F: np.ndarray
    fixedDOFs: np.ndarray
    mapUniqueKes: np.ndarray
    uniqueKesFree: np.ndarray
    uniqueKesFixed: np.ndarray
    eleWithFixedDOFs: np.ndarray
    cellSize: float
    nu: float
    super_element: bool = False
    I_mat: Optional[np.ndarray] = None


def adapt_bc_external(srcBC, nely, nelx, nelz, resX, resY, resZ):
    """AdaptBCExternalMdl (TOP_GPU.m:1743-1793): rows [node v1 v2 v3] of the unpadded
    (nely+1)(nelx+1)(nelz+1) node grid -> the padded (resY+1)(resX+1)(resZ+1) grid, ascending."""
    srcBC = srcBC[np.argsort(srcBC[:, 0], kind='stable')]
    vol = np.zeros((nely + 1) * (nelx + 1) * (nelz + 1))
    out = np.zeros_like(srcBC, dtype=np.float64)
    vol[srcBC[:, 0].astype(np.int64) - 1] = 1
    padded = np.zeros((resY + 1, resX + 1, resZ + 1))
    padded[:nely + 1, :nelx + 1, :nelz + 1] = vol.reshape((nely + 1, nelx + 1, nelz + 1), order='F')
    newNodes = np.flatnonzero(padded.reshape(-1, order='F'))
    out[:, 0] = newNodes + 1
    for c in range(1, 4):
        vol = np.zeros((nely + 1) * (nelx + 1) * (nelz + 1))
        vol[srcBC[:, 0].astype(np.int64) - 1] = srcBC[:, c]
        padded = np.zeros((resY + 1, resX + 1, resZ + 1))
        padded[:nely + 1, :nelx + 1, :nelz + 1] = vol.reshape((nely + 1, nelx + 1, nelz + 1), order='F')
        out[:, c] = padded.reshape(-1, order='F')[newNodes]
    return out


def build_file_model(nelx, nely, nelz, voxels, fixing, loads, nu=0.3, coarsest_resolution_control=20000,
                     non_dyadic=1, super_element=False):
    """CreateVoxelFEAmodel for a parsed .TopVoxel file (TOP_GPU.m:416-496): cellSize_ = 1 (TOP_GPU.m:67);
    fixations and loads are uniqued, sorted, moved to the padded grid and renumbered to solid nodes."""
    flat = np.zeros(nelx * nely * nelz, dtype=bool)
    flat[np.asarray(voxels, dtype=np.int64) - 1] = True
    vox = flat.reshape((nely, nelx, nelz), order='F')
    finest, numLevels, half, _ = voxel_discretization(vox, coarsest_resolution_control)
    levels, numLevels, nd = build_mesh_hierarchy(finest, numLevels, half, non_dyadic)

    def prep(a):
        a = np.asarray(a, dtype=np.float64)
        _, first = np.unique(a[:, 0], return_index=True)
        a = a[first]
        a = adapt_bc_external(a, nely, nelx, nelz, finest.resX, finest.resY, finest.resZ)
        a[:, 0] = finest.nodMapForward[a[:, 0].astype(np.int64) - 1]
        return a

    fixing = prep(fixing)
    # one [n x 4] block, or a list of them: the file's load cases (loadingCond_, TOP_GPU.m:447-461)
    blocks = loads if isinstance(loads, (list, tuple)) else [loads]
    fixing, ldl, F, fixed = apply_boundary_condition(finest, fixing, [prep(b) for b in blocks])
    finest.Ke = C.element_stiffness(nu, 1.0)
    I_mat = None
    if super_element:  # as for logical input (build_cuboid_model), with cellSize_ = 1
        I_mat = C.super_element_matrices(4, 1.0, nu)
        mapU, free, fixK, eleFix = setup_ke_with_fixed_dofs(finest, fixing, I_mat, 1.0, True)
    else:
        mapU, free, fixK, eleFix = setup_ke_with_fixed_dofs(finest, fixing, finest.Ke, 1.0, False)
    return Model(levels=levels, numLevels=numLevels, nonDyadic=nd, nelx=nelx, nely=nely, nelz=nelz,
                 fixing=fixing, loads=ldl, F=F, fixedDOFs=fixed, mapUniqueKes=mapU, uniqueKesFree=free,
                 uniqueKesFixed=fixK, eleWithFixedDOFs=eleFix, cellSize=1.0, nu=nu,
                 super_element=super_element, I_mat=I_mat)


def build_cuboid_model(nely, nelx, nelz, opt_case=1, cell_size=None, nu=0.3, super_element=False,
                       non_dyadic=1, coarsest_resolution_control=20000, voxels=None):
    """CreateVoxelFEAmodel (logical input) -> FEA_ApplyBoundaryCondition -> FEA_SetupVoxelBased
    (TOP_GPU.m:95-103, 497-535, 1444-1463).  cellSize_ = 1/nelx for logical input (TOP_GPU.m:61, 392)."""
    if voxels is None:
        voxels = np.ones((nely, nelx, nelz), dtype=bool)
    if cell_size is None:
        cell_size = 1.0 / nelx
    finest, numLevels, half, _ = voxel_discretization(voxels, coarsest_resolution_control)
    levels, numLevels, nd = build_mesh_hierarchy(finest, numLevels, half, non_dyadic)
    fixing, loading = builtin_case(finest, nely, nelx, nelz, opt_case)
    fixing, loads, F, fixed = apply_boundary_condition(finest, fixing, [loading])
    finest.Ke = C.element_stiffness(nu, cell_size)
    I_mat = None
    if super_element:
        I_mat = C.super_element_matrices(4, cell_size, nu)
        mapU, free, fixK, eleFix = setup_ke_with_fixed_dofs(finest, fixing, I_mat, 1.0, True)
    else:
        mapU, free, fixK, eleFix = setup_ke_with_fixed_dofs(finest, fixing, finest.Ke, 1.0, False)
    return Model(levels=levels, numLevels=numLevels, nonDyadic=nd, nelx=nelx, nely=nely, nelz=nelz,
                 fixing=fixing, loads=loads, F=F, fixedDOFs=fixed, mapUniqueKes=mapU,
                 uniqueKesFree=free, uniqueKesFixed=fixK, eleWithFixedDOFs=eleFix,
                 cellSize=cell_size, nu=nu, super_element=super_element, I_mat=I_mat)
