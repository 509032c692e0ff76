"""The .TopVoxel model format (CreateVoxelFEAmodel file branch, TOP_GPU.m:416-496, 1743-1813).

Layout (whitespace-separated tokens, as read by the reference's fscanf calls):

    <4 header words>
    Version: 1.0
    Resolution: nelx nely nelz
    <2 words> densityFlag              (1: each solid voxel line carries a density)
    <2 words> numSolidVoxels
    voxel [density]  ...               (1-based column-major indices of the nely x nelx x nelz box)
    Passive elements: n  e1 e2 ...
    Fixations: n   node fx fy fz ...   (nodes of the (nely+1)(nelx+1)(nelz+1) grid, flags 0/1)
    Loads: n       node Fx Fy Fz ...
    <2 words> numAdditionalLoads
      Loads: n idx  node Fx Fy Fz ...  (repeated)
"""
import numpy as np

from .model import Model


class TopVoxelError(ValueError):
    pass


def read_topvoxel(path):
    """Parse a .TopVoxel file.  Returns dict(nelx, nely, nelz, voxels [1-based], densities or None,
    passive [1-based], fixing [n x 4], loads [list of [n x 4]])."""
    toks = open(path).read().split()
    pos = [0]

    def word():
        if pos[0] >= len(toks):
            raise TopVoxelError('Incompatible Mesh Data Format!')
        t = toks[pos[0]]
        pos[0] += 1
        return t

    def num(kind=float):
        return kind(float(word()))

    def expect(w):
        if word() != w:
            raise TopVoxelError('Incompatible Mesh Data Format!')

    for _ in range(4):
        word()
    word()                                  # "Version:"
    version = num()
    if version != 1.0:
        raise TopVoxelError('Unsupported Data!')
    expect('Resolution:')
    nelx, nely, nelz = num(int), num(int), num(int)
    word(), word()
    has_density = num(int)
    word(), word()
    nsolid = num(int)
    if has_density:
        vals = np.array([float(word()) for _ in range(2 * nsolid)]).reshape(nsolid, 2)
        voxels, dens = vals[:, 0].astype(np.int64), vals[:, 1]
    else:
        voxels, dens = np.array([num(int) for _ in range(nsolid)], dtype=np.int64), None
    expect('Passive')
    expect('elements:')
    npas = num(int)
    passive = np.array([num(int) for _ in range(npas)], dtype=np.int64)
    expect('Fixations:')
    nfix = num(int)
    fixing = np.array([float(word()) for _ in range(4 * nfix)]).reshape(nfix, 4) if nfix else np.zeros((0, 4))
    expect('Loads:')
    nld = num(int)
    loads = [np.array([float(word()) for _ in range(4 * nld)]).reshape(nld, 4) if nld else np.zeros((0, 4))]
    word(), word()
    nadd = num(int)
    for _ in range(nadd):
        expect('Loads:')
        n = num(int)
        idx = num(int)
        arr = np.array([float(word()) for _ in range(4 * n)]).reshape(n, 4)
        while len(loads) < idx:
            loads.append(np.zeros((0, 4)))
        loads[idx - 1] = arr
    return dict(nelx=nelx, nely=nely, nelz=nelz, voxels=voxels, densities=dens, passive=passive, fixing=fixing,
                loads=loads)


def write_topvoxel(path, voxels_mask, fixing, loads, passive=(), densities=None, additional_loads=()):
    """Write a .TopVoxel file.  voxels_mask: bool [nely, nelx, nelz]; fixing/loads in node indices of the
    unpadded (nely+1)(nelx+1)(nelz+1) grid (1-based); additional_loads: further load cases, [n x 4] each,
    written as the reference reads them (TOP_GPU.m:451-460: 'Additional Loads: k', then per case
    'Loads: n index')."""
    nely, nelx, nelz = voxels_mask.shape
    solid = np.flatnonzero(np.asarray(voxels_mask).reshape(-1, order='F')) + 1
    with open(path, 'w') as f:
        f.write('TopVoxel voxel model file\nVersion: 1.0\n')
        f.write('Resolution: %d %d %d\n' % (nelx, nely, nelz))
        f.write('Density values: %d\n' % (1 if densities is not None else 0))
        f.write('Solid voxels: %d\n' % solid.size)
        if densities is not None:
            for v, d in zip(solid, densities):
                f.write('%d %.16e\n' % (v, d))
        else:
            f.write('\n'.join(str(v) for v in solid) + '\n')
        f.write('Passive elements: %d\n' % len(passive))
        if len(passive):
            f.write(' '.join(str(int(p)) for p in passive) + '\n')
        f.write('Fixations: %d\n' % len(fixing))
        for r in fixing:
            f.write('%d %d %d %d\n' % tuple(int(v) for v in r))
        f.write('Loads: %d\n' % len(loads))
        for r in loads:
            f.write('%d %.16e %.16e %.16e\n' % (int(r[0]), r[1], r[2], r[3]))
        f.write('Additional Loads: %d\n' % len(additional_loads))
        for k, block in enumerate(additional_loads):
            f.write('Loads: %d %d\n' % (len(block), k + 2))
            for r in block:
                f.write('%d %.16e %.16e %.16e\n' % (int(r[0]), r[1], r[2], r[3]))


def _unique_sorted_rows(a):
    """[~,i] = unique(a(:,1)); a = a(i,:)  (first occurrence, ascending node)."""
    if len(a) == 0:
        return a
    _, first = np.unique(a[:, 0], return_index=True)
    return a[first]


def _adapt_nodes(nodes, nely, nelx, nelz, model):
    """AdaptBCExternalMdl (TOP_GPU.m:1743-1793): node index of the unpadded grid -> padded grid."""
    info = model.level_info(1)
    NYp, NXp = info['resY'] + 1, info['resX'] + 1
    n0 = nodes.astype(np.int64) - 1
    iy = n0 % (nely + 1)
    ix = (n0 // (nely + 1)) % (nelx + 1)
    iz = n0 // ((nely + 1) * (nelx + 1))
    return 1 + iy + NYp * (ix + NXp * iz)


def load_topvoxel_model(path, super_element=False, nu=0.3, precision='fp64', **kw):
    """CreateVoxelFEAmodel(fileName) + FEA_ApplyBoundaryCondition + FEA_SetupVoxelBased; cellSize_ = 1
    for file input (TOP_GPU.m:67).  Returns a finalised Model with .passive (0-based element ids)."""
    d = read_topvoxel(path)
    nelx, nely, nelz = d['nelx'], d['nely'], d['nelz']
    flat = np.zeros(nelx * nely * nelz, dtype=bool)
    flat[d['voxels'] - 1] = True
    vox = flat.reshape((nely, nelx, nelz), order='F')
    m = Model(vox, cell_size=1.0, poisson_ratio=nu, super_element=super_element, precision=precision, **kw)
    fwd = m.get_map(1, 'nodMapForward')
    fixing = _unique_sorted_rows(d['fixing'])
    if len(fixing):
        fixing = fixing.copy()
        fixing[:, 0] = fwd[_adapt_nodes(fixing[:, 0], nely, nelx, nelz, m) - 1]
    def prep(block):  # uniqued by node, sorted, moved to the padded grid, solid node numbers
        loads = _unique_sorted_rows(block).copy()
        if len(loads):
            loads[:, 0] = fwd[_adapt_nodes(loads[:, 0], nely, nelx, nelz, m) - 1]
        return loads
    m.set_conditions(fixing, prep(d['loads'][0]))
    for block in d['loads'][1:]:  # further load cases (TOP_GPU.m:453-461)
        m.add_loads(prep(block))
    m.finalize()
    info = m.level_info(1)
    pas = np.sort(d['passive'])
    if len(pas):
        e0 = pas - 1
        iy, ix, iz = e0 % nely, (e0 // nely) % nelx, e0 // (nely * nelx)
        full = iy + info['resY'] * (ix + info['resX'] * iz)
        pas = m.get_map(1, 'eleMapForward')[full] - 1
        pas = np.unique(pas[pas >= 0])
    m.passive = pas
    m.densities = d['densities']
    return m
