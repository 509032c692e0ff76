"""Synthetic non-cuboid design domain: seeded value noise, thresholded, largest 26-connected component,
with the clamped-bottom / top-patch-load boundary conditions, written as a .TopVoxel file (the
reference's file input, TOP_GPU.m:416-496).

    mask = value_noise_mask(256, 128, 128)            # nx, ny, nz -> bool [ny, nx, nz]
    fixing, loads = bottom_clamp_top_patch(mask)
    write_topvoxel('domain.TopVoxel', mask, fixing, loads)

Value noise: for octave o (0 .. octaves-1) a lattice of period period/2^o voxels carries hashed values
in [0, 1) (an integer hash of the lattice point, the octave and the seed), trilinearly interpolated
at voxel centres and summed with amplitude 2^-o.  The solid set is the top `solid_fraction` of the
noise values; only its largest connected component (26-connected by default) is kept.
"""
import numpy as np
import scipy.ndimage

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _hash01(ix, iy, iz, octave, seed):
    """splitmix64 finaliser of the packed lattice coordinates -> uniform [0, 1)."""
    with np.errstate(over='ignore'):
        k = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
             ^ iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
             ^ iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)
             ^ np.uint64(octave) * np.uint64(0xD6E8FEB86659FD93)
             ^ np.uint64(seed) * np.uint64(0xFF51AFD7ED558CCD))
        k = (k ^ (k >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        k = (k ^ (k >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        k = k ^ (k >> np.uint64(31))
    return (k >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def value_noise(nx, ny, nz, seed=0, octaves=4, period=32):
    """Noise sampled at voxel centres, array [ny, nx, nz] (the reference's nely x nelx x nelz order)."""
    y, x, z = np.meshgrid(np.arange(ny) + 0.5, np.arange(nx) + 0.5, np.arange(nz) + 0.5, indexing='ij')
    total = np.zeros((ny, nx, nz))
    for o in range(octaves):
        p = period / 2.0 ** o
        gx, gy, gz = x / p, y / p, z / p
        ix, iy, iz = np.floor(gx).astype(np.int64), np.floor(gy).astype(np.int64), np.floor(gz).astype(np.int64)
        fx, fy, fz = gx - ix, gy - iy, gz - iz
        v = np.zeros_like(total)
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                    v += w * _hash01(ix + dx, iy + dy, iz + dz, o, seed)
        total += v * 0.5 ** o
    return total


def largest_component(mask, connectivity=26):
    """Largest connected component: 26 (faces, edges, corners) or 6 (faces only)."""
    st = np.ones((3, 3, 3), dtype=bool) if connectivity == 26 else scipy.ndimage.generate_binary_structure(3, 1)
    lab, n = scipy.ndimage.label(mask, structure=st)
    if n == 0:
        return mask.copy()
    sizes = np.bincount(lab.reshape(-1))
    sizes[0] = 0
    return lab == int(np.argmax(sizes))


def value_noise_mask(nx, ny, nz, seed=0, octaves=4, period=32, solid_fraction=0.4, connectivity=26):
    """bool [ny, nx, nz]: the top solid_fraction of the noise, largest connected component.  With the
    default 26-connectivity, voxel clusters joined only along an edge or at a corner stay in the
    domain; such hinges are mechanisms of the Hex8 mesh (K singular), on which CG runs to maxIT.
    connectivity=6 keeps face-connected voxels only."""
    v = value_noise(nx, ny, nz, seed, octaves, period)
    thr = np.quantile(v, 1.0 - solid_fraction)
    return largest_component(v > thr, connectivity)


def bottom_clamp_top_patch(mask, clamp_layers=4, patch=8, total_load=1.0):
    """Boundary conditions in .TopVoxel node numbering (1-based, unpadded (ny+1)(nx+1)(nz+1) grid):
    every mesh-boundary node of a solid voxel in the bottom `clamp_layers` z layers is fixed in x, y
    and z; a downward load totalling `total_load` is spread evenly over the top-face nodes of the
    highest solid voxel of each column of the centred patch x patch voxel window.  Only boundary nodes
    are clamped: the reference pairs boundary fixed nodes with the fixing rows in order, so interior
    fixations would be misapplied (TOP_GPU.m:1587-1604)."""
    ny, nx, nz = mask.shape
    NY, NX, NZ = ny + 1, nx + 1, nz + 1
    # number of solid voxels around every node
    cnt = np.zeros((NY, NX, NZ), dtype=np.int32)
    m = mask.astype(np.int32)
    for dy in (0, 1):
        for dx in (0, 1):
            for dz in (0, 1):
                cnt[dy:dy + ny, dx:dx + nx, dz:dz + nz] += m
    node_id = lambda y, x, z: 1 + y + NY * (x + NX * z)
    touched = np.zeros((NY, NX, NZ), dtype=bool)
    low = mask[:, :, :clamp_layers]
    for dy in (0, 1):
        for dx in (0, 1):
            for dz in (0, 1):
                touched[dy:dy + ny, dx:dx + nx, dz:dz + clamp_layers] |= low
    fixed = touched & (cnt > 0) & (cnt < 8)
    fy, fx, fz = np.nonzero(fixed)
    fix_ids = node_id(fy, fx, fz)
    # loads
    y0, x0 = (ny - patch) // 2, (nx - patch) // 2
    nodes = set()
    for yy in range(y0, y0 + patch):
        for xx in range(x0, x0 + patch):
            col = np.flatnonzero(mask[yy, xx, :])
            if col.size == 0:
                continue
            zt = int(col[-1]) + 1  # top face of the highest solid voxel
            for dy in (0, 1):
                for dx in (0, 1):
                    nodes.add(int(node_id(yy + dy, xx + dx, zt)))
    load_ids = np.array(sorted(nodes), dtype=np.int64)
    fix_ids = np.setdiff1d(fix_ids, load_ids)
    fixing = np.column_stack([fix_ids, np.ones((fix_ids.size, 3))]).astype(np.float64)
    f = total_load / max(load_ids.size, 1)
    loads = np.column_stack([load_ids, np.zeros(load_ids.size), np.zeros(load_ids.size),
                             -f * np.ones(load_ids.size)]).astype(np.float64)
    return fixing, loads
