"""CPU tests of the file formats: .TopVoxel models (parity of the library's file branch with the
oracle's restatement of CreateVoxelFEAmodel / AdaptBCExternalMdl) and the NIfTI / history writers."""
import os

import numpy as np
import pytest

from oracle import mesh as M
from tests.conftest import holey_mask

pytest.importorskip('top_gpu_on_matlab_amd._lib')


def _conditions(nely, nelx, nelz):
    """Clamp the x = 0 face, load the top-face nodes of a centred patch (unpadded node numbering)."""
    NY, NX = nely + 1, nelx + 1
    node = lambda y, x, z: 1 + y + NY * (x + NX * z)
    fix = [[node(y, 0, z), 1, 1, 1] for z in range(nelz + 1) for y in range(nely + 1)]
    cx, cz = nelx // 2, nelz // 2
    patch = [(x, z) for x in range(cx - 1, cx + 2) for z in range(cz - 1, cz + 2)]
    ld = [[node(0, x, z), 0.0, 0.0, -1.0 / len(patch)] for x, z in patch]
    return np.array(fix, float), np.array(ld, float)


def test_topvoxel_round_trip_and_parity(tmp_path):
    from top_gpu_on_matlab_amd.topvoxel import load_topvoxel_model, read_topvoxel, write_topvoxel
    nely, nelx, nelz = 9, 21, 11                       # padded to 16 x 24 x 16
    vox = holey_mask(nely, nelx, nelz, seed=4)
    fix, ld = _conditions(nely, nelx, nelz)
    # duplicated and unsorted rows are uniqued and sorted by the reader (TOP_GPU.m:470-486)
    fix_file = np.concatenate([fix[::-1], fix[:3]])
    path = os.path.join(tmp_path, 'model.TopVoxel')
    write_topvoxel(path, vox, fix_file, ld, passive=[5, 2])
    d = read_topvoxel(path)
    assert (d['nelx'], d['nely'], d['nelz']) == (nelx, nely, nelz)
    assert np.array_equal(np.sort(d['voxels']), np.flatnonzero(vox.reshape(-1, order='F')) + 1)
    assert d['fixing'].shape == fix_file.shape and np.allclose(d['loads'][0], ld)
    gm = load_topvoxel_model(path)
    om = M.build_file_model(nelx, nely, nelz, d['voxels'], d['fixing'], d['loads'][0])
    assert np.array_equal(gm.get_map(1, 'eNodMat'), om.levels[0].eNodMat)
    assert np.array_equal(gm.get_table(1, 'F'), om.F)
    assert np.array_equal(gm.get_table(1, 'fixedDOFs').astype(bool), om.fixedDOFs)
    assert np.array_equal(gm.get_map(1, 'mapUniqueKes'), om.mapUniqueKes)
    assert np.abs(gm.get_table(1, 'Ke') - om.levels[0].Ke).max() < 1e-16        # cellSize 1 for files
    assert list(gm.passive) == sorted(int(om.levels[0].eleMapForward[e]) - 1 for e in
                                      [(1 % nely) + 16 * ((1 // nely) % nelx), (4 % nely) + 16 * ((4 // nely) % nelx)])


def test_topvoxel_additional_load_cases(tmp_path):
    """A .TopVoxel file with two additional load cases ('Additional Loads: 2', TOP_GPU.m:451-460): the
    reader returns three load blocks, and the library's model has one load column per case, equal to
    the oracle's (each case's loaded nodes leave the fixations, TOP_GPU.m:1511-1524)."""
    from top_gpu_on_matlab_amd.topvoxel import load_topvoxel_model, read_topvoxel, write_topvoxel
    nely, nelx, nelz = 9, 21, 11
    vox = holey_mask(nely, nelx, nelz, seed=4)
    fix, ld = _conditions(nely, nelx, nelz)
    NY, NX = nely + 1, nelx + 1
    node = lambda y, x, z: 1 + y + NY * (x + NX * z)
    lateral = np.array([[node(y, nelx, z), 0.0, 1.0, 0.0] for y in range(3) for z in range(nelz + 1)])
    axial = np.array([[node(nely, nelx, nelz), 1.0, 0.0, 0.5]])
    path = os.path.join(tmp_path, 'loads.TopVoxel')
    write_topvoxel(path, vox, fix, ld, additional_loads=[lateral, axial])
    d = read_topvoxel(path)
    assert len(d['loads']) == 3 and np.allclose(d['loads'][1], lateral) and np.allclose(d['loads'][2], axial)
    gm = load_topvoxel_model(path)
    om = M.build_file_model(nelx, nely, nelz, d['voxels'], d['fixing'], d['loads'])
    F = gm.get_table(1, 'F')
    assert F.shape == om.F.shape == (3 * om.levels[0].numNodes, 3)
    assert np.array_equal(F, om.F)
    assert np.array_equal(gm.get_table(1, 'fixedDOFs').astype(bool), om.fixedDOFs)
    assert all(np.abs(F[:, k]).sum() > 0 for k in range(3))


def test_nifti_and_history_writers(tmp_path):
    from top_gpu_on_matlab_amd.io import read_history, read_nifti1, write_histories, write_nifti1
    V = np.random.default_rng(0).random((4, 6, 5))
    p = os.path.join(tmp_path, 'v.nii')
    write_nifti1(p, V)
    assert np.array_equal(read_nifti1(p), V)
    raw = open(p, 'rb').read()
    assert len(raw) == 352 + V.size * 8 and raw[344:347] == b'n+1'
    out = dict(lssIts=[3, 4], cHist=[1.25, 0.5], sharpHist=[0.4, 0.3], tHist=[[1, 2, 3, 4, 5], [6, 7, 8, 9, 10]])
    write_histories(str(tmp_path), out)
    assert read_history(str(tmp_path), 'c_Target.dat') == [1.25, 0.5]
    assert open(os.path.join(tmp_path, 'iters_Target.dat')).read() == '3\n4\n'
    line = open(os.path.join(tmp_path, 'c_Target.dat')).readline()
    assert line == '%30.16e\n' % 1.25 and len(line) == 31
    assert read_history(str(tmp_path), 'solvetiming_Target.dat') == [2.0, 7.0]
