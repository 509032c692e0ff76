"""Bit-exact parity of the library's host model (C++ hierarchy builder) with the oracle's
literal restatement of TOP_GPU.m's setup: every index map, coefficient table, load vector and
BC-modified element matrix."""
import numpy as np
import pytest

from oracle import mesh as M
from tests.conftest import holey_mask

pytest.importorskip('top_gpu_on_matlab_amd._lib')

CASES = [
    dict(shape=(8, 16, 8), case=1),
    dict(shape=(24, 48, 24), case=1),                   # C1
    dict(shape=(10, 60, 10), case=2),                   # padded to 16 x 64 x 16
    dict(shape=(10, 60, 10), case=2, se=True),          # C2, super-element
    dict(shape=(12, 24, 12), case=1, holes=0),
    dict(shape=(9, 21, 13), case=1, holes=5),           # odd sizes, padding on all axes
    dict(shape=(16, 32, 16), case=1, holes=7, nd=0),    # dyadic hierarchy (nonDyadic_ = 0)
    dict(shape=(40, 80, 40), case=1, holes=1),          # 4 levels before the non-dyadic merge
]


@pytest.mark.parametrize('c', CASES, ids=lambda c: '%s-%s' % ('x'.join(map(str, c['shape'])), c))
def test_maps_bit_exact(c):
    from top_gpu_on_matlab_amd.model import Model
    ny, nx, nz = c['shape']
    vox = holey_mask(ny, nx, nz, seed=c['holes']) if 'holes' in c else np.ones(c['shape'], bool)
    se = c.get('se', False)
    nd = c.get('nd', 1)
    om = M.build_cuboid_model(ny, nx, nz, opt_case=c['case'], voxels=vox, super_element=se, non_dyadic=nd)
    gm = Model(vox, super_element=se, non_dyadic=nd)
    gm.builtin_case(c['case'])
    gm.finalize()
    assert gm.num_levels == om.numLevels
    for level in range(1, om.numLevels + 1):
        lv = om.levels[level - 1]
        info = gm.level_info(level)
        assert (info['resX'], info['resY'], info['resZ'], info['numElements'], info['numNodes']) == \
               (lv.resX, lv.resY, lv.resZ, lv.numElements, lv.numNodes)
        names = ['eNodMat', 'eleMapBack', 'eleMapForward', 'nodMapBack', 'nodMapForward', 'nodesOnBoundary']
        names += ['elementsOnBoundary'] if level == 1 else ['elementUpwardMap', 'transferMat',
                                                             'solidNodeMapCoarser2Finer']
        for n in names:
            a, b = gm.get_map(level, n), getattr(lv, n)
            assert a.dtype == np.int32 and a.shape == b.shape and np.array_equal(a, b), (level, n)
        if level > 1:
            assert info['intermediateNumNodes'] == lv.intermediateNumNodes
            assert np.array_equal(gm.get_table(level, 'transferMatCoeffi'), lv.transferMatCoeffi)
    assert np.array_equal(gm.get_map(1, 'mapUniqueKes'), om.mapUniqueKes)
    assert np.array_equal(gm.get_map(1, 'eleWithFixedDOFs'), om.eleWithFixedDOFs)
    assert np.array_equal(gm.get_table(1, 'fixedDOFs').astype(bool), om.fixedDOFs)
    assert np.array_equal(gm.get_table(1, 'F'), om.F)
    for n, ref in (('uniqueKesFree', om.uniqueKesFree), ('uniqueKesFixed', om.uniqueKesFixed)):
        a = gm.get_table(1, n)
        assert a.shape == ref.shape and np.abs(a - ref).max() <= 1e-15 * np.abs(ref).max()
    assert np.abs(gm.get_table(1, 'Ke') - om.levels[0].Ke).max() < 1e-16
    if se:
        assert np.abs(gm.get_table(1, 'I_mat') - om.I_mat).max() < 1e-16
