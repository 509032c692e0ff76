"""GPU parity of the six MEX entry points (through the C ABI) against the oracle's restatement
of the reference kernels, on reference-shaped inputs."""
import numpy as np
import pytest

from oracle import mesh as M
from oracle import solver as S
from tests.conftest import random_modulus

pytestmark = pytest.mark.gpu


def _ga(a, dtype=None):
    """Column-major numpy array -> GpuArray on the device."""
    import torch
    from top_gpu_on_matlab_amd.mex import GpuArray
    a = np.asarray(a)
    if dtype is not None:
        a = a.astype(dtype)
    flat = torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order='F'))).cuda()
    return GpuArray(flat, a.shape if a.ndim > 1 else (a.shape[0], 1))


def _np(ga):
    return ga.t.cpu().numpy().reshape(ga.dims, order='F')


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_gather_scatter(dtype):
    from top_gpu_on_matlab_amd import mex
    rng = np.random.default_rng(0)
    Nn, Ne = 1000, 700
    eNod = rng.integers(1, Nn + 1, size=(Ne, 8)).astype(np.int32)
    u = rng.standard_normal((Nn, 3)).astype(dtype)
    gU = _ga(u)
    gM = _ga(np.zeros((Ne, 24), dtype))
    mex.Gathering_inplace(gM, gU, _ga(eNod))
    assert np.array_equal(_np(gM), S.gathering(u, eNod))
    mex.Gathering_inplace(gM, gU, _ga(eNod.astype(np.uint32)))      # uint32 indices accepted
    assert np.array_equal(_np(gM), S.gathering(u, eNod))
    V = rng.standard_normal((Ne, 24)).astype(dtype)
    Y0 = rng.standard_normal((Nn, 3)).astype(dtype)
    eBad = eNod.copy()
    eBad[3, 2] = Nn + 5                                              # skipped (Scattering_inplace.cu:70)
    gY = _ga(Y0)
    mex.Scattering_inplace(gY, _ga(eBad), _ga(V))
    ref = S.scattering(Y0.astype(np.float64), eBad, V.astype(np.float64))
    tol = 1e-13 if dtype == np.float64 else 1e-5
    assert np.abs(_np(gY) - ref).max() < tol * np.abs(ref).max()
    keys = rng.integers(0, 130, size=900).astype(np.int32)
    Tm = rng.standard_normal((900, 3)).astype(dtype)
    X0 = np.zeros((120, 3), dtype)
    gX = _ga(X0)
    mex.scatter_accum3_inplace(_ga(keys), _ga(Tm), gX)
    ref = S.scatter_accum3(keys, Tm.astype(np.float64), np.zeros((120, 3)))
    assert np.abs(_np(gX) - ref).max() < tol * np.abs(ref).max()
    # deterministic: a second run gives identical bits
    gX2 = _ga(X0)
    mex.scatter_accum3_inplace(_ga(keys), _ga(Tm), gX2)
    assert np.array_equal(_np(gX), _np(gX2))


def test_assemble_entries_match_oracle():
    from top_gpu_on_matlab_amd import mex
    m = M.build_cuboid_model(8, 16, 8, opt_case=1)
    lv1, lv2, lv3 = m.levels
    E = random_modulus(lv1.numElements, seed=5)
    Pch = S.build_pch(4)
    ref2 = S.assemble_ks_level2(lv2.elementUpwardMap, E, lv1.Ke.reshape(-1, order='F'), Pch, m.mapUniqueKes,
                                m.uniqueKesFree, m.uniqueKesFixed)
    dKs = _ga(np.zeros((24, 24, lv2.numElements)))
    mex.assembleKs_level2_inplace(dKs, _ga(lv2.elementUpwardMap), _ga(E), _ga(lv1.Ke.reshape(-1, order='F')),
                                  _ga(Pch), _ga(m.mapUniqueKes), _ga(m.uniqueKesFree), _ga(m.uniqueKesFixed))
    got2 = _np(dKs)
    assert np.abs(got2 - ref2).max() < 1e-14 * np.abs(ref2).max()
    # empty boundary matrices: every child uses E*K0
    ref2b = S.assemble_ks_level2(lv2.elementUpwardMap, E, lv1.Ke.reshape(-1, order='F'), Pch)
    dKs2 = _ga(np.zeros((24, 24, lv2.numElements)))
    empty = _ga(np.zeros((576, 0)))
    mex.assembleKs_level2_inplace(dKs2, _ga(lv2.elementUpwardMap), _ga(E), _ga(lv1.Ke.reshape(-1, order='F')),
                                  _ga(Pch), _ga(m.mapUniqueKes), empty, empty)
    assert np.abs(_np(dKs2) - ref2b).max() < 1e-14 * np.abs(ref2b).max()
    Psub = S.build_pch(2)
    ref3 = S.assemble_ks_higher_level(lv3.elementUpwardMap, ref2, Psub)
    dKs3 = _ga(np.zeros((24, 24, lv3.numElements)))
    mex.assembleKs_higherLevel_inplace(dKs3, _ga(lv3.elementUpwardMap), _ga(ref2), _ga(Psub))
    assert np.abs(_np(dKs3) - ref3).max() < 1e-14 * np.abs(ref3).max()


def test_assemble_super_element_entry():
    from top_gpu_on_matlab_amd import mex
    m = M.build_cuboid_model(10, 20, 10, opt_case=2, super_element=True)
    lv1, lv2 = m.levels[0], m.levels[1]
    Esub = random_modulus(64 * lv1.numElements, seed=6).reshape(64, -1, order='F')
    Kmat = m.I_mat @ Esub
    Pch = S.build_pch(4)
    ref = S.assemble_ks_level2_super(lv2.elementUpwardMap, Kmat, Pch, m.mapUniqueKes, Esub, m.uniqueKesFree,
                                     m.uniqueKesFixed)
    for free in (m.uniqueKesFree, m.uniqueKesFree.reshape(576 * 64, -1, order='F')):   # 3-D and 2-D forms
        dKs = _ga(np.zeros((24, 24, lv2.numElements)))
        mex.assembleKs_level2_superEle_inplace(dKs, _ga(lv2.elementUpwardMap), _ga(Kmat), _ga(Pch),
                                               _ga(m.mapUniqueKes), _ga(Esub), _ga(free), _ga(m.uniqueKesFixed))
        assert np.abs(_np(dKs) - ref).max() < 1e-13 * np.abs(ref).max()


def test_entry_errors_on_device_inputs():
    from top_gpu_on_matlab_amd import mex
    from top_gpu_on_matlab_amd._lib import TopGpuError
    u = _ga(np.zeros((10, 3)))
    with pytest.raises(TopGpuError) as ei:
        mex.Gathering_inplace(_ga(np.zeros((4, 24), np.float32)), u, _ga(np.ones((4, 8), np.int32)))
    assert ei.value.identifier == 'Gathering_inplace:Type'
    with pytest.raises(TopGpuError) as ei:
        mex.scatter_accum3_inplace(_ga(np.ones(5, np.int32)), _ga(np.zeros((4, 3))), _ga(np.zeros((3, 3))))
    assert ei.value.identifier == 'scatter_accum3:shape'


def test_entries_accept_empty_inputs():
    """Zero elements (and zero keys) are valid inputs for all six entries: nothing is launched and the
    accumulated outputs keep their values (the reference's kernels run zero blocks)."""
    from top_gpu_on_matlab_amd import mex
    Y0 = np.arange(30, dtype=np.float64).reshape(10, 3)
    u = _ga(Y0)
    gM = _ga(np.zeros((0, 24)))
    mex.Gathering_inplace(gM, u, _ga(np.zeros((0, 8), np.int32)))
    gY = _ga(Y0)
    mex.Scattering_inplace(gY, _ga(np.zeros((0, 8), np.int32)), _ga(np.zeros((0, 24))))
    assert np.array_equal(_np(gY), Y0)
    gX = _ga(Y0)
    mex.scatter_accum3_inplace(_ga(np.zeros(0, np.int32)), _ga(np.zeros((0, 3))), gX)
    assert np.array_equal(_np(gX), Y0)
    m = M.build_cuboid_model(8, 16, 8, opt_case=1)
    lv1 = m.levels[0]
    E = random_modulus(lv1.numElements, seed=5)
    Pch, Psub = S.build_pch(4), S.build_pch(2)
    empty_ks = _ga(np.zeros((24, 24, 0)))
    mex.assembleKs_level2_inplace(empty_ks, _ga(np.zeros((0, 64), np.int32)), _ga(E),
                                  _ga(lv1.Ke.reshape(-1, order='F')), _ga(Pch), _ga(m.mapUniqueKes),
                                  _ga(m.uniqueKesFree), _ga(m.uniqueKesFixed))
    mex.assembleKs_higherLevel_inplace(empty_ks, _ga(np.zeros((0, 8), np.int32)),
                                       _ga(np.zeros((24, 24, 5))), _ga(Psub))
    ms = M.build_cuboid_model(10, 20, 10, opt_case=2, super_element=True)
    Esub = random_modulus(64 * ms.levels[0].numElements, seed=6).reshape(64, -1, order='F')
    mex.assembleKs_level2_superEle_inplace(empty_ks, _ga(np.zeros((0, 64), np.int32)), _ga(ms.I_mat @ Esub),
                                           _ga(Pch), _ga(ms.mapUniqueKes), _ga(Esub), _ga(ms.uniqueKesFree),
                                           _ga(ms.uniqueKesFixed))


def test_scatter_incidence_cache_follows_the_connectivity():
    """The scatter entry points keep the sorted incidence per connectivity; reuse must follow the
    array's content (a connectivity rewritten in place is sorted again) and stay bitwise repeatable."""
    import torch
    from top_gpu_on_matlab_amd import mex
    rng = np.random.default_rng(3)
    Nn, Ne = 3000, 2000
    eA = rng.integers(1, Nn + 1, size=(Ne, 8)).astype(np.int32)
    eB = rng.integers(1, Nn + 1, size=(Ne, 8)).astype(np.int32)
    V = rng.standard_normal((Ne, 24))
    gV = _ga(V)
    gA, gB = _ga(eA), _ga(eB)

    def run(gE):
        gY = _ga(np.zeros((Nn, 3)))
        mex.Scattering_inplace(gY, gE, gV)
        return _np(gY)

    rA, rB = S.scattering(np.zeros((Nn, 3)), eA, V), S.scattering(np.zeros((Nn, 3)), eB, V)
    yA = run(gA)
    assert np.abs(yA - rA).max() < 1e-13 * np.abs(rA).max()
    assert np.array_equal(run(gA), yA)                              # cached: identical bits
    assert np.abs(run(gB) - rB).max() < 1e-13 * np.abs(rB).max()    # a second connectivity
    assert np.array_equal(run(gA), yA)
    gA.t.copy_(torch.from_numpy(np.ascontiguousarray(eB.reshape(-1, order='F'))).cuda())  # rewritten in place
    assert np.abs(run(gA) - rB).max() < 1e-13 * np.abs(rB).max()
    # scatter_accum3: the same keys array rewritten in place
    keys = rng.integers(1, 400, size=5000).astype(np.int32)
    Tm = rng.standard_normal((5000, 3))
    gk, gT = _ga(keys), _ga(Tm)
    for it in range(2):
        gX = _ga(np.zeros((390, 3)))
        mex.scatter_accum3_inplace(gk, gT, gX)
        ref = S.scatter_accum3(keys, Tm, np.zeros((390, 3)))
        assert np.abs(_np(gX) - ref).max() < 1e-13 * np.abs(ref).max()
        keys = rng.integers(1, 400, size=5000).astype(np.int32)
        gk.t.copy_(torch.from_numpy(keys).cuda())
