"""Scattering_inplace_noalloc (the allocation-free variant of Scattering_inplace) against
Scattering_inplace through the C ABI: two calls in a row on the same GPU, the second with a smaller
input, each bit-exact against the allocating entry point (both sum every node's entries in ascending
entry order), plus the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ga(a):
    import torch
    from top_gpu_on_matlab_amd.mex import GpuArray
    a = np.asarray(a)
    flat = torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order='F'))).cuda()
    return GpuArray(flat, a.shape)


def _np(ga):
    return ga.t.cpu().numpy().reshape(ga.dims, order='F')


def _case(rng, Ne, Nn, dtype):
    eNod = rng.integers(1, Nn + 1, size=(Ne, 8)).astype(np.int32)
    eNod[rng.integers(0, Ne, 5), rng.integers(0, 8, 5)] = Nn + 3     # out of range: skipped
    eNod[rng.integers(0, Ne, 5), rng.integers(0, 8, 5)] = 0
    V = rng.standard_normal((Ne, 24)).astype(dtype)
    Y0 = rng.standard_normal((Nn, 3)).astype(dtype)
    return eNod, V, Y0


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_noalloc_matches_scattering_inplace_on_two_calls(dtype):
    from top_gpu_on_matlab_amd import mex
    rng = np.random.default_rng(11)
    for Ne, Nn in ((60000, 70000), (9000, 12000)):                 # second call smaller
        print('noalloc dtype=%s Ne=%d Nn=%d' % (np.dtype(dtype).name, Ne, Nn), flush=True)
        eNod, V, Y0 = _case(rng, Ne, Nn, dtype)
        gE, gV = _ga(eNod), _ga(V)
        ya, yb = _ga(Y0), _ga(Y0)
        mex.Scattering_inplace(ya, gE, gV)
        mex.Scattering_inplace_noalloc(yb, gE, gV)
        assert np.array_equal(_np(ya), _np(yb))
        # and both against a plain accumulation (out-of-range ids skipped)
        ref = Y0.astype(np.float64).copy()
        for j in range(8):
            ok = (eNod[:, j] >= 1) & (eNod[:, j] <= Nn)
            for c in range(3):
                np.add.at(ref[:, c], eNod[ok, j] - 1, V[ok, c + 3 * j].astype(np.float64))
        tol = 1e-13 if dtype == np.float64 else 1e-5
        assert np.abs(_np(yb) - ref).max() < tol * np.abs(ref).max()


def test_noalloc_rejects_inputs_beyond_its_scratch():
    from top_gpu_on_matlab_amd import _lib
    L = _lib.lib()
    Ne, Nn = 8_000_000, 100
    args = (_lib.TgArray * 3)()
    for a, (cls, dims) in zip(args, ((_lib.TG_DOUBLE, (Nn, 3)), (_lib.TG_INT32, (Ne, 8)), (_lib.TG_DOUBLE, (Ne, 24)))):
        a.data = None
        a.cls = cls
        a.ndim = 2
        a.dims[0], a.dims[1] = dims
    st = L.Scattering_inplace_noalloc(0, (_lib.TgArray * 1)(), 3, args, None)
    assert st == 3
    assert L.tg_last_error_id().decode() == 'Scattering_inplace:shape'
    assert b'scratch' in L.tg_last_error_message()
