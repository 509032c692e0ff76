"""C ABI checks that need no GPU: the library loads, exports every symbol declared in
include/topgpu.h, builds models on the host, and rejects bad MEX arguments with the
reference's error identifiers before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
lib_mod = pytest.importorskip('top_gpu_on_matlab_amd._lib')


def declared_symbols():
    src = open(os.path.join(ROOT, 'include', 'topgpu.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b([A-Za-z_][A-Za-z0-9_]*)\s*\(', src)) - {'if', 'sizeof'})


def test_python_constants_match_the_header():
    """Every TG_* enumerator of include/topgpu.h that the Python binding names has the header's value
    (status codes, classes, precisions, tuning keys)."""
    src = re.sub(r'/\*.*?\*/', '', open(os.path.join(ROOT, 'include', 'topgpu.h')).read(), flags=re.S)
    header = {k: int(v) for k, v in re.findall(r'\b(TG_[A-Z0-9_]+)\s*=\s*(\d+)', src)}
    named = {k: getattr(lib_mod, k) for k in dir(lib_mod) if k.startswith('TG_') and isinstance(getattr(lib_mod, k), int)}
    for k in ('TG_ERR_HIP', 'TG_TUNE_COARSE_PERSISTENT', 'TG_TUNE_COARSE_POLLS', 'TG_TUNE_KFINE_MFMA'):
        assert k in header and k in named, k
    wrong = {k: (v, header[k]) for k, v in named.items() if k in header and header[k] != v}
    assert not wrong, wrong


def test_every_declared_symbol_is_exported():
    L = lib_mod.lib()
    names = declared_symbols()
    assert 'Gathering_inplace' in names and 'tg_solver_pcg' in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(lib_mod.SIGNATURES), set(names) - set(lib_mod.SIGNATURES)
    assert b'gfx950' in L.tg_version()


def _desc(cls, dims, ptr=None):
    a = lib_mod.TgArray()
    a.data = ptr
    a.cls = cls
    a.ndim = len(dims)
    for i, d in enumerate(dims):
        a.dims[i] = d
    return a


def _call(name, args, nlhs=0):
    L = lib_mod.lib()
    arr = (lib_mod.TgArray * max(len(args), 1))(*args)
    out = (lib_mod.TgArray * 1)()
    st = getattr(L, name)(nlhs, out, len(args), arr, None)
    return st, L.tg_last_error_id().decode()


D, S, I32 = lib_mod.TG_DOUBLE, lib_mod.TG_SINGLE, lib_mod.TG_INT32


@pytest.mark.parametrize('name,args,nlhs,ident', [
    ('Gathering_inplace', [_desc(D, (4, 24)), _desc(D, (10, 3))], 0, 'Gathering_inplace:Args'),
    ('Gathering_inplace', [_desc(D, (4, 24)), _desc(D, (10, 3)), _desc(I32, (4, 8))], 1, 'Gathering_inplace:Args'),
    ('Gathering_inplace', [_desc(S, (4, 24)), _desc(D, (10, 3)), _desc(I32, (4, 8))], 0, 'Gathering_inplace:Type'),
    ('Gathering_inplace', [_desc(D, (4, 24)), _desc(D, (10, 3)), _desc(D, (4, 8))], 0, 'Gathering_inplace:Type'),
    ('Gathering_inplace', [_desc(D, (4, 24)), _desc(D, (10, 2)), _desc(I32, (4, 8))], 0, 'Gathering_inplace:Shape'),
    ('Gathering_inplace', [_desc(D, (5, 24)), _desc(D, (10, 3)), _desc(I32, (4, 8))], 0, 'Gathering_inplace:Shape'),
    ('Scattering_inplace', [_desc(D, (10, 3)), _desc(I32, (4, 8))], 0, 'Scattering_inplace:arity'),
    ('Scattering_inplace', [_desc(D, (10, 3)), _desc(D, (4, 8)), _desc(D, (4, 24))], 0, 'Scattering_inplace:type'),
    ('Scattering_inplace', [_desc(D, (10, 4)), _desc(I32, (4, 8)), _desc(D, (4, 24))], 0, 'Scattering_inplace:shape'),
    ('Scattering_inplace', [_desc(D, (10, 3)), _desc(I32, (4, 8)), _desc(D, (4, 23))], 0, 'Scattering_inplace:shape'),
    ('scatter_accum3_inplace', [_desc(I32, (5, 1)), _desc(D, (5, 3))], 0, 'scatter_accum3:args'),
    ('scatter_accum3_inplace', [_desc(I32, (5, 1)), _desc(D, (5, 3)), _desc(D, (4, 3))], 1, 'scatter_accum3:nlhs'),
    ('assembleKs_level2_inplace', [_desc(D, (24, 24, 3))] * 7, 0, 'assembleKs_level2_inplace:Input'),
    ('assembleKs_level2_inplace', [_desc(D, (24, 24, 3)), _desc(D, (3, 64))] + [_desc(D, (1, 1))] * 6, 0,
     'assembleKs_level2_inplace:Type'),
    ('assembleKs_higherLevel_inplace', [_desc(D, (24, 24, 3))] * 3, 0, 'assembleKs_higherLevel_inplace:Input'),
    ('assembleKs_higherLevel_inplace', [_desc(D, (24, 23, 3)), _desc(I32, (3, 8)), _desc(D, (24, 24, 5)),
                                        _desc(D, (24, 24, 8))], 0, 'assembleKs_higherLevel_inplace:Shape'),
    ('assembleKs_level2_superEle_inplace', [_desc(D, (24, 24, 3))] * 4, 0,
     'assembleKs_level2_superEle_inplace:Input'),
])
def test_mex_argument_errors(name, args, nlhs, ident):
    st, got = _call(name, args, nlhs)
    assert st != 0 and got == ident


def test_model_errors_and_pairing():
    from top_gpu_on_matlab_amd._lib import TopGpuError
    from top_gpu_on_matlab_amd.model import Model
    with pytest.raises(TopGpuError) as ei:
        Model(np.ones((2, 8, 8), bool))
    assert ei.value.identifier == 'CreateVoxelFEAmodel:Input'
    m = Model(np.ones((8, 8, 8), bool))
    with pytest.raises(TopGpuError) as ei:
        m.finalize()
    assert 'State' in ei.value.identifier
    # an interior fixed node sorted before a boundary one breaks the pairing of fixedNodesStruct
    # with fixingCond_ rows (TOP_GPU.m:1587-1604): the library refuses it
    interior = int(m.get_map(1, 'nodMapForward')[4 + 9 * (4 + 9 * 4)])
    m.set_conditions([[interior, 1, 1, 1], [m.num_nodes, 1, 1, 1]], [[1, 0, 0, -1]])
    with pytest.raises(TopGpuError) as ei:
        m.finalize()
    assert ei.value.identifier == 'Solving_SetupKeWithFixedDOFs:Pairing'
    # an interior fixation sorted after every boundary one is silently left out of the element
    # matrices, as in the reference
    m.set_conditions([[1, 1, 1, 1], [interior, 1, 1, 1]], [[m.num_nodes, 0, 0, -1]])
    m.finalize()
    assert m.get_table(1, 'fixedDOFs')[3 * (interior - 1)] == 1


def test_solver_level_entry_points_reject_bad_arguments_without_a_device():
    """Null handles and bad options are rejected before any device work (status + identifier)."""
    import ctypes
    L = lib_mod.lib()
    c = ctypes.c_double()
    its = ctypes.c_int32()
    st = L.tg_compliance_grad(None, None, None, 1e-6, 1.0, 3.0, 1e-3, 800, 1e-4, 0, None, ctypes.byref(c), None,
                              ctypes.byref(its))
    assert st == 1 and L.tg_last_error_id() == b'tg_compliance_grad:Input'
    assert L.tg_solver_set_tuning(None, 1, 0) == 1
    assert L.tg_solver_set_jacobi_weight(None, 0.5) == 1
    assert L.tg_last_error_id() == b'tg_solver_set_jacobi_weight:Input'
    n = ctypes.c_int64()
    assert L.tg_solver_kfine_timing(None, 0, 0, ctypes.byref(n), ctypes.byref(c)) == 1
    assert L.tg_last_error_id() == b'tg_solver_kfine_timing:Input'
