"""ctypes binding of libtopgpu.so (include/topgpu.h).

Loading fails loudly when the library has not been built: there is no fallback path.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, 'libtopgpu.so')

TG_OK = 0
TG_ERR_ARGS, TG_ERR_TYPE, TG_ERR_SHAPE, TG_ERR_HIP, TG_ERR_STATE, TG_ERR_NUMERIC = 1, 2, 3, 4, 5, 6
TG_SINGLE, TG_DOUBLE, TG_INT32, TG_UINT32 = 1, 2, 3, 4
TG_PREC_FP64, TG_PREC_MIXED = 0, 1

MAPS = dict(eNodMat=1, eleMapBack=2, eleMapForward=3, nodMapBack=4, nodMapForward=5, nodesOnBoundary=6,
            elementsOnBoundary=7, elementUpwardMap=8, transferMat=9, solidNodeMapCoarser2Finer=10,
            mapUniqueKes=11, eleWithFixedDOFs=12)
TABLES = dict(transferMatCoeffi=1, Ke=2, F=3, uniqueKesFree=4, uniqueKesFixed=5, I_mat=6, fixedDOFs=7)


class TopGpuError(RuntimeError):
    """Error raised by the HIP library; .identifier carries the reference-style '<Entry>:<Kind>' id."""

    def __init__(self, code, identifier, message):
        super().__init__('%s: %s (status %d)' % (identifier, message, code))
        self.code = code
        self.identifier = identifier


class TgArray(ctypes.Structure):
    _fields_ = [('data', ctypes.c_void_p), ('cls', ctypes.c_int32), ('ndim', ctypes.c_int32),
                ('dims', ctypes.c_int64 * 4)]


class TgParams(ctypes.Structure):
    _fields_ = [('youngs_modulus', ctypes.c_double), ('poisson_ratio', ctypes.c_double),
                ('cell_size', ctypes.c_double), ('weight_factor_jacobi', ctypes.c_double),
                ('coarsest_resolution_control', ctypes.c_int32), ('non_dyadic', ctypes.c_int32),
                ('vcycle_type', ctypes.c_int32), ('super_element', ctypes.c_int32), ('precision', ctypes.c_int32)]


class TgLevelInfo(ctypes.Structure):
    _fields_ = [('resX', ctypes.c_int32), ('resY', ctypes.c_int32), ('resZ', ctypes.c_int32),
                ('spanWidth', ctypes.c_int32), ('numElements', ctypes.c_int64), ('numNodes', ctypes.c_int64),
                ('numDOFs', ctypes.c_int64), ('intermediateNumNodes', ctypes.c_int64)]


# symbol -> (restype, argtypes); every symbol declared in include/topgpu.h
_VP, _I32, _I64, _D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_PA = ctypes.POINTER(TgArray)
SIGNATURES = {
    'tg_last_error_id': (ctypes.c_char_p, []),
    'tg_last_error_message': (ctypes.c_char_p, []),
    'tg_version': (ctypes.c_char_p, []),
    'Gathering_inplace': (ctypes.c_int, [ctypes.c_int, _PA, ctypes.c_int, _PA, _VP]),
    'Scattering_inplace': (ctypes.c_int, [ctypes.c_int, _PA, ctypes.c_int, _PA, _VP]),
    'scatter_accum3_inplace': (ctypes.c_int, [ctypes.c_int, _PA, ctypes.c_int, _PA, _VP]),
    'Scattering_inplace_noalloc': (ctypes.c_int, [ctypes.c_int, _PA, ctypes.c_int, _PA, _VP]),
    'assembleKs_level2_inplace': (ctypes.c_int, [ctypes.c_int, _PA, ctypes.c_int, _PA, _VP]),
    'assembleKs_higherLevel_inplace': (ctypes.c_int, [ctypes.c_int, _PA, ctypes.c_int, _PA, _VP]),
    'assembleKs_level2_superEle_inplace': (ctypes.c_int, [ctypes.c_int, _PA, ctypes.c_int, _PA, _VP]),
    'tg_params_default': (None, [ctypes.POINTER(TgParams)]),
    'tg_model_create': (ctypes.c_int, [_VP, _I32, _I32, _I32, ctypes.POINTER(TgParams), ctypes.POINTER(_VP)]),
    'tg_model_destroy': (None, [_VP]),
    'tg_model_builtin_case': (ctypes.c_int, [_VP, _I32]),
    'tg_model_set_conditions': (ctypes.c_int, [_VP, _VP, _I64, _VP, _I64]),
    'tg_model_add_loads': (ctypes.c_int, [_VP, _VP, _I64]),
    'tg_model_finalize': (ctypes.c_int, [_VP]),
    'tg_model_num_levels': (_I32, [_VP]),
    'tg_model_level_info': (ctypes.c_int, [_VP, _I32, ctypes.POINTER(TgLevelInfo)]),
    'tg_model_map_size': (_I64, [_VP, _I32, _I32]),
    'tg_model_get_map': (ctypes.c_int, [_VP, _I32, _I32, _VP]),
    'tg_model_table_size': (_I64, [_VP, _I32, _I32]),
    'tg_model_get_table': (ctypes.c_int, [_VP, _I32, _I32, _VP]),
    'tg_solver_create': (ctypes.c_int, [_VP, _VP, ctypes.POINTER(_VP)]),
    'tg_solver_destroy': (None, [_VP]),
    'tg_solver_set_modulus': (ctypes.c_int, [_VP, _VP]),
    'tg_solver_assemble': (ctypes.c_int, [_VP]),
    'tg_solver_kbyu': (ctypes.c_int, [_VP, _I32, _VP, _VP]),
    'tg_solver_vcycle': (ctypes.c_int, [_VP, _VP, _VP]),
    'tg_solver_restrict': (ctypes.c_int, [_VP, _I32, _VP, _VP]),
    'tg_solver_interpolate': (ctypes.c_int, [_VP, _I32, _VP, _VP]),
    'tg_solver_coarse_solve': (ctypes.c_int, [_VP, _VP, _VP]),
    'tg_solver_pcg': (ctypes.c_int, [_VP, _VP, _VP, _D, _I32, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_D)]),
    'tg_solver_get_diag': (ctypes.c_int, [_VP, _I32, _VP]),
    'tg_solver_get_ks': (ctypes.c_int, [_VP, _I32, _VP]),
    'tg_solver_last_timing': (ctypes.c_int, [_VP, ctypes.POINTER(_D)]),
    'tg_solver_time_kbyu': (ctypes.c_int, [_VP, _I32, _I32, _I32, ctypes.POINTER(_D)]),
    'tg_solver_set_tuning': (ctypes.c_int, [_VP, _I32, _I32]),
    'tg_solver_set_jacobi_weight': (ctypes.c_int, [_VP, _D]),
    'tg_solver_kfine_timing': (ctypes.c_int, [_VP, _I32, _I32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_D)]),
    'tg_solver_unit_compliance': (ctypes.c_int, [_VP, _VP, _VP]),
    'tg_solver_subvoxel_energy': (ctypes.c_int, [_VP, _VP, _VP]),
    'tg_compliance_grad': (ctypes.c_int, [_VP, _VP, _VP, _D, _D, _D, _D, _I32, _D, _I32, _VP, ctypes.POINTER(_D), _VP,
                                           ctypes.POINTER(_I32)]),
    'tg_simp_modulus': (ctypes.c_int, [_VP, _I64, _D, _D, _D, _VP, _VP]),
    'tg_sensitivity': (ctypes.c_int, [_VP, _VP, _I64, _I32, _D, _D, _D, _D, _VP, _VP]),
    'tg_elementwise': (ctypes.c_int, [_I32, _VP, _VP, _D, _I64, _VP, _VP]),
    'tg_reduce': (ctypes.c_int, [_I32, _VP, _VP, _I64, ctypes.POINTER(_D), _VP]),
    'tg_oc_update': (ctypes.c_int, [_VP, _VP, _VP, _VP, _I64, _D, _D, _VP, ctypes.POINTER(_D), ctypes.POINTER(_I32), _VP]),
    'tg_filter_create': (ctypes.c_int, [_VP, _D, _VP, ctypes.POINTER(_VP)]),
    'tg_filter_destroy': (None, [_VP]),
    'tg_filter_apply': (ctypes.c_int, [_VP, _VP, _VP, _D, _I32, ctypes.POINTER(_I32)]),
    'tg_convfilter_create': (ctypes.c_int, [_I32, _I32, _I32, _D, _I32, _VP, ctypes.POINTER(_VP)]),
    'tg_convfilter_destroy': (None, [_VP]),
    'tg_convfilter_forward': (ctypes.c_int, [_VP, _VP, _VP]),
    'tg_convfilter_adjoint': (ctypes.c_int, [_VP, _VP, _VP]),
}
TG_EW_MUL, TG_EW_DIV_FLOOR, TG_EW_SET_WHERE, TG_EW_AXPY = 0, 1, 2, 3
(TG_TUNE_KFINE_KERNEL, TG_TUNE_KFINE_ZCHUNK, TG_TUNE_TIME_KFINE, TG_TUNE_KELEM_PARTS, TG_TUNE_COARSE_GRAPH,
 TG_TUNE_KFINE_MFMA, TG_TUNE_COARSE_PERSISTENT, TG_TUNE_COARSE_STREAMING, TG_TUNE_COARSE_POLLS) = \
    1, 2, 3, 4, 5, 6, 7, 8, 9
TG_RED_SUM, TG_RED_DOT, TG_RED_MAXABSDIFF, TG_RED_SHARPNESS = 0, 1, 2, 3

_lib = None


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError('libtopgpu.so not built (%s); run `python -m top_gpu_on_matlab_amd.build`' % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != TG_OK:
        L = lib()
        raise TopGpuError(status, L.tg_last_error_id().decode(), L.tg_last_error_message().decode())
    return status
