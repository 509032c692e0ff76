"""Host model: voxel discretisation, multigrid hierarchy and boundary conditions (libtopgpu).

Mirrors CreateVoxelFEAmodel / FEA_VoxelBasedDiscretization / Solving_BuildingMeshHierarchy /
FEA_ApplyBoundaryCondition / Solving_SetupKeWithFixedDOFs (TOP_GPU.m:403-536, 1276-1727,
2019-2119).  Index maps come back exactly as the reference stores them: int32, 1-based,
column-major shapes.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

PRECISIONS = {'fp64': _lib.TG_PREC_FP64, 'mixed': _lib.TG_PREC_MIXED}


class Model:
    """meshHierarchy_ and the boundary-condition globals of TOP_GPU.m, built by libtopgpu."""

    def __init__(self, voxels, cell_size=None, poisson_ratio=0.3, youngs_modulus=1.0, super_element=False,
                 precision='fp64', non_dyadic=1, coarsest_resolution_control=20000,
                 weight_factor_jacobi=0.6, vcycle_type='Standard'):
        vox = np.asarray(voxels)
        if vox.ndim != 3:
            raise ValueError('voxels must be a 3-D array [nely, nelx, nelz]')
        self.nely, self.nelx, self.nelz = (int(v) for v in vox.shape)
        p = _lib.TgParams()
        lib().tg_params_default(ctypes.byref(p))
        p.youngs_modulus = youngs_modulus
        p.poisson_ratio = poisson_ratio
        p.cell_size = (1.0 / self.nelx) if cell_size is None else float(cell_size)
        p.weight_factor_jacobi = weight_factor_jacobi
        p.coarsest_resolution_control = coarsest_resolution_control
        p.non_dyadic = non_dyadic
        p.vcycle_type = {'Standard': 0, 'Adapted': 1}[vcycle_type]
        p.super_element = 1 if super_element else 0
        p.precision = PRECISIONS[precision]
        self.params = p
        self.super_element = bool(super_element)
        self.precision = precision
        buf = np.ascontiguousarray(vox.astype(np.uint8).reshape(-1, order='F'))
        h = ctypes.c_void_p()
        check(lib().tg_model_create(buf.ctypes.data, self.nely, self.nelx, self.nelz, ctypes.byref(p),
                                    ctypes.byref(h)))
        self.h = h
        self.finalized = False

    def __del__(self):
        h = getattr(self, 'h', None)
        if h:
            self.h = None
            try:
                lib().tg_model_destroy(h)
            except Exception:   # interpreter shutdown: module globals already torn down
                pass

    # ---- conditions ----
    def builtin_case(self, opt_case):
        check(lib().tg_model_builtin_case(self.h, int(opt_case)))
        self.finalized = False

    def set_conditions(self, fixing, loads):
        fx = np.asfortranarray(np.asarray(fixing, dtype=np.float64))
        ld = np.asfortranarray(np.asarray(loads, dtype=np.float64))
        check(lib().tg_model_set_conditions(self.h, fx.ctypes.data, fx.shape[0], ld.ctypes.data, ld.shape[0]))
        self.finalized = False

    def add_loads(self, loads):
        """One more load case [n x 4] (node, Fx, Fy, Fz) after set_conditions: a .TopVoxel file's further
        load blocks (loadingCond_{2..}, TOP_GPU.m:453-461), solved and weighted equally by the
        optimisation (objWeightingList_, TOP_GPU.m:461, 599)."""
        ld = np.asfortranarray(np.asarray(loads, dtype=np.float64).reshape(-1, 4))
        check(lib().tg_model_add_loads(self.h, ld.ctypes.data, ld.shape[0]))
        self.finalized = False

    def finalize(self):
        check(lib().tg_model_finalize(self.h))
        self.finalized = True

    # ---- queries ----
    @property
    def num_levels(self):
        return int(lib().tg_model_num_levels(self.h))

    def level_info(self, level):
        info = _lib.TgLevelInfo()
        check(lib().tg_model_level_info(self.h, int(level), ctypes.byref(info)))
        return {f: getattr(info, f) for f, _ in _lib.TgLevelInfo._fields_}

    def get_map(self, level, name):
        which = _lib.MAPS[name]
        n = lib().tg_model_map_size(self.h, int(level), which)
        if n < 0:
            raise KeyError('map %s not available on level %d' % (name, level))
        out = np.empty(n, dtype=np.int32)
        check(lib().tg_model_get_map(self.h, int(level), which, out.ctypes.data))
        info = self.level_info(level)
        Ne = info['numElements']
        span = info['spanWidth']
        if name == 'eNodMat':
            return out.reshape((Ne, 8), order='F')
        if name == 'elementUpwardMap':
            return out.reshape((Ne, span ** 3), order='F')
        if name == 'transferMat':
            return out.reshape(((span + 1) ** 3, Ne), order='F')
        return out

    def get_table(self, level, name):
        which = _lib.TABLES[name]
        n = lib().tg_model_table_size(self.h, int(level), which)
        if n < 0:
            raise KeyError('table %s not available on level %d' % (name, level))
        out = np.empty(n, dtype=np.float64)
        check(lib().tg_model_get_table(self.h, int(level), which, out.ctypes.data))
        if name == 'Ke':
            return out.reshape((24, 24), order='F')
        if name == 'I_mat':
            return out.reshape((576, 64), order='F')
        if name in ('uniqueKesFree',):
            nU = self.map_size(1, 'eleWithFixedDOFs')
            return out.reshape((576, 64, nU), order='F') if self.super_element else out.reshape((576, nU), order='F')
        if name == 'uniqueKesFixed':
            return out.reshape((576, -1), order='F')
        if name == 'F':
            return out.reshape((3 * self.level_info(1)['numNodes'], -1), order='F')
        return out

    def map_size(self, level, name):
        return int(lib().tg_model_map_size(self.h, int(level), _lib.MAPS[name]))

    @property
    def num_elements(self):
        return self.level_info(1)['numElements']

    @property
    def num_nodes(self):
        return self.level_info(1)['numNodes']


def cuboid_model(nely, nelx, nelz, opt_case=1, **kw):
    """Built-in cuboid case (TOP_GPU(true(nely,nelx,nelz), 'optCase', opt_case))."""
    m = Model(np.ones((nely, nelx, nelz), dtype=bool), **kw)
    m.builtin_case(opt_case)
    m.finalize()
    return m
