"""The six MEX entry points, callable with torch CUDA tensors.

Each function takes its arguments in the reference's prhs order; a tensor argument stands for
a gpuArray with the same column-major shape (pass `shape=` on a 1-D tensor, or a tensor whose
.T is contiguous).  Results are written in place exactly as the MEX functions do.
"""
import ctypes

import torch

from . import _lib
from ._lib import check, lib

_CLS = {torch.float32: _lib.TG_SINGLE, torch.float64: _lib.TG_DOUBLE, torch.int32: _lib.TG_INT32}
if hasattr(torch, 'uint32'):
    _CLS[torch.uint32] = _lib.TG_UINT32


class GpuArray:
    """A device buffer with a MATLAB class and column-major dims (an mxGPUArray stand-in)."""

    def __init__(self, tensor, dims):
        if not tensor.is_cuda:
            raise ValueError('GpuArray needs a device tensor')
        self.t = tensor
        self.dims = tuple(int(d) for d in dims)

    @staticmethod
    def from_fortran(t):
        """Wrap a tensor holding a column-major array: t.permute(reversed) must be contiguous."""
        tt = t.permute(*reversed(range(t.dim())))
        if not tt.is_contiguous():
            raise ValueError('tensor is not column-major contiguous')
        return GpuArray(tt.reshape(-1), t.shape)

    def desc(self):
        a = _lib.TgArray()
        a.data = self.t.data_ptr() if self.t.numel() else None
        a.cls = _CLS.get(self.t.dtype, 0)
        a.ndim = len(self.dims)
        for i, d in enumerate(self.dims[:4]):
            a.dims[i] = d
        return a


def _call(fn, args, nlhs=0, stream=None):
    arr = (_lib.TgArray * len(args))(*[a.desc() for a in args])
    out = (_lib.TgArray * 1)()
    s = stream if stream is not None else torch.cuda.current_stream()
    check(getattr(lib(), fn)(nlhs, out, len(args), arr, ctypes.c_void_p(s.cuda_stream)))


def Gathering_inplace(gUMat, gU, gE, stream=None):
    _call('Gathering_inplace', [gUMat, gU, gE], stream=stream)


def Scattering_inplace(Y, eNodMat, uMat, stream=None):
    _call('Scattering_inplace', [Y, eNodMat, uMat], stream=stream)


def Scattering_inplace_noalloc(Y, eNodMat, uMat, stream=None):
    """Scattering_inplace without any device allocation (fixed library scratch; see include/topgpu.h)."""
    _call('Scattering_inplace_noalloc', [Y, eNodMat, uMat], stream=stream)


def scatter_accum3_inplace(keys, Tmp, xFiner, stream=None):
    _call('scatter_accum3_inplace', [keys, Tmp, xFiner], stream=stream)


def assembleKs_level2_inplace(d_Ks, d_upMap, d_eleModulus, d_K_0, d_Pch, d_mapUnique, d_uniqueFree, d_uniqueFixed,
                              stream=None):
    _call('assembleKs_level2_inplace', [d_Ks, d_upMap, d_eleModulus, d_K_0, d_Pch, d_mapUnique, d_uniqueFree,
                                        d_uniqueFixed], stream=stream)


def assembleKs_higherLevel_inplace(d_Ks, d_upMap, d_KsPrev, d_Psub, stream=None):
    _call('assembleKs_higherLevel_inplace', [d_Ks, d_upMap, d_KsPrev, d_Psub], stream=stream)


def assembleKs_level2_superEle_inplace(d_Ks, d_upMap, d_Ks0, d_Psub, d_mapUnique, d_eleModulus, d_uniqueFree,
                                       d_uniqueFixed, stream=None):
    _call('assembleKs_level2_superEle_inplace', [d_Ks, d_upMap, d_Ks0, d_Psub, d_mapUnique, d_eleModulus,
                                                 d_uniqueFree, d_uniqueFixed], stream=stream)
