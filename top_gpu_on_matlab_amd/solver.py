"""Device solver (libtopgpu tg_solver_*): assembly, matrix-free K*u, V-cycle, PCG.

Vectors are torch CUDA float64 tensors in the reference's layout (DOF-interleaved, solid-node
numbering).  torch only supplies device memory and the stream; all arithmetic runs in the
library's HIP kernels.
"""
import ctypes

import torch

from ._lib import check, lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _dev(x, n=None):
    t = torch.as_tensor(x, dtype=torch.float64, device='cuda').contiguous()
    if n is not None and t.numel() != n:
        raise ValueError('expected %d values, got %d' % (n, t.numel()))
    return t


class Solver:
    """Device-resident counterpart of the solver group of TOP_GPU.m (TOP_GPU.m:602-1267)."""

    def __init__(self, model, stream=None):
        if not torch.cuda.is_available():
            raise RuntimeError('tg_solver needs a HIP device')
        self.model = model            # the model must outlive the solver
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        check(lib().tg_solver_create(model.h, _stream_handle(self.stream), ctypes.byref(h)))
        self.h = h
        self.levels = [model.level_info(l) for l in range(1, model.num_levels + 1)]
        self.num_levels = len(self.levels)

    def __del__(self):
        h = getattr(self, 'h', None)
        if h:
            self.h = None
            try:
                lib().tg_solver_destroy(h)
            except Exception:   # interpreter shutdown: module globals already torn down
                pass

    def ndof(self, level=1):
        return 3 * self.levels[level - 1]['numNodes']

    def set_modulus(self, E):
        n = self.levels[0]['numElements'] * (64 if self.model.super_element else 1)
        self._E = _dev(E, n)
        check(lib().tg_solver_set_modulus(self.h, _ptr(self._E)))

    def assemble(self):
        check(lib().tg_solver_assemble(self.h))

    def kbyu(self, u, level=1):
        u = _dev(u, self.ndof(level))
        y = torch.empty_like(u)
        check(lib().tg_solver_kbyu(self.h, int(level), _ptr(u), _ptr(y)))
        return y

    def vcycle(self, r):
        r = _dev(r, self.ndof(1))
        z = torch.empty_like(r)
        check(lib().tg_solver_vcycle(self.h, _ptr(r), _ptr(z)))
        return z

    def restrict(self, rf, coarse_level):
        rf = _dev(rf, self.ndof(coarse_level - 1))
        rc = torch.empty(self.ndof(coarse_level), dtype=torch.float64, device='cuda')
        check(lib().tg_solver_restrict(self.h, int(coarse_level), _ptr(rf), _ptr(rc)))
        return rc

    def interpolate(self, xc, coarse_level):
        xc = _dev(xc, self.ndof(coarse_level))
        xf = torch.empty(self.ndof(coarse_level - 1), dtype=torch.float64, device='cuda')
        check(lib().tg_solver_interpolate(self.h, int(coarse_level), _ptr(xc), _ptr(xf)))
        return xf

    def coarse_solve(self, r):
        r = _dev(r, self.ndof(self.num_levels))
        x = torch.empty_like(r)
        check(lib().tg_solver_coarse_solve(self.h, _ptr(r), _ptr(x)))
        return x

    def pcg(self, b, y0=None, tol=1e-3, max_it=800):
        """Solving_PCG(@Solving_KbyU_MatrixFree, @Solving_Vcycle, b, tol, maxIT[, y0]).

        Returns (y, its, resnorm)."""
        b = _dev(b, self.ndof(1))
        y = torch.zeros_like(b) if y0 is None else _dev(y0, self.ndof(1)).clone()
        its = ctypes.c_int32()
        res = ctypes.c_double()
        check(lib().tg_solver_pcg(self.h, _ptr(b), _ptr(y), float(tol), int(max_it), 0 if y0 is None else 1,
                                  ctypes.byref(its), ctypes.byref(res)))
        return y, its.value, res.value

    def diag(self, level):
        d = torch.empty(self.ndof(level), dtype=torch.float64, device='cuda')
        check(lib().tg_solver_get_diag(self.h, int(level), _ptr(d)))
        return d

    def ks(self, level):
        Ne = self.levels[level - 1]['numElements']
        k = torch.empty(576 * Ne, dtype=torch.float64, device='cuda')
        check(lib().tg_solver_get_ks(self.h, int(level), _ptr(k)))
        return k.view(Ne, 24, 24).transpose(1, 2)  # [e][row][col] view of the column-major pages

    def time_kbyu(self, level=1, precision='fp64', reps=50):
        """Average K*u kernel duration (ms) at `level`, HIP events on the solver stream."""
        ms = ctypes.c_double()
        check(lib().tg_solver_time_kbyu(self.h, int(level), 0 if precision == 'fp64' else 1, int(reps),
                                        ctypes.byref(ms)))
        return ms.value

    def set_tuning(self, kfine_kernel=None, kfine_zchunk=None, kelem_parts=None, coarse_graph=None, kfine_mfma=None,
                   coarse_persistent=None, coarse_streaming=None, coarse_polls=None):
        """Finest-level K*u launch choice: kernel 'tiled' (default) or 'gather'; z planes per tile;
        threads per element of the FP64 per-element-matrix K*u (1, the default: the packed upper triangles,
        one thread per element; 4, 6, 8, 12 or 24: the full matrices split over that many threads); coarsest solve as one persistent launch
        (True, default; coarse_streaming=True, the default, reads its factors with non-temporal loads) or step by step,
        then either replayed as a captured HIP graph (coarse_graph=True)
        or launched kernel by kernel; the mixed V-cycle's finest element products: 'walsh' (the Walsh
        form on the VALU, default), 'mfma32' (dense, MFMA with FP32 operands), 'bf16' (dense, MFMA with
        BF16 operands and FP32 accumulation) or 'bf16x3' (as 'bf16' with each operand split into two BF16
        parts, three products)."""
        from ._lib import (TG_TUNE_COARSE_GRAPH, TG_TUNE_COARSE_PERSISTENT, TG_TUNE_COARSE_STREAMING,
                           TG_TUNE_KELEM_PARTS, TG_TUNE_KFINE_KERNEL, TG_TUNE_KFINE_MFMA, TG_TUNE_KFINE_ZCHUNK)
        if coarse_polls is not None:
            from ._lib import TG_TUNE_COARSE_POLLS
            check(lib().tg_solver_set_tuning(self.h, TG_TUNE_COARSE_POLLS, int(coarse_polls)))
        if coarse_streaming is not None:
            check(lib().tg_solver_set_tuning(self.h, TG_TUNE_COARSE_STREAMING, int(bool(coarse_streaming))))
        if coarse_persistent is not None:
            check(lib().tg_solver_set_tuning(self.h, TG_TUNE_COARSE_PERSISTENT, int(bool(coarse_persistent))))
        if kfine_mfma is not None:
            mode = ({'walsh': 0, 'mfma32': 1, 'bf16': 2, 'bf16x3': 3}[kfine_mfma] if isinstance(kfine_mfma, str)
                    else int(kfine_mfma))
            check(lib().tg_solver_set_tuning(self.h, TG_TUNE_KFINE_MFMA, mode))
        if coarse_graph is not None:
            check(lib().tg_solver_set_tuning(self.h, TG_TUNE_COARSE_GRAPH, int(bool(coarse_graph))))
        if kelem_parts is not None:
            check(lib().tg_solver_set_tuning(self.h, TG_TUNE_KELEM_PARTS, int(kelem_parts)))
        if kfine_kernel is not None:
            check(lib().tg_solver_set_tuning(self.h, TG_TUNE_KFINE_KERNEL, {'tiled': 0, 'gather': 1}[kfine_kernel]))
        if kfine_zchunk is not None:
            check(lib().tg_solver_set_tuning(self.h, TG_TUNE_KFINE_ZCHUNK, int(kfine_zchunk)))

    def set_jacobi_weight(self, omega):
        """Damped-Jacobi weight of the V-cycle smoother (weightFactorJacobi_, TOP_GPU.m:397; 0.6)."""
        check(lib().tg_solver_set_jacobi_weight(self.h, float(omega)))

    def time_kfine(self, on=True, only_operator=False):
        """Start (clearing the record) or stop bracketing each finest-level K*u launch with HIP events
        (only_operator: just the plain products K*u, i.e. the PCG operator, not the V-cycle's)."""
        from ._lib import TG_TUNE_TIME_KFINE
        check(lib().tg_solver_set_tuning(self.h, TG_TUNE_TIME_KFINE, (2 if only_operator else 1) if on else 0))

    def kfine_timing(self, precision='fp64', epilogue='ku'):
        """(launches, total ms) of the timed finest K*u launches; epilogue 'ku', 'resid', 'jacobi' or
        'resid_jac0' (residual with the pre-smoothing fused in)."""
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        check(lib().tg_solver_kfine_timing(self.h, 0 if precision == 'fp64' else 1,
                                           {'ku': 0, 'resid': 1, 'jacobi': 2, 'resid_jac0': 3}[epilogue], ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def last_pcg_ms(self):
        ms = ctypes.c_double()
        check(lib().tg_solver_last_timing(self.h, ctypes.byref(ms)))
        return ms.value
