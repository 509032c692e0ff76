"""Host mirror of TOP_GPU.m: the SIMP compliance optimisation driver (TOP_GPU.m:1-292).

Every numerical step runs in libtopgpu's HIP kernels: assembly, MG-PCG, unit compliance,
sensitivities, PDE / convolution filters, OC update and the reductions.  torch only provides
the device buffers.  The call sequence and histories follow TOP_GPU.m line by line.

    out = TOP_GPU(np.ones((24, 48, 24), bool), optCase=1, V0=0.12)
"""
import ctypes
import math
import os
import time

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .model import Model
from .solver import Solver


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class PDEFilter:
    """TopOpti_SetupPDEfilter_matrixFree / TopOpti_ConductPDEFiltering_matrixFree (TOP_GPU.m:880-1006)."""

    def __init__(self, model, rMin, tol=1e-3 / 10, max_it=800):
        h = ctypes.c_void_p()
        check(lib().tg_filter_create(model.h, float(rMin), _stream(), ctypes.byref(h)))
        self.h = h
        self.model = model
        self.tol = tol
        self.max_it = max_it
        self.last_its = 0

    def __del__(self):
        h = getattr(self, 'h', None)
        if h:
            self.h = None
            try:
                lib().tg_filter_destroy(h)
            except Exception:   # interpreter shutdown: module globals already torn down
                pass

    def __call__(self, src):
        src = src.contiguous()
        out = torch.empty_like(src)
        its = ctypes.c_int32()
        check(lib().tg_filter_apply(self.h, _p(src), _p(out), float(self.tol), int(self.max_it), ctypes.byref(its)))
        self.last_its = its.value
        return out


class ConvolutionFilter:
    """imfilter density filter of super-element mode (TOP_GPU.m:112-121, 220-227, 262-265)."""

    def __init__(self, nelx, nely, nelz, rMin_subvoxels, num=4):
        h = ctypes.c_void_p()
        check(lib().tg_convfilter_create(nelx, nely, nelz, float(rMin_subvoxels), num, _stream(), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, 'h', None)
        if h:
            self.h = None
            try:
                lib().tg_convfilter_destroy(h)
            except Exception:   # interpreter shutdown: module globals already torn down
                pass

    def forward(self, x):
        out = torch.empty_like(x)
        check(lib().tg_convfilter_forward(self.h, _p(x.contiguous()), _p(out)))
        return out

    def adjoint(self, dc):
        out = torch.empty_like(dc)
        check(lib().tg_convfilter_adjoint(self.h, _p(dc.contiguous()), _p(out)))
        return out


def reduce(op, a, b=None):
    out = ctypes.c_double()
    check(lib().tg_reduce(op, _p(a), _p(b) if b is not None else None, a.numel(), ctypes.byref(out), _stream()))
    return out.value


def simp(x, Emin, E0, penal):
    E = torch.empty_like(x)
    check(lib().tg_simp_modulus(_p(x), x.numel(), Emin, E0, penal, _p(E), _stream()))
    return E


def sensitivity(x, ce, cSolid, Emin, E0, penal):
    dc = torch.empty_like(x)
    per = x.numel() // ce.numel()
    check(lib().tg_sensitivity(_p(x), _p(ce), x.numel(), per, Emin, E0, penal, cSolid, _p(dc), _stream()))
    return dc


def elementwise(op, a, b, scalar=0.0):
    out = torch.empty_like(a)
    check(lib().tg_elementwise(op, _p(a), _p(b), float(scalar), a.numel(), _p(out), _stream()))
    return out


def oc_update(x, dc, dv, fval, move=0.2, passive=None):
    """passive: None or a float64 indicator like x (nonzero: xnew held at 1, TOP_GPU.m:239, 247)."""
    xnew = torch.empty_like(x)
    change = ctypes.c_double()
    steps = ctypes.c_int32()
    check(lib().tg_oc_update(_p(x), _p(dc), _p(dv), _p(passive) if passive is not None else None, x.numel(),
                             float(fval), float(move), _p(xnew), ctypes.byref(change), ctypes.byref(steps),
                             _stream()))
    return xnew, change.value


def passive_indicator(model, shape, super_element, num=4):
    """Indicator of the passive entries of the design array (None when the model has none): element e
    itself, or in super-element mode only sub-voxel row Num_ = 4 of element e (the reference's
    xnew(Num_, passiveElements) and xPhys(Num_, passiveElements), TOP_GPU.m:247, 273)."""
    pas = np.asarray(getattr(model, 'passive', ()), dtype=np.int64)
    if pas.size == 0:
        return None
    ind = np.zeros(int(np.prod(shape)))
    if super_element:
        ind[num ** 3 * pas + (num - 1)] = 1.0
    else:
        ind[pas] = 1.0
    return torch.from_numpy(ind).to('cuda')


def unit_compliance(solver, U):
    ce = torch.empty(solver.levels[0]['numElements'], dtype=torch.float64, device='cuda')
    check(lib().tg_solver_unit_compliance(solver.h, _p(U), _p(ce)))
    return ce


def subvoxel_energy(solver, U):
    dc = torch.empty(64 * solver.levels[0]['numElements'], dtype=torch.float64, device='cuda')
    check(lib().tg_solver_subvoxel_energy(solver.h, _p(U), _p(dc)))
    return dc


def weigh_cases(parts):
    """sum_k parts[k] / n over the load cases (ceList * objWeightingList_ with objWeightingList_ =
    ones / n, TOP_GPU.m:461, 599, 2644, 2671), on the device; a single case as it is."""
    if len(parts) == 1:
        return parts[0]
    out = torch.zeros_like(parts[0])
    for p in parts:
        out = elementwise(_lib.TG_EW_AXPY, out, p, 1.0 / len(parts))
    return out


def TOP_GPU(inputModel, optCase=1, V0=0.12, rMin=math.sqrt(3), nLoop=50, ft=1, mixed_Precision=0,
            super_element=0, *, Emin=1e-6, E0=1.0, penal=3.0, nu=0.3, tol=1e-3, maxIT=800, x0=None,
            out_path=None, verbose=False, model=None, tuning=None, weightFactorJacobi=0.6):
    """TOP_GPU(inputModel, 'optCase', ..., 'V0', ..., 'rMin', ..., 'nLoop', ..., 'ft', ...,
    'mixed_Precision', ..., 'super_element', ...)  (TOP_GPU.m:1-292).

    inputModel: logical voxel array [nely, nelx, nelz] or a .TopVoxel path.  weightFactorJacobi is
    the V-cycle's damped-Jacobi weight (the reference's global of that name, 0.6); ragged domains
    need a smaller one (PCG stalls when it exceeds 2 / lambda_max(D^-1 K)).  Returns a dict with
    the histories (cHist, volHist, sharpHist, consHist, lssIts, tHist), cSolid and xPhys."""
    num = 4
    se = bool(super_element)
    if model is None:
        if isinstance(inputModel, (str, os.PathLike)):
            from .topvoxel import load_topvoxel_model
            model = load_topvoxel_model(inputModel, super_element=se, nu=nu,
                                        precision='mixed' if mixed_Precision else 'fp64')
        else:
            vox = np.asarray(inputModel, dtype=bool)
            nely_, nelx_, nelz_ = vox.shape
            if nelx_ < 3 or nely_ < 3 or nelz_ < 3:
                raise ValueError('Inappropriate Input Model!')
            model = Model(vox, cell_size=1.0 / nelx_, poisson_ratio=nu, youngs_modulus=E0, super_element=se,
                          precision='mixed' if mixed_Precision else 'fp64')
            model.builtin_case(optCase)
            model.finalize()
    if se and ft != 2:
        raise ValueError('super-element mode requires ft = 2 (README: ft=2 is required)')
    solver = Solver(model)
    if tuning:
        solver.set_tuning(**tuning)
    if weightFactorJacobi != 0.6:
        solver.set_jacobi_weight(weightFactorJacobi)
    Ne = model.num_elements
    dev = 'cuda'
    if se:
        if Ne != model.nely * model.nelx * model.nelz:
            raise ValueError('super-element mode supports box-shaped domains only')
        filt = ConvolutionFilter(model.nelx, model.nely, model.nelz, num * rMin, num)
        shape = (num ** 3 * Ne,)
    else:
        filt = PDEFilter(model, rMin, tol=tol / 10, max_it=maxIT)
        shape = (Ne,)
    # one PCG solve per load case (TOP_GPU.m:155, 186)
    Fs = [torch.from_numpy(np.ascontiguousarray(col)).to(dev) for col in model.get_table(1, 'F').T]
    if x0 is None:
        x = torch.full(shape, float(V0), dtype=torch.float64, device=dev)
    else:
        x = torch.as_tensor(np.asarray(x0, dtype=np.float64).reshape(-1, order='F'), device=dev).clone()
    xPhys = x.clone()
    nel = x.numel()
    # 4. fully solid domain (TOP_GPU.m:142-166)
    E1 = torch.full(shape, E0, dtype=torch.float64, device=dev)
    solver.set_modulus(E1)
    solver.assemble()
    Us = [solver.pcg(F, tol=tol, max_it=maxIT)[0] for F in Fs]
    ce = weigh_cases([unit_compliance(solver, U) for U in Us])
    cSolid = reduce(_lib.TG_RED_DOT, E1, ce) if not se else reduce(_lib.TG_RED_SUM, ce)
    out = dict(cHist=[], volHist=[], sharpHist=[], consHist=[], lssIts=[], tHist=[], changeHist=[],
               cSolid=cSolid, pcg_ms=[], filter_its=[])
    passive = passive_indicator(model, shape, se, num)
    loop, change, sharp = 0, 1.0, 1.0
    ones = torch.ones(shape, dtype=torch.float64, device=dev)
    while loop < nLoop and change > 0.0001 and sharp > 0.01:
        t0 = time.perf_counter()
        loop += 1
        E = simp(xPhys, Emin, E0, penal)
        solver.set_modulus(E)
        solver.assemble()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for k, F in enumerate(Fs):  # each case warm-started from its previous solution
            Us[k], its, _ = solver.pcg(F, y0=Us[k], tol=tol, max_it=maxIT)
            out['lssIts'].append(its)
            out['pcg_ms'].append(solver.last_pcg_ms())
        t2 = time.perf_counter()
        ce = weigh_cases([unit_compliance(solver, U) for U in Us])
        if not se:
            cObj = reduce(_lib.TG_RED_DOT, E, ce) / cSolid
            cDesign = cObj * cSolid
            V = reduce(_lib.TG_RED_SUM, xPhys) / Ne
            dc = sensitivity(xPhys, ce, cSolid, Emin, E0, penal)
            dv = ones
        else:
            cObj = reduce(_lib.TG_RED_SUM, ce) / cSolid
            cDesign = cObj * cSolid
            V = reduce(_lib.TG_RED_SUM, xPhys) / (Ne * num ** 3)
            dc = sensitivity(xPhys, weigh_cases([subvoxel_energy(solver, U) for U in Us]), cSolid, Emin, E0, penal)
            dv = ones
        t3 = time.perf_counter()
        fits = []
        if ft == 1:
            dc = elementwise(_lib.TG_EW_DIV_FLOOR, filt(elementwise(_lib.TG_EW_MUL, x, dc)), x, 1e-3)
            fits.append(filt.last_its)
        elif ft == 2:
            if not se:
                dc = filt(dc)
                fits.append(filt.last_its)
                dv = filt(dv)
                fits.append(filt.last_its)
            else:
                dc = filt.adjoint(dc)
                dv = filt.adjoint(dv)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        fval = reduce(_lib.TG_RED_SUM, xPhys) / nel - V0
        xnew, change = oc_update(x, dc, dv, fval, 0.2, passive)
        x = xnew
        t5 = time.perf_counter()
        if ft == 1:
            xPhys = xnew
        elif ft == 2:
            if not se:
                xPhys = filt(xnew)
                fits.append(filt.last_its)
            else:
                xPhys = filt.forward(xnew)
        if passive is not None:       # xPhys(passiveElements) = 1 (TOP_GPU.m:270, 273)
            xPhys = elementwise(_lib.TG_EW_SET_WHERE, xPhys, passive, 1.0)
        sharp = 4 * reduce(_lib.TG_RED_SHARPNESS, xPhys) / nel
        t6 = time.perf_counter()
        out['cHist'].append(cDesign)
        out['volHist'].append(V)
        out['consHist'].append(fval)
        out['sharpHist'].append(sharp)
        out['changeHist'].append(change)
        out['filter_its'].append(fits)
        out['tHist'].append([t1 - t0, t2 - t1, (t3 - t2) + (t5 - t4), (t4 - t3) + (t6 - t5), t6 - t0])
        if verbose:
            print(' It.:%4i Obj.:%16.8e Vol.:%6.4e Sharp.:%6.4e Cons.:%4.2e Ch.:%4.2e PCG its %d'
                  % (loop, cDesign, V, sharp, fval, change, its), flush=True)
    out['xPhys'] = xPhys
    out['x'] = x
    out['U'] = Us[0]
    out['Us'] = Us
    out['model'] = model
    if out_path:
        from .io import write_histories
        write_histories(out_path, out)
    return out
