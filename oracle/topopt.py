"""Compliance, sensitivities, PDE Helmholtz filter and the optimality-criteria driver (oracle).

Literal FP64 restatement of TOP_GPU.m:1-292 (driver), 580-600 (unit compliance),
880-1031 (PDE filter), 2622-2697 (super-element compliance / sensitivities / index maps)
and the imfilter density filter of super-element mode (TOP_GPU.m:112-121, 220-227, 262-265).
Test infrastructure only (see oracle/__init__.py).
"""
import math
import time

import numpy as np
import scipy.ndimage

from . import constants as C
from . import mesh as M
from . import solver as S


def unit_compliance(lv1, U, Ke):
    """TopOpti_ComputeUnitCompliance (TOP_GPU.m:580-600): ce_e = u_e' * Ke * u_e (unmodified Ke)."""
    uMat = S.gathering(np.asarray(U).reshape(lv1.numNodes, 3), lv1.eNodMat)
    return np.sum((uMat @ Ke) * uMat, axis=1)


def unit_compliance_super(lv1, U, K_mat):
    """TopOpti_ComputeUnitCompliance_Super (TOP_GPU.m:2622-2645): u_e' * K_mat_e * u_e."""
    uMat = S.gathering(np.asarray(U).reshape(lv1.numNodes, 3), lv1.eNodMat)
    KU = np.einsum('ije,ej->ei', K_mat.reshape(24, 24, -1, order='F'), uMat)
    return np.sum(KU * uMat, axis=1)


def pre_compute_der(lv1, U, I_mat):
    """pre_compute_der (TOP_GPU.m:2646-2671): dc(s, e) = u_e' * I_s * u_e, [64 x Ne]."""
    uMat = S.gathering(np.asarray(U).reshape(lv1.numNodes, 3), lv1.eNodMat)
    Is = I_mat.reshape(24, 24, -1, order='F')
    KU = np.einsum('ijs,ej->sei', Is, uMat)
    return np.sum(KU * uMat[None, :, :], axis=2)


class PDEFilter:
    """TopOpti_SetupPDEfilter_matrixFree / TopOpti_ConductPDEFiltering_matrixFree /
    MatTimesVec_matrixFree_B (TOP_GPU.m:880-1031), standard (non-super-element) mode."""

    def __init__(self, lv1, filter_radius, tol=1e-3, maxIT=800):
        self.lv = lv1
        self.kernel = C.pde_filter_kernel(filter_radius, lv1.eleSize[0])
        eNod = lv1.eNodMat.astype(np.int64)
        d = np.bincount((eNod - 1).reshape(-1), weights=np.tile(np.diag(self.kernel), lv1.numElements),
                        minlength=lv1.numNodes)
        self.diagPrecond = d ** (-1)
        self.tol = tol / 10
        self.maxIT = maxIT
        self.last_its = 0

    def matvec(self, u):
        eNod = self.lv.eNodMat.astype(np.int64) - 1
        sub = np.asarray(u)[eNod] @ self.kernel
        return np.bincount(eNod.reshape(-1), weights=sub.reshape(-1), minlength=self.lv.numNodes)

    def element_to_node(self, src):
        eNod = self.lv.eNodMat.astype(np.int64) - 1
        vals = np.asarray(src, dtype=np.float64).reshape(-1) * (1 / 8)
        return np.bincount(eNod.reshape(-1), weights=np.repeat(vals, 8), minlength=self.lv.numNodes)

    def node_to_element(self, tar):
        eNod = self.lv.eNodMat.astype(np.int64) - 1
        return np.sum(np.asarray(tar)[eNod], axis=1) * (1 / 8)

    def __call__(self, src):
        rhs = self.element_to_node(src)
        tar, its, _ = S.pcg(self.matvec, lambda x: self.diagPrecond * x, rhs, self.tol, self.maxIT)
        self.last_its = its
        return self.node_to_element(tar)


def gen_index_mapping(X, Y, Z, n):
    """gen_index_mapping (TOP_GPU.m:2673-2697): sub-voxel grid <-> element-major order (1-based)."""
    nX, nY, nZ = n * X, n * Y, n * Z
    N = nX * nY * nZ
    i_old = np.arange(N, dtype=np.int64)
    iy = i_old % nY
    ix = (i_old // nY) % nX
    iz = i_old // (nY * nX)
    big = iy // n + Y * (ix // n + X * (iz // n))
    loc = iy % n + n * (ix % n + n * (iz % n))
    old2new = (loc + n ** 3 * big + 1).astype(np.int32)
    new2old = np.zeros(N, dtype=np.int32)
    new2old[old2new - 1] = np.arange(1, N + 1, dtype=np.int32)
    return old2new, new2old


class ConvolutionFilter:
    """Super-element density filter: imfilter with a linear cone on the num x sub-voxel grid
    (TOP_GPU.m:112-121, 220-227, 262-265), zero boundary (bcF = 0)."""

    def __init__(self, nelx, nely, nelz, rMin, num=4):
        self.num = num
        self.r = num * rMin
        c = math.ceil(self.r)
        v = np.arange(-c + 1, c)
        dy, dx, dz = np.meshgrid(v, v, v)
        self.h = np.maximum(0.0, self.r - np.sqrt(dx ** 2 + dy ** 2 + dz ** 2))
        self.shape = (num * nely, num * nelx, num * nelz)
        self.Hs = self._imfilter(np.ones(self.shape))
        self.old2new, self.new2old = gen_index_mapping(nelx, nely, nelz, num)

    def _imfilter(self, A):
        return scipy.ndimage.correlate(A, self.h, mode='constant', cval=0.0)

    def _to_grid(self, v):
        return np.asarray(v).reshape(-1, order='F')[self.old2new - 1].reshape(self.shape, order='F')

    def _from_grid(self, g):
        return g.reshape(-1, order='F')[self.new2old - 1].reshape(self.num ** 3, -1, order='F')

    def adjoint(self, dc):
        """dc <- imfilter(dc ./ Hs, h) in element-major [64 x Ne] layout."""
        return self._from_grid(self._imfilter(self._to_grid(dc) / self.Hs))

    def forward(self, x):
        """xPhys <- imfilter(x, h) ./ Hs."""
        return self._from_grid(self._imfilter(self._to_grid(x)) / self.Hs)


def run_top(model, V0=0.12, rMin=math.sqrt(3), nLoop=50, ft=1, Emin=1e-6, E0=1.0, penal=3.0,
            tol=1e-3, maxIT=800, verbose=False, x_init=None, weight_factor_jacobi=0.6, passive=None):
    """Optimization driver TOP_GPU (TOP_GPU.m:93-292), FP64.

    passive: 0-based solid element ids held at density 1 (passiveElements_, TOP_GPU.m:123-125).  As in
    the reference they are set inside every bisection step (so they enter gt, TOP_GPU.m:239, 247) and
    again after the density filter (:270, :273); in super-element mode only sub-voxel row Num_ = 4
    of each passive element is set (xnew(Num_, passiveElements), TOP_GPU.m:247, 273).

    Returns dict with cHist, volHist, sharpHist, consHist, lssIts (every solve, load cases in turn),
    cSolid, xPhys, U (the first load case's solution), Us (every case's), times.
    """
    pas = np.asarray(passive if passive is not None else [], dtype=np.int64)
    sv = S.Solver(model, weight_factor_jacobi=weight_factor_jacobi)
    lv1 = model.levels[0]
    Ne = lv1.numElements
    se = model.super_element
    num = 4
    if not se:
        filt = PDEFilter(lv1, rMin, tol, maxIT)
        x = np.full(Ne, V0) if x_init is None else np.array(x_init, dtype=np.float64)
    else:
        filt = ConvolutionFilter(model.nelx, model.nely, model.nelz, rMin, num)
        x = np.full((num ** 3, Ne), V0) if x_init is None else np.array(x_init, dtype=np.float64)
    xPhys = x.copy()
    # one PCG solve per load case (TOP_GPU.m:155, 186), their unit compliances weighted equally
    # (objWeightingList_, TOP_GPU.m:461; ceList * objWeightingList_, :599, 2644, 2671)
    F = model.F
    nC = F.shape[1]

    def weigh(parts):
        if nC == 1:
            return parts[0]
        out = np.zeros_like(parts[0])
        for p in parts:
            out = out + (1.0 / nC) * p
        return out

    # fully solid reference (TOP_GPU.m:142-166)
    E1 = np.full((num ** 3, Ne), E0) if se else np.full(Ne, E0)
    sv.assemble(E1)
    Us = [sv.pcg(F[:, k], tol, maxIT)[0] for k in range(nC)]
    if not se:
        ce = weigh([unit_compliance(lv1, u, lv1.Ke) for u in Us])
        cSolid = float(E1 @ ce)
    else:
        ce = weigh([unit_compliance_super(lv1, u, sv.K_mat) for u in Us])
        cSolid = float(np.sum(ce))

    def simp(xp):
        return Emin + xp ** penal * (E0 - Emin)

    def desimp(xp):
        return penal * (E0 - Emin) * xp ** (penal - 1)

    out = dict(cHist=[], volHist=[], sharpHist=[], consHist=[], lssIts=[], changeHist=[], cSolid=cSolid,
               tHist=[], resTail=[], ocMargin=[])
    loop, change, sharp = 0, 1.0, 1.0
    while loop < nLoop and change > 0.0001 and sharp > 0.01:
        t0 = time.perf_counter()
        loop += 1
        E = simp(xPhys)
        sv.assemble(E)
        t1 = time.perf_counter()
        for k in range(nC):  # each case warm-started from its previous solution
            Us[k], its, res = sv.pcg(F[:, k], tol, maxIT, y0=Us[k])
            out['lssIts'].append(its)
            out['resTail'].append([float(v) for v in res[-2:]])   # relative residuals deciding the stop
        t2 = time.perf_counter()
        if not se:
            ce = weigh([unit_compliance(lv1, u, lv1.Ke) for u in Us])
            ceNorm = ce / cSolid
            cObj = float(E @ ceNorm)
            cDesign = cObj * cSolid
            V = float(np.sum(xPhys) / Ne)
            dc = -desimp(xPhys) * ceNorm
            dv = np.ones(Ne)
        else:
            ce = weigh([unit_compliance_super(lv1, u, sv.K_mat) for u in Us])
            ceNorm = ce / cSolid
            cObj = float(np.sum(ceNorm))
            cDesign = cObj * cSolid
            V = float(np.sum(xPhys) / (Ne * num ** 3))
            dc = weigh([pre_compute_der(lv1, u, model.I_mat) for u in Us]) / cSolid
            dc = -desimp(xPhys) * dc
            dv = np.ones((num ** 3, Ne))
        t3 = time.perf_counter()
        if ft == 1:
            dc = filt(x.reshape(-1) * dc.reshape(-1)) / np.maximum(1e-3, x.reshape(-1))
        elif ft == 2:
            if not se:
                dc = filt(dc)
                dv = filt(dv)
            else:
                dc = filt.adjoint(dc)
                dv = filt.adjoint(dv)
        t4 = time.perf_counter()
        fval = float(np.mean(xPhys)) - V0
        l1, l2, move = 0.0, 1e9, 0.2
        margin = math.inf   # smallest |fval + mean(xnew - x)| the bisection decided on
        while (l2 - l1) / (l1 + l2) > 1e-6:
            lmid = 0.5 * (l2 + l1)
            xnew = np.maximum(0, np.maximum(x - move, np.minimum(1, np.minimum(x + move, x * np.sqrt(-dc / dv / lmid)))))
            if pas.size:
                if not se:
                    xnew[pas] = 1.0
                else:
                    xnew[num - 1, pas] = 1.0
            gt = fval + float(np.mean(xnew - x))
            margin = min(margin, abs(gt))
            if gt > 0:
                l1 = lmid
            else:
                l2 = lmid
        change = float(np.max(np.abs(xnew - x)))
        x = xnew
        t5 = time.perf_counter()
        if ft == 1:
            xPhys = xnew
        elif ft == 2:
            xPhys = filt(xnew) if not se else filt.forward(xnew)
        if pas.size:
            xPhys = xPhys.copy()
            if not se:
                xPhys[pas] = 1.0
            else:
                xPhys[num - 1, pas] = 1.0
        if not se:
            sharp = 4 * float(np.sum(xPhys * (1 - xPhys))) / Ne
        else:
            sharp = 4 * float(np.sum(xPhys * (1 - xPhys))) / (Ne * num ** 3)
        t6 = time.perf_counter()
        out['cHist'].append(cDesign)
        out['volHist'].append(V)
        out['consHist'].append(fval)
        out['ocMargin'].append(margin)
        out['sharpHist'].append(sharp)
        out['changeHist'].append(change)
        out['tHist'].append([t1 - t0, t2 - t1, (t3 - t2) + (t5 - t4), (t4 - t3) + (t6 - t5), t6 - t0])
        if verbose:
            print(' It.:%4i Obj.:%16.8e Vol.:%6.4e Sharp.:%6.4e Cons.:%4.2e Ch.:%4.2e its %d'
                  % (loop, cDesign, V, sharp, fval, change, its))
    out['xPhys'] = xPhys
    out['x'] = x
    out['U'] = Us[0]
    out['Us'] = Us
    return out
