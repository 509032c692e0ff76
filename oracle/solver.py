"""Solver restatement: MEX kernels, matrix-free K*u, V-cycle, PCG, Galerkin assembly (oracle).

Literal FP64 numpy restatement of TOP_GPU.m's solver group and of the six CUDA MEX
kernels.  Accumulations that the reference performs with atomics are done here in a fixed
serial order (numpy bincount / add.at), so the oracle is deterministic.
Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np
import scipy.linalg
import scipy.sparse
import scipy.sparse.linalg

from . import constants as C


# ---------------------------------------------------------------------------------------
# The MEX kernels
# ---------------------------------------------------------------------------------------

def gathering(u, eNodMat):
    """Gathering_inplace (Gathering_inplace.cu:22-54): uMat(e, 3n+c) = u(eNod(e,n), c).

    u: [Nn x 3]; eNodMat: [Ne x 8] 1-based.  Returns uMat [Ne x 24].
    """
    idx = np.asarray(eNodMat, dtype=np.int64) - 1
    return np.asarray(u)[idx, :].reshape(idx.shape[0], 24)


def scattering(Y, eNodMat, uMat):
    """Scattering_inplace (Scattering_inplace.cu:45-78): Y(eNod(e,j), c) += uMat(e, c+3j).

    Node ids outside [1, Nn] are skipped (Scattering_inplace.cu:70).  Y [Nn x 3] is updated
    in place and returned.
    """
    Nn = Y.shape[0]
    nid = np.asarray(eNodMat, dtype=np.int64).reshape(-1) - 1          # e-major, j fastest
    vals = np.asarray(uMat).reshape(-1, 8, 3)                           # (e, j, c)
    ok = (nid >= 0) & (nid < Nn)
    for c in range(3):
        Y[:, c] += np.bincount(nid[ok], weights=vals[:, :, c].reshape(-1)[ok], minlength=Nn)
    return Y


def scatter_accum3(keys, Tmp, xF):
    """scatter_accum3_inplace (scatter_accum3_inplace.cu:32-56): xF(key_i, c) += Tmp(i, c).

    keys: [N] 1-based int; Tmp: [N x 3]; xF: [M x 3], updated in place; out-of-range keys skipped.
    """
    M = xF.shape[0]
    k = np.asarray(keys, dtype=np.int64).reshape(-1) - 1
    ok = (k >= 0) & (k < M)
    for c in range(3):
        xF[:, c] += np.bincount(k[ok], weights=np.asarray(Tmp)[ok, c], minlength=M)
    return xF


def build_pch(span):
    """Pch(:,:,s) = interpolatingKe(eDofMat(s,:), :) (assembleKs_level2_gpu.m:10-14,
    TOP_GPU.m:1056-1060).  Returns [24 x 24 x span^3]."""
    iKe = C.restriction_interpolation_operator(span, in_dof=True)
    sub = C.sub_element_node_table(span)
    eDof = np.concatenate([3 * sub - 2, 3 * sub - 1, 3 * sub], axis=1)[:, C.RE_ORDERING - 1]
    P = np.zeros((24, 24, sub.shape[0]))
    for s in range(sub.shape[0]):
        P[:, :, s] = iKe[eDof[s] - 1, :]
    return P


def assemble_ks_level2(upMap, eleModulus, K0col, Pch, mapUnique=None, uniqueFree=None, uniqueFixed=None):
    """assembleKs_level2_inplace (assembleKs_level2_inplace.cu:58-160).

    Ks_e = sum_s Pch_s' * Kch_s * Pch_s, Kch = E_f*K0 or E_f*Ufree[:,u] + Ufixed[:,u] when
    mapUnique(f) = u > 0; children with upMap <= 0 are skipped.  Returns [24 x 24 x Ne].
    """
    Ne, nCh = upMap.shape
    K0 = np.asarray(K0col).reshape(24, 24, order='F')
    hasFixed = uniqueFree is not None and uniqueFree.size > 0 and mapUnique is not None
    Ks = np.zeros((24, 24, Ne))
    for e in range(Ne):
        Kout = np.zeros((24, 24))
        for s in range(nCh):
            f = int(upMap[e, s])
            if f <= 0:
                continue
            mod = eleModulus[f - 1]
            uid = int(mapUnique[f - 1]) - 1 if hasFixed else -1
            if uid >= 0:
                Kch = (uniqueFree[:, uid] * mod + uniqueFixed[:, uid]).reshape(24, 24, order='F')
            else:
                Kch = K0 * mod
            B = Pch[:, :, s]
            Kout += B.T @ (Kch @ B)
        Ks[:, :, e] = Kout
    return Ks


def assemble_ks_higher_level(upMap, KsPrev, Psub):
    """assembleKs_higherLevel_inplace (assembleKs_higherLevel_inplace.cu:59-129):
    Ks_e = sum_s Psub_s' * KsPrev[:,:,upMap(e,s)] * Psub_s."""
    Ne, nSub = upMap.shape
    Ks = np.zeros((24, 24, Ne))
    for e in range(Ne):
        Kout = np.zeros((24, 24))
        for s in range(nSub):
            f = int(upMap[e, s])
            if f <= 0:
                continue
            B = Psub[:, :, s]
            Kout += B.T @ (KsPrev[:, :, f - 1] @ B)
        Ks[:, :, e] = Kout
    return Ks


def assemble_ks_level2_super(upMap, Ks0, Psub, mapUnique, eleModulus, uniqueFree, uniqueFixed):
    """assembleKs_level2_superEle_inplace (assembleKs_level2_superEle_inplace.cu:70-173).

    Ksub = Ks0[:, f], or sum_q Ufree[:, q, u]*E[q, f] + Ufixed[:, u] when mapUnique(f) = u > 0.
    uniqueFree: [576 x 64 x nU]; eleModulus: [64 x Nf]; Ks0: [576 x Nf].
    """
    Ne, nSub = upMap.shape
    hasFixed = uniqueFree is not None and uniqueFree.size > 0
    Ks = np.zeros((24, 24, Ne))
    for e in range(Ne):
        Kout = np.zeros((24, 24))
        for s in range(nSub):
            f = int(upMap[e, s])
            if f <= 0:
                continue
            uid = int(mapUnique[f - 1]) if hasFixed else 0
            if uid > 0:
                Ksub = uniqueFree[:, :, uid - 1] @ eleModulus[:, f - 1] + uniqueFixed[:, uid - 1]
            else:
                Ksub = Ks0[:, f - 1]
            Ksub = Ksub.reshape(24, 24, order='F')
            B = Psub[:, :, s]
            Kout += B.T @ (Ksub @ B)
        Ks[:, :, e] = Kout
    return Ks


def assemble_ks_level2_fast(upMap, eleModulus, K0col, Pch, mapUnique=None, uniqueFree=None, uniqueFixed=None):
    """Vectorised assemble_ks_level2 for large meshes: Ks_e = sum_s E_f(e,s) * (P_s' K0 P_s) plus
    the boundary children's corrections P_s' (E (Ufree - K0) + Ufixed) P_s.  Equal to the literal
    loop up to rounding (tests/test_oracle_solver.py pins it)."""
    Ne, nCh = upMap.shape
    K0 = np.asarray(K0col).reshape(24, 24, order='F')
    G = np.einsum('kis,kl,ljs->sij', Pch, K0, Pch)
    up = upMap.astype(np.int64)
    Ech = np.where(up > 0, eleModulus[np.maximum(up, 1) - 1], 0.0)
    Ks = np.ascontiguousarray((Ech @ G.reshape(nCh, 576)).T.reshape(24, 24, Ne))
    if uniqueFree is not None and uniqueFree.size > 0 and mapUnique is not None:
        es, ss = np.nonzero(up > 0)
        f = up[es, ss] - 1
        uid = mapUnique[f].astype(np.int64) - 1
        sel = uid >= 0
        for e, s, ff, u in zip(es[sel], ss[sel], f[sel], uid[sel]):
            D = ((uniqueFree[:, u] - K0.reshape(-1, order='F')) * eleModulus[ff] + uniqueFixed[:, u]).reshape(
                24, 24, order='F')
            B = Pch[:, :, s]
            Ks[:, :, e] += B.T @ D @ B
    return Ks


def assemble_ks_higher_level_fast(upMap, KsPrev, Psub):
    """Vectorised assemble_ks_higher_level (batched matmuls per child slot)."""
    Ne, nSub = upMap.shape
    out = np.zeros((Ne, 24, 24))
    Kp = np.moveaxis(KsPrev, 2, 0)
    for s in range(nSub):
        f = upMap[:, s].astype(np.int64)
        ok = f > 0
        B = Psub[:, :, s]
        out[ok] += B.T @ Kp[f[ok] - 1] @ B
    return np.moveaxis(out, 0, 2).copy()


# ---------------------------------------------------------------------------------------
# Solver state (the MATLAB globals the hot path reads)
# ---------------------------------------------------------------------------------------

class Solver:
    """Holds meshHierarchy_ + the BC element matrices and implements the solver group of
    TOP_GPU.m (Solving_KbyU_MatrixFree, Solving_Vcycle, Solving_RestrictResidual,
    Solving_InterpolationDeviation, Solving_PCG, Solving_AssembleFEAstencil)."""

    def __init__(self, model, weight_factor_jacobi=0.6, vcycle_type='Standard'):
        self.m = model
        self.levels = model.levels
        self.L = model.numLevels
        self.omega = weight_factor_jacobi
        self.vtype = vcycle_type
        self.K_mat = None
        self.coarse_solve = None
        self.stats = {}

    # --- K*u ---------------------------------------------------------------------------
    def kbyu(self, u, level=1):
        """Solving_KbyU_MatrixFree (TOP_GPU.m:669-765); level is 1-based."""
        lv = self.levels[level - 1]
        uVec = np.asarray(u).reshape(lv.numNodes, 3)
        uMat = gathering(uVec, lv.eNodMat)
        if level == 1:
            E = lv.eleModulus
            b = self.m.eleWithFixedDOFs.astype(np.int64) - 1
            Nb = b.size
            if not self.m.super_element:
                Kf = self.m.uniqueKesFree.reshape(24, 24, Nb, order='F')
                Kx = self.m.uniqueKesFixed.reshape(24, 24, Nb, order='F')
                Kc = Kf * E[b][None, None, :] + Kx
                Ufix = uMat[b, :]
                Yfix = np.einsum('ijb,bj->bi', Kc, Ufix)
                uMat = (uMat @ lv.Ks) * E[:, None]
                uMat[b, :] = Yfix
            else:
                Kf = self.m.uniqueKesFree                                  # [576 x 64 x Nb]
                Kc = np.einsum('iqb,qb->ib', Kf, E[:, b]).reshape(24, 24, Nb, order='F')
                Kc = Kc + self.m.uniqueKesFixed.reshape(24, 24, Nb, order='F')
                Ke3D = self.K_mat.reshape(24, 24, -1, order='F').copy()
                Ke3D[:, :, b] = Kc
                uMat = np.einsum('ije,ej->ei', Ke3D, uMat)
        else:
            uMat = np.einsum('ei,ije->ej', uMat, lv.Ks)
        Y = np.zeros((lv.numNodes, 3))
        scattering(Y, lv.eNodMat, uMat)
        return Y.reshape(-1)

    # --- transfer operators -------------------------------------------------------------
    def restrict(self, rFiner, ii):
        """Solving_RestrictResidual (TOP_GPU.m:817-847); ii = coarse level (1-based)."""
        lv, fine = self.levels[ii - 1], self.levels[ii - 2]
        rF = np.asarray(rFiner).reshape(fine.numNodes, 3)
        r1 = np.zeros((lv.intermediateNumNodes, 3))
        r1[lv.solidNodeMapCoarser2Finer.astype(np.int64) - 1, :] = rF
        r1 = r1 / lv.transferMatCoeffi[:, None]
        rC = np.zeros((lv.numNodes, 3))
        RI = lv.multiGridOperatorRI
        Ne = lv.numElements
        Fev = r1[lv.transferMat.reshape(-1, order='F').astype(np.int64) - 1, :]   # [(nInt*Ne) x 3]
        pages = Fev.reshape(RI.shape[0], Ne, 3, order='F')
        P3 = np.einsum('in,iec->nec', RI, pages)                                     # [8 x Ne x 3]
        P2 = P3.transpose(1, 0, 2).reshape(Ne * 8, 3, order='F')
        scatter_accum3(lv.eNodMat.reshape(-1, order='F'), P2, rC)
        return rC.reshape(-1)

    def interpolate(self, xCoarser, ii):
        """Solving_InterpolationDeviation (TOP_GPU.m:848-878); ii = coarse level (1-based)."""
        lv, fine = self.levels[ii - 1], self.levels[ii - 2]
        xC = np.asarray(xCoarser).reshape(lv.numNodes, 3)
        xF = np.zeros((lv.intermediateNumNodes, 3))
        RI = lv.multiGridOperatorRI
        Ne = lv.numElements
        Ev = xC[lv.eNodMat.reshape(-1, order='F').astype(np.int64) - 1, :]         # [(Ne*8) x 3]
        pages = Ev.reshape(Ne, 8, 3, order='F').transpose(1, 0, 2)                  # [8 x Ne x 3]
        Tmp = np.einsum('in,nec->iec', RI, pages)                                   # [nInt x Ne x 3]
        Tmp2 = Tmp.reshape(RI.shape[0] * Ne, 3, order='F')
        scatter_accum3(lv.transferMat.reshape(-1, order='F'), Tmp2, xF)
        xF = xF / lv.transferMatCoeffi[:, None]
        xF = xF[lv.solidNodeMapCoarser2Finer.astype(np.int64) - 1, :]
        return xF.reshape(-1)

    # --- V-cycle ------------------------------------------------------------------------
    def vcycle(self, r):
        """Solving_Vcycle (TOP_GPU.m:766-816), FP64."""
        L = self.L
        x = [None] * L
        rr = [None] * L
        rr[0] = np.asarray(r, dtype=np.float64)
        w = self.omega
        for ii in range(L - 1):
            x[ii] = w * (rr[ii] / self.levels[ii].diagK)
            if self.vtype == 'Adapted':
                rr[ii + 1] = self.restrict(rr[ii], ii + 2)
            else:
                d = rr[ii] - self.kbyu(x[ii], ii + 1)
                rr[ii + 1] = self.restrict(d, ii + 2)
        x[L - 1] = self.coarse_solve(rr[L - 1])
        for ii in range(L - 2, -1, -1):
            x[ii] = x[ii] + self.interpolate(x[ii + 1], ii + 2)
            if self.vtype == 'Adapted':
                x[ii] = x[ii] + w * rr[ii] / self.levels[ii].diagK
            else:
                x[ii] = x[ii] + w * (rr[ii] - self.kbyu(x[ii], ii + 1)) / self.levels[ii].diagK
        return x[0]

    # --- assembly -----------------------------------------------------------------------
    def assemble(self, eleModulus, fast=None):
        """Solving_AssembleFEAstencil (TOP_GPU.m:1033-1267) for the current moduli.

        eleModulus: [Ne] (standard) or [64 x Ne] (super-element mode).  fast selects the
        vectorised Galerkin products (default: on for meshes with more than 4096 coarse elements).
        """
        m = self.m
        lv1 = self.levels[0]
        lv1.eleModulus = np.asarray(eleModulus, dtype=np.float64)
        lv1.Ks = lv1.Ke
        for ii in range(2, self.L + 1):
            lv = self.levels[ii - 1]
            P = build_pch(lv.spanWidth)
            vec = fast if fast is not None else lv.numElements > 4096
            if ii == 2:
                if not m.super_element:
                    f = assemble_ks_level2_fast if vec else assemble_ks_level2
                    lv.Ks = f(lv.elementUpwardMap, lv1.eleModulus, lv1.Ke.reshape(-1, order='F'),
                              P, m.mapUniqueKes, m.uniqueKesFree, m.uniqueKesFixed)
                else:
                    self.K_mat = m.I_mat @ lv1.eleModulus
                    lv.Ks = assemble_ks_level2_super(lv.elementUpwardMap, self.K_mat, P, m.mapUniqueKes,
                                                     lv1.eleModulus, m.uniqueKesFree, m.uniqueKesFixed)
            else:
                f = assemble_ks_higher_level_fast if vec else assemble_ks_higher_level
                lv.Ks = f(lv.elementUpwardMap, self.levels[ii - 2].Ks, P)
            if ii < self.L:
                blk = lv.Ks.reshape(576, -1, order='F')[0::25, :].T                  # [Ne x 24]
                dk = np.zeros((lv.numNodes, 3))
                scattering(dk, lv.eNodMat, blk)
                lv.diagK = dk.reshape(-1)
        # finest Jacobi diagonal (TOP_GPU.m:1168-1239)
        b = m.eleWithFixedDOFs.astype(np.int64) - 1
        uid = m.mapUniqueKes[b].astype(np.int64) - 1
        if not m.super_element:
            blk = lv1.eleModulus[:, None] * np.diag(lv1.Ke)[None, :]
            if b.size:
                dfree = m.uniqueKesFree[0::25, :][:, uid]
                dfix = m.uniqueKesFixed[0::25, :][:, uid]
                blk[b, :] = (dfree * lv1.eleModulus[b][None, :] + dfix).T
        else:
            blk = self.K_mat[0::25, :].T.copy()
            for k, e in enumerate(b):
                comb = (m.uniqueKesFree[:, :, uid[k]] @ lv1.eleModulus[:, e]).reshape(24, 24, order='F')
                kfix = m.uniqueKesFixed[:, uid[k]].reshape(24, 24, order='F')
                blk[e, :] = np.diag(comb) + np.diag(kfix)
        dk = np.zeros((lv1.numNodes, 3))
        scattering(dk, lv1.eNodMat, blk)
        lv1.diagK = dk.reshape(-1)
        # coarsest assembly + factorisation (TOP_GPU.m:1240-1256)
        self.coarse_matrix = coarsest_matrix(self.levels[-1])
        self.coarse_solve = factorize_coarsest(self.coarse_matrix)

    # --- PCG ----------------------------------------------------------------------------
    def pcg(self, b, tol=1e-3, maxIT=800, y0=None, AtX=None, PtV=None):
        """Solving_PCG (TOP_GPU.m:602-667).  Returns (y, its, resnorm history)."""
        AtX = AtX or self.kbyu
        PtV = PtV or self.vcycle
        return pcg(AtX, PtV, b, tol, maxIT, y0)


def pcg(AtX, PtV, b, tol, maxIT, y0=None):
    """Solving_PCG (TOP_GPU.m:602-667): preconditioned CG on the recursively updated residual,
    stopping when ||r||/||b|| < tol.  Returns (y, its, list of resnorms)."""
    b = np.asarray(b, dtype=np.float64)
    normB = np.linalg.norm(b)
    y = np.zeros_like(b) if y0 is None else np.array(y0, dtype=np.float64)
    r = b - AtX(y)
    z = PtV(r)
    p = z.copy()
    x1 = float(z @ r)
    hist = []
    its = 0
    for its in range(1, maxIT + 1):
        v = AtX(p)
        lam = x1 / float(p @ v)
        y = y + lam * p
        r = r - lam * v
        res = np.linalg.norm(r) / normB
        hist.append(res)
        if res < tol:
            break
        z = PtV(r)
        x2 = float(z @ r)
        p = z + x2 / x1 * p
        x1 = x2
    return y, its, hist


def coarsest_matrix(lv):
    """Global stiffness of the coarsest level: sparse(iK, jK, sK') (TOP_GPU.m:1242-1254)."""
    eNod = lv.eNodMat.astype(np.int64)
    eDof = np.concatenate([3 * eNod - 2, 3 * eNod - 1, 3 * eNod], axis=1)[:, C.RE_ORDERING - 1]
    rowI = np.tile(np.arange(24), 24)
    colI = np.repeat(np.arange(24), 24)
    iK = eDof[:, rowI].reshape(-1) - 1
    jK = eDof[:, colI].reshape(-1) - 1
    sK = lv.Ks.reshape(576, -1, order='F').T.reshape(-1)
    n = 3 * lv.numNodes
    return scipy.sparse.csc_matrix((sK, (iK, jK)), shape=(n, n))


def factorize_coarsest(K):
    """[L,~,S] = chol(K,'lower') and x = S*(L'\\(L\\(S'*r))) (TOP_GPU.m:1255, 796-802).

    Any exact SPD factorisation restates it up to rounding; dense Cholesky below ~8k DOFs,
    a sparse LU above.
    """
    n = K.shape[0]
    if n <= 8000:
        cf = scipy.linalg.cho_factor(K.toarray(), lower=True)
        return lambda r: scipy.linalg.cho_solve(cf, np.asarray(r, dtype=np.float64))
    lu = scipy.sparse.linalg.splu(K.tocsc())
    return lambda r: lu.solve(np.asarray(r, dtype=np.float64))
