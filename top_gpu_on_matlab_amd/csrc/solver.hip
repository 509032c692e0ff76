// Device-resident multigrid-preconditioned CG for the TOP_GPU hot path.
//
// Per level the solver keeps structured full-grid data (kernels.h): node masks, Jacobi
// diagonals, element matrices (entry-major), transfer coefficients and V-cycle vectors.
// FP64 mode mirrors the reference's double path.  Mixed mode runs the V-cycle in FP32
// (coarsest solve in FP64) inside an FP64 PCG whose operator, dots and updates stay FP64.
#include <cmath>

#include "coarse.h"
#include "galerkin.h"
#include "kernels.h"
#include "kfine_tiled.h"
#include "model.h"
#include "pcg.h"
#include "solver.h"

namespace tg {
namespace {

// ---- Galerkin sources, structured layouts -------------------------------------------------
struct ChildGrid {
    int span, cEY, cEX, fEY, fEX;
    const uint8_t *fineSolid;
    int64_t ce;
    __device__ int child(int s) const {
        const int a = s % span, b = (s / span) % span, c = s / (span * span);
        const int cy = (int)(ce % cEY), cx = (int)((ce / cEY) % cEX), cz = (int)(ce / ((int64_t)cEY * cEX));
        const int64_t f = (int64_t)(span * cy + a) + (int64_t)fEY * ((span * cx + b) + (int64_t)fEX * (span * cz + c));
        return fineSolid[f] ? (int)f : -1;
    }
};

struct SrcDevL2 : ChildGrid {
    const double *E, *K0, *uf, *ux;
    const int32_t *uid;
    __device__ double kch(int f, int i) const {
        const int32_t u = uid[f];
        if (u > 0) return uf[(size_t)(u - 1) * 576 + i] * E[f] + ux[(size_t)(u - 1) * 576 + i];
        return K0[i] * E[f];
    }
};

struct SrcDevHigher : ChildGrid {
    const double *prev;
    int64_t ldp;
    __device__ double kch(int f, int i) const { return prev[(int64_t)i * ldp + f]; }
};

struct SrcDevSuper : ChildGrid {
    const double *Kmat, *Esub, *uf, *ux;
    int64_t ldk;
    const int32_t *uid;
    __device__ double kch(int f, int i) const {
        const int32_t u = uid[f];
        if (u > 0) {
            const size_t fb = (size_t)(u - 1) * 576 * 64, mb = (size_t)f * 64;
            double acc = 0;
            for (int q = 0; q < 64; ++q) acc += uf[fb + (size_t)q * 576 + i] * Esub[mb + q];
            return acc + ux[(size_t)(u - 1) * 576 + i];
        }
        return Kmat[(int64_t)i * ldk + f];
    }
};

template <class Src>
__global__ __launch_bounds__(GK_THREADS) void k_galerkin_dev(Src src, const uint8_t *__restrict__ cSolid,
                                                             int nCh, const double *__restrict__ P,
                                                             double *__restrict__ Ks, int64_t ldk) {
    __shared__ double sB[576], sK[576], sT[576];
    const int64_t ce = blockIdx.x;
    if (!cSolid[ce]) return;
    src.ce = ce;
    double acc[9];
    galerkin_block(acc, src, nCh, P, sB, sK, sT);
#pragma unroll
    for (int i = 0; i < 9; ++i) Ks[(int64_t)galerkin_entry(i) * ldk + ce] = acc[i];
}

// The same product restricted to a list of coarse elements (those with boundary children).
template <class Src>
__global__ __launch_bounds__(GK_THREADS) void k_galerkin_list(Src src, const int64_t *__restrict__ list, int nCh,
                                                              const double *__restrict__ P, double *__restrict__ Ks,
                                                              int64_t ldk) {
    __shared__ double sB[576], sK[576], sT[576];
    const int64_t ce = list[blockIdx.x];
    src.ce = ce;
    double acc[9];
    galerkin_block(acc, src, nCh, P, sB, sK, sT);
#pragma unroll
    for (int i = 0; i < 9; ++i) Ks[(int64_t)galerkin_entry(i) * ldk + ce] = acc[i];
}

// Children moduli of every coarse element, child-major: Ech[ce + ldk * s] = E(child s) (0 when void).
__global__ void k_gather_children(const double *__restrict__ E, const uint8_t *__restrict__ fineSolid, int span,
                                  int cEY, int cEX, int64_t nC, int fEY, int fEX, int nCh, double *__restrict__ Ech,
                                  int64_t ldk) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nC * nCh) return;
    const int64_t ce = t % nC;
    const int s = (int)(t / nC);
    const int a = s % span, b = (s / span) % span, c = s / (span * span);
    const int cy = (int)(ce % cEY), cx = (int)((ce / cEY) % cEX), cz = (int)(ce / ((int64_t)cEY * cEX));
    const int64_t f = (int64_t)(span * cy + a) + (int64_t)fEY * ((span * cx + b) + (int64_t)fEX * (span * cz + c));
    Ech[ce + ldk * s] = fineSolid[f] ? E[f] : 0.0;
}

// Level-2 Galerkin product on the matrix cores: Ks[j * ldk + e] = sum_k Ech[e + ldk * k] G[k * 576 + j]
// (k < 4 KS children, j < 576 entries), i.e. Ks_e = sum_s E_s (P_s' K0 P_s) for every coarse element.
// Each MFMA f64 16x16x4 takes A = a 16-entry x 4-child block of G' and B = a 4-child x 16-element block of
// Ech', so its 16 x 16 result is entries x elements and every store row is 16 consecutive elements.
// Lane layouts of that instruction: lane l holds A(l % 16, l / 16), B(l / 16, l % 16) and result registers
// r = 0..3 at D(4 r + l / 16, l % 16).  A
// wavefront owns 16 NG consecutive elements, keeps their children moduli in registers and walks the 36
// entry tiles; each G value loaded feeds NG MFMAs.  e ranges over the padded count ldk (a multiple of
// 64; padded elements have zero moduli).
typedef double tg_dbl4 __attribute__((ext_vector_type(4)));
template <int KS, int NG>
__global__ __launch_bounds__(256) void k_galerkin_l2_mfma(const double *__restrict__ Ech, const double *__restrict__ G,
                                                          double *__restrict__ Ks, int64_t ldk) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int64_t e0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * (16 * NG);
    if (e0 >= ldk) return;
    double b[NG][KS];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) b[g][kk] = Ech[(e0 + 16 * g + li) + ldk * (lk + 4 * kk)];
    const double *gp = G + (int64_t)lk * 576 + li;
    double a[KS];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) a[kk] = gp[(int64_t)4 * kk * 576];
    for (int t = 0; t < 36; ++t) {
        tg_dbl4 acc[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) acc[g] = tg_dbl4{0.0, 0.0, 0.0, 0.0};
        double an[KS];  // next tile's G values, loaded while this tile's MFMAs run
        const int tn = t + 1 < 36 ? t + 1 : t;
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) an[kk] = gp[(int64_t)4 * kk * 576 + 16 * tn];
#pragma unroll
        for (int kk = 0; kk < KS; ++kk)
#pragma unroll
            for (int g = 0; g < NG; ++g) acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], b[g][kk], acc[g], 0, 0, 0);
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) Ks[(int64_t)(16 * t + 4 * r + lk) * ldk + e0 + 16 * g + li] = acc[g][r];
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) a[kk] = an[kk];
    }
}

// K_mat = I_mat * E_sub per element (TOP_GPU.m:1101-1103), entry-major output.
__global__ void k_kmat(const double *__restrict__ Imat, const double *__restrict__ Esub, double *__restrict__ Km,
                       int64_t ldk, int64_t nE) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nE * 576) return;
    const int64_t e = t % nE, i = t / nE;
    double acc = 0;
    for (int q = 0; q < 64; ++q) acc += Imat[i + 576 * q] * Esub[q + 64 * e];
    Km[i * ldk + e] = acc;
}

// Jacobi diagonal: sum over solid elements around the node of the element matrix diagonal.
// MODE 0: E_e * K0(j,j) (finest, standard); 1: Km entry-major (super-element finest / coarse).
// Fixed DOFs (finest level) get the identity row's 1; nodes outside the mesh get 1.
template <int MODE>
__global__ void k_diag(double *__restrict__ diag, const double *__restrict__ E, const double *__restrict__ K0,
                       const double *__restrict__ Km, int64_t ldk, const uint8_t *__restrict__ eleMask,
                       const uint8_t *__restrict__ mask, const Grid g) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes()) return;
    const uint8_t mk = mask[n];
    if (!(mk & NODE_ACTIVE)) {
        diag[3 * n] = diag[3 * n + 1] = diag[3 * n + 2] = 1.0;
        return;
    }
    const int iy = (int)(n % g.NY);
    const int64_t t = n / g.NY;
    const int ix = (int)(t % g.NX), iz = (int)(t / g.NX);
    double a0 = 0, a1 = 0, a2 = 0;
    for (int oz = -1; oz <= 0; ++oz)
        for (int ox = -1; ox <= 0; ++ox)
            for (int oy = -1; oy <= 0; ++oy) {
                const int ey = iy + oy, ex = ix + ox, ez = iz + oz;
                if ((unsigned)ey >= (unsigned)g.EY || (unsigned)ex >= (unsigned)g.EX || (unsigned)ez >= (unsigned)g.EZ)
                    continue;
                const int64_t e = ey + (int64_t)g.EY * (ex + (int64_t)g.EX * ez);
                if (!eleMask[e]) continue;
                const int a = vtx_of(-oy, -ox, -oz);
                if (MODE == 0) {
                    a0 += K0[(3 * a + 0) * 25] * E[e];
                    a1 += K0[(3 * a + 1) * 25] * E[e];
                    a2 += K0[(3 * a + 2) * 25] * E[e];
                } else {
                    a0 += Km[(int64_t)((3 * a + 0) * 25) * ldk + e];
                    a1 += Km[(int64_t)((3 * a + 1) * 25) * ldk + e];
                    a2 += Km[(int64_t)((3 * a + 2) * 25) * ldk + e];
                }
            }
    if (mk & 1) a0 = 1.0;
    if (mk & 2) a1 = 1.0;
    if (mk & 4) a2 = 1.0;
    diag[3 * n] = a0;
    diag[3 * n + 1] = a1;
    diag[3 * n + 2] = a2;
}

__global__ void k_scatter_modulus(const double *__restrict__ Ec, double *__restrict__ Ef,
                                  const int32_t *__restrict__ eleFull0, int64_t Ne, int per) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Ne * per) return;
    const int64_t k = t / per, q = t % per;
    Ef[(int64_t)eleFull0[k] * per + q] = Ec[t];
}

}  // namespace

// ---------------------------------------------------------------------------------------------
void build_pch(int span, std::vector<double> &P) {
    const int nb = span + 1, s3 = span * span * span;
    // child order of Solving_SubEleNodMat (TOP_GPU.m:1938-2017): span 4 column-major, span 2 per
    // z-layer in hex-vertex order (0,0),(1,0),(1,1),(0,1)
    std::vector<int> cy, cx, cz;
    for (int ez = 0; ez < span; ++ez) {
        if (span == 2) {
            const int oy[4] = {0, 1, 1, 0}, ox[4] = {0, 0, 1, 1};
            for (int k = 0; k < 4; ++k) { cy.push_back(oy[k]); cx.push_back(ox[k]); cz.push_back(ez); }
        } else {
            for (int ex = 0; ex < span; ++ex)
                for (int ey = 0; ey < span; ++ey) { cy.push_back(ey); cx.push_back(ex); cz.push_back(ez); }
        }
    }
    P.assign((size_t)576 * s3, 0.0);
    for (int s = 0; s < s3; ++s)
        for (int v = 0; v < 8; ++v) {
            const int a = cy[s] + VY[v], b = cx[s] + VX[v], c = cz[s] + VZ[v];
            (void)nb;
            for (int j = 0; j < 8; ++j) {
                const double w = ri_weight(span, a, b, c, j);
                for (int d = 0; d < 3; ++d) P[(size_t)576 * s + (3 * v + d) + 24 * (3 * j + d)] = w;
            }
        }
}

static Grid grid_of(const HostLevel &L) {
    Grid g;
    g.EY = L.resY;
    g.EX = L.resX;
    g.EZ = L.resZ;
    g.NY = L.resY + 1;
    g.NX = L.resX + 1;
    g.NZ = L.resZ + 1;
    return g;
}

Solver::Solver(const tg_model &m, hipStream_t s) : model(m), st(s) {
    const char *id = "tg_solver_create:Input";
    if (!m.finalized) fail(TG_ERR_STATE, id, "model has no boundary conditions (call tg_model_finalize first)");
    L = m.numLevels;
    prec = m.prm.precision;
    omega = m.prm.weight_factor_jacobi;
    adapted = m.prm.vcycle_type == 1;
    const_k0 = m.prm.poisson_ratio == 0.3;
    super = m.prm.super_element != 0;
    lv.resize(L);
    for (int i = 0; i < L; ++i) {
        const HostLevel &H = m.lv[i];
        DevLevel &D = lv[i];
        D.g = grid_of(H);
        D.span = H.span;
        D.Nn = H.Nn;
        D.Ne = H.Ne;
        D.ldk = ((D.g.elems() + 63) / 64) * 64;
        std::vector<uint8_t> mask(D.g.nodes(), 0);
        std::vector<int32_t> nf(H.Nn), ef(H.Ne);
        for (int64_t k = 0; k < H.Nn; ++k) {
            nf[k] = H.nodFull[k] - 1;
            mask[nf[k]] = NODE_ACTIVE | (i == 0 ? (m.nodeFixMask[k] & 7) : 0);
        }
        for (int64_t k = 0; k < H.Ne; ++k) ef[k] = H.eleMapBack[k] - 1;
        D.nodeMask.upload(mask, st);
        D.eleMask.upload(H.solid, st);
        D.nodFull0.upload(nf, st);
        D.eleFull0.upload(ef, st);
        const int64_t nv = 3 * D.g.nodes();
        D.diag64.alloc(nv);
        D.r64.alloc(nv);
        D.x64.alloc(nv);
        D.t64.alloc(nv);
        if (prec == TG_PREC_MIXED) {
            D.diag32.alloc(nv);
            D.r32.alloc(nv);
            D.x32.alloc(nv);
            D.t32.alloc(nv);
        }
        if (i >= 1) {
            D.Ks64.alloc(576 * D.ldk);
            D.fv64.alloc(24 * D.ldk);
            D.Ks64.zero(st);
            if (prec == TG_PREC_MIXED) {
                D.Ks32.alloc(KSYM_N * D.ldk);  // packed symmetric (FP32 V-cycle)
                D.fv32.alloc(24 * D.ldk);
            } else if (i < L - 1) {
                D.Ks64p.alloc(KSYM_N * D.ldk);  // packed symmetric (FP64 V-cycle; the coarsest level has no K*u)
            }
            // transfer coefficients on the finer grid (1 where a node is unused): counts of at most 8
            // coarse elements, so one byte each, converted exactly to T where the restriction divides
            const HostLevel &F = m.lv[i - 1];
            std::vector<uint8_t> cf(F.nodesFull(), 1);
            for (int64_t n = 0; n < F.nodesFull(); ++n)
                if (H.fineInterId[n]) {
                    const double c = H.coeff[H.fineInterId[n] - 1];
                    if (!(c >= 1 && c <= 8 && c == (double)(int)c))
                        fail(TG_ERR_STATE, id, "transfer coefficient %g out of range", c);
                    cf[n] = (uint8_t)c;
                }
            D.coeffCnt.upload(cf, st);
            std::vector<double> P;
            build_pch(H.span, P);
            D.P.upload(P, st);
        }
    }
    DevLevel &F = lv[0];
    K0_64.upload(m.Ke, st);
    {
        // Walsh-basis form of the finest element matrix for the tiled K*u (isotropic: 45 entries)
        KUnit k0{};
        for (int i = 0; i < 576; ++i) k0.v[i] = m.Ke[i];
        const KUnit kw = make_kwalsh(k0);
        walsh_ok = true;
        for (int i = 0; i < 576; ++i) {
            const int c = KW_PATTERN.idx[i];
            if (c < 0 && kw.v[i] != 0) walsh_ok = false;
            if (c >= 0) {
                kw64.v[c] = kw.v[i];
                kw32.v[c] = (float)kw.v[i];
            }
        }
    }
    if (prec == TG_PREC_MIXED) {
        std::vector<float> k32(m.Ke.begin(), m.Ke.end());
        K0_32.upload(k32, st);
    }
    const int64_t nEf = F.g.elems();
    std::vector<int32_t> uid(nEf, 0);
    for (int64_t k = 0; k < F.Ne; ++k) uid[m.lv[0].eleMapBack[k] - 1] = m.mapUniqueKes[k];
    fullUid.upload(uid, st);
    if (!super && L >= 2) {
        // Level-2 Galerkin as a GEMM on the matrix cores: Ks_e = sum_s E_s G_s with G_s = Pch_s' K0 Pch_s
        // (linear in the children moduli) for children without fixed DOFs; coarse elements with a
        // boundary child are recomputed child by child.
        DevLevel &C = lv[1];
        const int span = C.span, nCh = span * span * span;
        std::vector<double> Ph;
        build_pch(span, Ph);
        std::vector<double> gt((size_t)nCh * 576, 0.0), tmp(576);
        for (int sc = 0; sc < nCh; ++sc) {
            const double *Pc = Ph.data() + (size_t)576 * sc;
            for (int c = 0; c < 24; ++c)          // tmp = K0 Pch_s
                for (int a = 0; a < 24; ++a) {
                    double v = 0;
                    for (int b = 0; b < 24; ++b) v += m.Ke[a + 24 * b] * Pc[b + 24 * c];
                    tmp[a + 24 * c] = v;
                }
            for (int c = 0; c < 24; ++c)          // G_s = Pch_s' tmp
                for (int r = 0; r < 24; ++r) {
                    double v = 0;
                    for (int a = 0; a < 24; ++a) v += Pc[a + 24 * r] * tmp[a + 24 * c];
                    gt[(size_t)576 * sc + (r + 24 * c)] = v;
                }
        }
        Gl2.upload(gt, st);
        std::vector<int64_t> bnd;
        const Grid &cgd = C.g, &fgd = F.g;
        for (int64_t ce = 0; ce < cgd.elems(); ++ce) {
            const int cy = (int)(ce % cgd.EY), cx = (int)((ce / cgd.EY) % cgd.EX), cz = (int)(ce / ((int64_t)cgd.EY * cgd.EX));
            bool b = false;
            for (int sc = 0; sc < nCh && !b; ++sc) {
                const int a = sc % span, bb2 = (sc / span) % span, c = sc / (span * span);
                const int64_t f = (int64_t)(span * cy + a) + (int64_t)fgd.EY * ((span * cx + bb2) + (int64_t)fgd.EX * (span * cz + c));
                b = uid[f] > 0;
            }
            if (b) bnd.push_back(ce);
        }
        nBnd = (int64_t)bnd.size();
        if (nBnd) bndList.upload(bnd, st);
        Ech.alloc((size_t)nCh * C.ldk);
        Ech.zero(st);
    }
    uFree.upload(m.uniqueFree, st);
    uFixed.upload(m.uniqueFixed, st);
    if (super) {
        Imat.upload(m.Imat, st);
        Esub.alloc(64 * nEf);
        Esub.zero(st);
        F.Ks64.alloc(576 * F.ldk);  // K_mat on the finest level
        F.fv64.alloc(24 * F.ldk);
        F.Ks64.zero(st);
        if (prec == TG_PREC_MIXED) {
            F.Ks32.alloc(KSYM_N * F.ldk);  // packed symmetric (FP32 V-cycle)
            F.fv32.alloc(24 * F.ldk);
        } else {
            F.Ks64p.alloc(KSYM_N * F.ldk);  // packed symmetric (FP64 V-cycle and PCG operator)
        }
    } else {
        E64.alloc(nEf);
        E64.zero(st);
        if (prec == TG_PREC_MIXED) E32.alloc(nEf);
    }
    const int64_t nv = 3 * F.g.nodes();
    b.alloc(nv);
    y.alloc(nv);
    pw.init(nv, st);
    if (prec == TG_PREC_MIXED) {
        rin32.alloc(nv);
        zout32.alloc(nv);
        pw.r32 = rin32.p;
        pw.z32 = zout32.p;
    }
    TG_HIP(hipEventCreate(&ev0), id);
    TG_HIP(hipEventCreate(&ev1), id);
    coarse.setup(m.lv[L - 1], lv[L - 1].g, st);
    TG_HIP(hipStreamSynchronize(st), id);
}

Solver::~Solver() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (auto &t : tpool) {
        (void)hipEventDestroy(t.ev[0]);
        (void)hipEventDestroy(t.ev[1]);
    }
}

hipEvent_t *Solver::timing_slot(int key) {
    if (tused == tpool.size()) {
        TimedLaunch t;
        TG_HIP(hipEventCreate(&t.ev[0]), "tg_solver:timing");
        TG_HIP(hipEventCreate(&t.ev[1]), "tg_solver:timing");
        tpool.push_back(t);
    }
    TimedLaunch &t = tpool[tused++];
    t.key = key;
    return t.ev;
}

void Solver::set_modulus(const double *dE) {
    const char *id = "tg_solver_set_modulus:Input";
    DevLevel &F = lv[0];
    if (super) {
        Esub.zero(st);
        k_scatter_modulus<<<grid_for(F.Ne * 64, 256), 256, 0, st>>>(dE, Esub, F.eleFull0, F.Ne, 64);
    } else {
        E64.zero(st);
        k_scatter_modulus<<<grid_for(F.Ne, 256), 256, 0, st>>>(dE, E64, F.eleFull0, F.Ne, 1);
    }
    TG_LAUNCH_CHECK(id);
    has_modulus = true;
    assembled = false;
}

void Solver::assemble() {
    const char *id = "tg_solver_assemble:State";
    if (!has_modulus) fail(TG_ERR_STATE, id, "set the element moduli first");
    DevLevel &F = lv[0];
    if (super) {
        k_kmat<<<grid_for(F.g.elems() * 576, 256), 256, 0, st>>>(Imat, Esub, F.Ks64, F.ldk, F.g.elems());
        TG_LAUNCH_CHECK(id);
    }
    for (int i = 1; i < L; ++i) {
        DevLevel &C = lv[i], &Fi = lv[i - 1];
        const int nCh = C.span * C.span * C.span;
        ChildGrid cg{C.span, C.g.EY, C.g.EX, Fi.g.EY, Fi.g.EX, Fi.eleMask.p, 0};
        const unsigned grid = (unsigned)C.g.elems();
        if (i == 1 && !super) {
            SrcDevL2 src;
            static_cast<ChildGrid &>(src) = cg;
            src.E = E64; src.K0 = K0_64; src.uf = uFree; src.ux = uFixed; src.uid = fullUid;
            const int64_t nC = C.g.elems();
            k_gather_children<<<grid_for(nC * nCh, 256), 256, 0, st>>>(E64, Fi.eleMask, C.span, C.g.EY, C.g.EX, nC,
                                                                      Fi.g.EY, Fi.g.EX, nCh, Ech, C.ldk);
            TG_LAUNCH_CHECK(id);
            // ldk is a multiple of 64: 64 elements per wavefront, 256 per workgroup
            if (nCh == 64)
                k_galerkin_l2_mfma<16, 4><<<grid_for(C.ldk, 256), 256, 0, st>>>(Ech, Gl2, C.Ks64, C.ldk);
            else if (nCh == 8)
                k_galerkin_l2_mfma<2, 4><<<grid_for(C.ldk, 256), 256, 0, st>>>(Ech, Gl2, C.Ks64, C.ldk);
            else
                fail(TG_ERR_ARGS, id, "level-2 Galerkin: %d children per coarse element", nCh);
            TG_LAUNCH_CHECK(id);
            if (nBnd) k_galerkin_list<<<(unsigned)nBnd, GK_THREADS, 0, st>>>(src, bndList, nCh, C.P, C.Ks64, C.ldk);
        } else if (i == 1) {
            SrcDevSuper src;
            static_cast<ChildGrid &>(src) = cg;
            src.Kmat = F.Ks64; src.ldk = F.ldk; src.Esub = Esub; src.uf = uFree; src.ux = uFixed; src.uid = fullUid;
            k_galerkin_dev<<<grid, GK_THREADS, 0, st>>>(src, C.eleMask, nCh, C.P, C.Ks64, C.ldk);
        } else {
            SrcDevHigher src;
            static_cast<ChildGrid &>(src) = cg;
            src.prev = Fi.Ks64; src.ldp = Fi.ldk;
            k_galerkin_dev<<<grid, GK_THREADS, 0, st>>>(src, C.eleMask, nCh, C.P, C.Ks64, C.ldk);
        }
        TG_LAUNCH_CHECK(id);
    }
    // Jacobi diagonals on levels 1..L-1 (TOP_GPU.m:1149-1239)
    for (int i = 0; i < L - 1; ++i) {
        DevLevel &D = lv[i];
        const unsigned gr = grid_for(D.g.nodes(), 256);
        if (i == 0 && !super)
            k_diag<0><<<gr, 256, 0, st>>>(D.diag64, E64, K0_64, nullptr, 0, D.eleMask, D.nodeMask, D.g);
        else
            k_diag<1><<<gr, 256, 0, st>>>(D.diag64, nullptr, nullptr, D.Ks64, D.ldk, D.eleMask, D.nodeMask, D.g);
        TG_LAUNCH_CHECK(id);
    }
    coarse.factor(lv[L - 1].Ks64, lv[L - 1].ldk, lv[L - 1].eleMask, lv[L - 1].nodeMask, st);
    if (prec == TG_PREC_MIXED) {
        for (int i = 0; i < L; ++i) {
            DevLevel &D = lv[i];
            if (i < L - 1)
                k_convert<double, float><<<grid_for(3 * D.g.nodes(), 256), 256, 0, st>>>(D.diag64, D.diag32, 3 * D.g.nodes());
            if ((i >= 1 && i < L - 1) || (i == 0 && super))
                k_pack_sym<double, float><<<grid_for(KSYM_N * D.ldk, 256), 256, 0, st>>>(D.Ks64, D.Ks32, D.ldk);
        }
        if (!super) k_convert<double, float><<<grid_for(F.g.elems(), 256), 256, 0, st>>>(E64, E32, F.g.elems());
        TG_LAUNCH_CHECK(id);
    } else {
        for (int i = 0; i < L - 1; ++i) {
            DevLevel &D = lv[i];
            if (D.Ks64p.n) k_pack_sym<double, double><<<grid_for(KSYM_N * D.ldk, 256), 256, 0, st>>>(D.Ks64, D.Ks64p, D.ldk);
        }
        TG_LAUNCH_CHECK(id);
    }
    assembled = true;
}

// A fused dot's partial sums: left for the PCG's next update to finish when it takes them, else summed
// into pw.scal[slot] here.
void Solver::finish_dot(const double *parts, int n, int slot) {
    if (pw.defer_ok && !pw.deferred.parts) pw.deferred = DeferredDot{parts, n, slot};
    else k_reduce_final<<<1, RED_THREADS, 0, st>>>(parts, n, pw.scal, slot);
}

// Grid of k_kelem_gather: one row per node plane (its in-plane and element indices are 32-bit).
static dim3 gather_grid(const Grid &g) {
    if ((int64_t)g.NY * g.NX >= (1ll << 31) || g.elems() >= (1ll << 31) || g.NZ > 65535)
        fail(TG_ERR_SHAPE, "tg_solver:kbyu", "grid too large for the element-gather kernel's indices");
    return dim3(grid_for((int64_t)g.NY * g.NX, 256), (unsigned)g.NZ);
}

template <typename T, bool MASKED>
static void launch_kelem_vec(int parts, hipStream_t st, const T *u, T *fv, const T *Km, int64_t ldk,
                             const uint8_t *eleMask, const uint8_t *nodeMask, const Grid &g) {
    if (parts != 4 && parts != 6 && parts != 8 && parts != 12) parts = 24;  // the grid must match the variant
    const unsigned ge = grid_for(parts * ldk, 256);
    switch (parts) {
        case 4: k_kelem_vec<T, MASKED, 4><<<ge, 256, 0, st>>>(u, fv, Km, ldk, eleMask, nodeMask, g); break;
        case 6: k_kelem_vec<T, MASKED, 6><<<ge, 256, 0, st>>>(u, fv, Km, ldk, eleMask, nodeMask, g); break;
        case 8: k_kelem_vec<T, MASKED, 8><<<ge, 256, 0, st>>>(u, fv, Km, ldk, eleMask, nodeMask, g); break;
        case 12: k_kelem_vec<T, MASKED, 12><<<ge, 256, 0, st>>>(u, fv, Km, ldk, eleMask, nodeMask, g); break;
        default: k_kelem_vec<T, MASKED, 24><<<ge, 256, 0, st>>>(u, fv, Km, ldk, eleMask, nodeMask, g); break;
    }
}

template <typename T, bool CK, bool DOT>
static void launch_walsh_k(int epi, unsigned gt, hipStream_t st, const T *u, T *out, const T *E, const uint8_t *mask,
                           const KWVals<T> &kw, T cs, const Grid &g, const KTiles &kt, const T *r, const T *d, T w,
                           double *partial, T *xout, const float *mk, int mf) {
    if constexpr (sizeof(T) == 4 && !CK) {
        if (mk && mf == 1) {  // element products on the matrix cores, FP32
            if (epi == EPI_KU)
                k_kfine_walsh<T, EPI_KU, false, DOT, false, 1><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else if (epi == EPI_RESID)
                k_kfine_walsh<T, EPI_RESID, false, false, false, 1><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else if (epi == EPI_RESID_JAC0)
                k_kfine_walsh<T, EPI_RESID_JAC0, false, false, false, 1><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else
                k_kfine_walsh<T, EPI_JACOBI, false, DOT, false, 1><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            return;
        }
        if (mk && mf == 2) {  // element products on the matrix cores, BF16 operands, FP32 accumulation
            if (epi == EPI_KU)
                k_kfine_walsh<T, EPI_KU, false, DOT, false, 2><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else if (epi == EPI_RESID)
                k_kfine_walsh<T, EPI_RESID, false, false, false, 2><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else if (epi == EPI_RESID_JAC0)
                k_kfine_walsh<T, EPI_RESID_JAC0, false, false, false, 2><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else
                k_kfine_walsh<T, EPI_JACOBI, false, DOT, false, 2><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            return;
        }
        if (mk && mf == 3) {  // as 2, each operand split into two BF16 parts (three products)
            if (epi == EPI_KU)
                k_kfine_walsh<T, EPI_KU, false, DOT, false, 3><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else if (epi == EPI_RESID)
                k_kfine_walsh<T, EPI_RESID, false, false, false, 3><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else if (epi == EPI_RESID_JAC0)
                k_kfine_walsh<T, EPI_RESID_JAC0, false, false, false, 3><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            else
                k_kfine_walsh<T, EPI_JACOBI, false, DOT, false, 3><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, mk);
            return;
        }
    }
    if (epi == EPI_KU)
        k_kfine_walsh<T, EPI_KU, CK, DOT><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr);
    else if (epi == EPI_RESID)
        k_kfine_walsh<T, EPI_RESID, CK, false><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr);
    else if (epi == EPI_RESID_JAC0)
        k_kfine_walsh<T, EPI_RESID_JAC0, CK, false><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr);
    else
        k_kfine_walsh<T, EPI_JACOBI, CK, DOT><<<gt, 256, 0, st>>>(u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr);
}

// mk, mf (FP32 only): the element matrix in FP32 and the matrix-core mode (1 FP32, 2 BF16 operands) of the
// element products; mf = 0 (or mk null): the Walsh form
template <typename T>
static void launch_walsh(int epi, bool ck, unsigned gt, hipStream_t st, const T *u, T *out, const T *E,
                         const uint8_t *mask, const KWVals<T> &kw, T cs, const Grid &g, const KTiles &kt,
                         const T *r, const T *d, T w, double *partial, T *xout, const float *mk = nullptr, int mf = 0) {
    if (!mf) mk = nullptr;
    if (ck && !mk) {
        if (partial) launch_walsh_k<T, true, true>(epi, gt, st, u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, 0);
        else launch_walsh_k<T, true, false>(epi, gt, st, u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, nullptr, 0);
    } else {
        if (partial) launch_walsh_k<T, false, true>(epi, gt, st, u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, mk, mf);
        else launch_walsh_k<T, false, false>(epi, gt, st, u, out, E, mask, kw, cs, g, kt, r, d, w, partial, xout, mk, mf);
    }
}

// dot_slot >= 0 (finest level, tiled kernel, EPI_KU or EPI_JACOBI only): also leave u.out (K*u) or
// r.out (Jacobi) in pw.scal[dot_slot] and return true; otherwise return false.
template <typename T>
bool Solver::kbyu_level(int i, const T *u, T *out, int epi, const T *rr, const T *dg, int dot_slot, T *xout) {
    bool dotted = false;
    DevLevel &D = lv[i];
    const unsigned gr = grid_for(D.g.nodes(), 256);
    const T w = (T)omega;
    const bool f64 = sizeof(T) == 8;
    if (i == 0 && !super) {
        const T *E = f64 ? (const T *)E64.p : (const T *)E32.p;
        const T *K0 = f64 ? (const T *)K0_64.p : (const T *)K0_32.p;
        const T cs = (T)model.prm.cell_size;
        if (kfine_kernel == 0 && walsh_ok) {
            const KTiles kt = ktiles_for(D.g, kfine_zchunk);
            const unsigned gt = (unsigned)kt.count();
            hipEvent_t *ev = (time_fine == 1 || (time_fine == 2 && epi == EPI_KU)) ? timing_slot((f64 ? 0 : 4) + epi) : nullptr;
            double *dp = nullptr;
            if (dot_slot >= 0 && epi != EPI_RESID) {
                if ((int64_t)dotp.n < (int64_t)gt) dotp.alloc(gt);
                dp = dotp.p;
            }
            if (ev) TG_HIP(hipEventRecord(ev[0], st), "tg_solver:kbyu");
            if (f64) {
                const KWVals<double> &k = kw64;
                const double *uu = (const double *)u, *r2 = (const double *)rr, *d2 = (const double *)dg;
                const double *E2 = (const double *)E;
                double *o2 = (double *)out;
                launch_walsh<double>(epi, const_k0, gt, st, uu, o2, E2, D.nodeMask, k, model.prm.cell_size, D.g, kt, r2, d2, omega, dp,
                                     (double *)xout);
            } else {
                const KWVals<float> &k = kw32;
                const float *uu = (const float *)u, *r2 = (const float *)rr, *d2 = (const float *)dg;
                const float *E2 = (const float *)E;
                float *o2 = (float *)out;
                const float w32 = (float)omega;
                launch_walsh<float>(epi, const_k0, gt, st, uu, o2, E2, D.nodeMask, k, (float)model.prm.cell_size, D.g, kt, r2, d2, w32, dp,
                                    (float *)xout, K0_32.p, kfine_mfma);
            }
            if (ev) TG_HIP(hipEventRecord(ev[1], st), "tg_solver:kbyu");
            if (dp) {
                finish_dot(dp, (int)gt, dot_slot);
                dotted = true;
            }
        } else if (const_k0) {
            if (epi == EPI_KU) k_kfine<T, EPI_KU, true><<<gr, 256, 0, st>>>(u, out, E, D.nodeMask, K0, cs, D.g, rr, dg, w);
            else if (epi == EPI_RESID) k_kfine<T, EPI_RESID, true><<<gr, 256, 0, st>>>(u, out, E, D.nodeMask, K0, cs, D.g, rr, dg, w);
            else k_kfine<T, EPI_JACOBI, true><<<gr, 256, 0, st>>>(u, out, E, D.nodeMask, K0, cs, D.g, rr, dg, w);
        } else {
            if (epi == EPI_KU) k_kfine<T, EPI_KU, false><<<gr, 256, 0, st>>>(u, out, E, D.nodeMask, K0, (T)1, D.g, rr, dg, w);
            else if (epi == EPI_RESID) k_kfine<T, EPI_RESID, false><<<gr, 256, 0, st>>>(u, out, E, D.nodeMask, K0, (T)1, D.g, rr, dg, w);
            else k_kfine<T, EPI_JACOBI, false><<<gr, 256, 0, st>>>(u, out, E, D.nodeMask, K0, (T)1, D.g, rr, dg, w);
        }
    } else {
        // FP64: the packed upper triangles (one thread per element) unless a thread split of the full
        // matrices is asked for; FP32: always packed
        const bool packed = !f64 || (kelem_parts == 1 && D.Ks64p.n);
        const T *Km = f64 ? (const T *)(packed ? D.Ks64p.p : D.Ks64.p) : (const T *)D.Ks32.p;
        T *fv = f64 ? (T *)D.fv64.p : (T *)D.fv32.p;
        const unsigned gs = grid_for(D.g.elems(), 256);
        const dim3 grp = gather_grid(D.g);
        if (i == 0) {
            if (!packed) launch_kelem_vec<T, true>(kelem_parts, st, u, fv, Km, D.ldk, D.eleMask, D.nodeMask, D.g);
            else k_kelem_sym<T, true><<<gs, 256, 0, st>>>(u, fv, Km, D.ldk, D.eleMask, D.nodeMask, D.g);
            if (epi == EPI_KU) k_kelem_gather<T, EPI_KU, true><<<grp, 256, 0, st>>>(fv, D.ldk, out, u, D.eleMask, D.nodeMask, D.g, rr, dg, w);
            else if (epi == EPI_RESID) k_kelem_gather<T, EPI_RESID, true><<<grp, 256, 0, st>>>(fv, D.ldk, out, u, D.eleMask, D.nodeMask, D.g, rr, dg, w);
            else k_kelem_gather<T, EPI_JACOBI, true><<<grp, 256, 0, st>>>(fv, D.ldk, out, u, D.eleMask, D.nodeMask, D.g, rr, dg, w);
        } else {
            if (!packed) launch_kelem_vec<T, false>(kelem_parts, st, u, fv, Km, D.ldk, D.eleMask, D.nodeMask, D.g);
            else k_kelem_sym<T, false><<<gs, 256, 0, st>>>(u, fv, Km, D.ldk, D.eleMask, D.nodeMask, D.g);
            if (epi == EPI_KU) k_kelem_gather<T, EPI_KU, false><<<grp, 256, 0, st>>>(fv, D.ldk, out, u, D.eleMask, D.nodeMask, D.g, rr, dg, w);
            else if (epi == EPI_RESID) k_kelem_gather<T, EPI_RESID, false><<<grp, 256, 0, st>>>(fv, D.ldk, out, u, D.eleMask, D.nodeMask, D.g, rr, dg, w);
            else k_kelem_gather<T, EPI_JACOBI, false><<<grp, 256, 0, st>>>(fv, D.ldk, out, u, D.eleMask, D.nodeMask, D.g, rr, dg, w);
        }
    }
    TG_LAUNCH_CHECK("tg_solver:kbyu");
    return dotted;
}

template <typename T>
void Solver::restrict_level(int c, const T *rf, T *rc, const PlanePack *pk) {
    DevLevel &C = lv[c], &F = lv[c - 1];
    const uint8_t *cf = C.coeffCnt.p;
    T *fv = sizeof(T) == 8 ? (T *)C.fv64.p : (T *)C.fv32.p;
    const unsigned gt = (unsigned)(((C.g.EY + RT_E - 1) / RT_E) * ((C.g.EX + RT_E - 1) / RT_E) *
                                   ((C.g.EZ + RT_Z - 1) / RT_Z));
    if (C.span == 4) k_restrict_tiled<T, 4><<<gt, RT_THREADS, 0, st>>>(rf, cf, C.g, F.g, fv, C.ldk);
    else k_restrict_tiled<T, 2><<<gt, RT_THREADS, 0, st>>>(rf, cf, C.g, F.g, fv, C.ldk);
    k_kelem_gather<T, EPI_KU, false><<<gather_grid(C.g), 256, 0, st>>>(
        fv, C.ldk, rc, nullptr, C.eleMask, C.nodeMask, C.g, nullptr, nullptr, (T)0, pk ? *pk : PlanePack{});
    TG_LAUNCH_CHECK("tg_solver:restrict");
}

template <typename T>
void Solver::prolong_level(int c, const T *xc, T *xf, bool add) {
    DevLevel &C = lv[c], &F = lv[c - 1];
    const unsigned gr = grid_for(F.g.nodes(), 256);
    if (C.span == 4) {
        if (add) k_prolong<T, 4, true><<<gr, 256, 0, st>>>(xc, xf, F.nodeMask, C.eleMask, C.g, F.g);
        else k_prolong<T, 4, false><<<gr, 256, 0, st>>>(xc, xf, F.nodeMask, C.eleMask, C.g, F.g);
    } else {
        if (add) k_prolong<T, 2, true><<<gr, 256, 0, st>>>(xc, xf, F.nodeMask, C.eleMask, C.g, F.g);
        else k_prolong<T, 2, false><<<gr, 256, 0, st>>>(xc, xf, F.nodeMask, C.eleMask, C.g, F.g);
    }
    TG_LAUNCH_CHECK("tg_solver:prolong");
}

template <typename T>
static T *vec(DevLevel &D, int which) {
    if (sizeof(T) == 8) return (T *)(which == 0 ? D.r64.p : (which == 1 ? D.x64.p : D.t64.p));
    return (T *)(which == 0 ? D.r32.p : (which == 1 ? D.x32.p : D.t32.p));
}

template <typename T>
static const T *diag_of(DevLevel &D) {
    return sizeof(T) == 8 ? (const T *)D.diag64.p : (const T *)D.diag32.p;
}

// Solving_Vcycle (TOP_GPU.m:766-816) on full-grid vectors of precision T.
template <typename T>
void Solver::vcycle_t(const T *r0, T *z0) {
    const T w = (T)omega;
    std::vector<const T *> rr(L);
    std::vector<T *> xx(L);
    rr[0] = r0;
    for (int i = 1; i < L; ++i) rr[i] = vec<T>(lv[i], 0);
    for (int i = 0; i < L; ++i) xx[i] = vec<T>(lv[i], 1);
    // the restriction into the coarsest level also packs its persistent solve's right-hand side
    PlanePack pk;
    const bool packed = coarse.plane_pack(pk, st);
    for (int i = 0; i < L - 1; ++i) {
        const PlanePack *pki = (packed && i + 1 == L - 1) ? &pk : nullptr;
        DevLevel &D = lv[i];
        const int64_t nv = 3 * D.g.nodes();
        if (i == 0 && !adapted && fine_tiled()) {
            // pre-smoothing x = omega r ./ diag fused into the residual product
            T *d = vec<T>(D, 2);
            kbyu_level<T>(0, nullptr, d, EPI_RESID_JAC0, rr[0], diag_of<T>(D), -1, xx[0]);
            restrict_level<T>(1, d, (T *)rr[1], pki);
            continue;
        }
        k_jacobi0<T><<<grid_for(nv, 256), 256, 0, st>>>(rr[i], diag_of<T>(D), xx[i], w, nv);
        if (adapted) {
            restrict_level<T>(i + 1, rr[i], (T *)rr[i + 1], pki);
        } else {
            T *d = vec<T>(D, 2);
            kbyu_level<T>(i, xx[i], d, EPI_RESID, rr[i], nullptr);
            restrict_level<T>(i + 1, d, (T *)rr[i + 1], pki);
        }
    }
    coarse.solve<T>(rr[L - 1], xx[L - 1], lv[L - 1].nodeMask, st, packed);
    for (int i = L - 2; i >= 0; --i) {
        DevLevel &D = lv[i];
        prolong_level<T>(i + 1, xx[i + 1], xx[i], true);
        // the smoothed iterate goes to a second buffer (neighbours still read the old one);
        // on the finest level straight into the caller's output
        T *out = (i == 0) ? z0 : vec<T>(D, 2);
        if (i == 0 && sizeof(T) == 4 && r64_dot) {
            // mixed: z = x + omega (r - K x) ./ d in FP32, with z.r against the FP64 r
            const KTiles kt = ktiles_for(D.g, kfine_zchunk);
            const unsigned gt = (unsigned)kt.count();
            if ((int64_t)dotp.n < (int64_t)gt) dotp.alloc(gt);
            const float *x32 = (const float *)xx[0], *r32 = (const float *)rr[0], *d32 = D.diag32.p;
            if (kfine_mfma == 1)
                k_kfine_walsh<float, EPI_JACOBI, false, true, true, 1><<<gt, 256, 0, st>>>(
                    x32, (float *)out, E32.p, D.nodeMask, kw32, (float)model.prm.cell_size, D.g, kt, r32, d32,
                    (float)omega, dotp.p, nullptr, r64_dot, K0_32.p);
            else if (kfine_mfma == 2)
                k_kfine_walsh<float, EPI_JACOBI, false, true, true, 2><<<gt, 256, 0, st>>>(
                    x32, (float *)out, E32.p, D.nodeMask, kw32, (float)model.prm.cell_size, D.g, kt, r32, d32,
                    (float)omega, dotp.p, nullptr, r64_dot, K0_32.p);
            else if (kfine_mfma == 3)
                k_kfine_walsh<float, EPI_JACOBI, false, true, true, 3><<<gt, 256, 0, st>>>(
                    x32, (float *)out, E32.p, D.nodeMask, kw32, (float)model.prm.cell_size, D.g, kt, r32, d32,
                    (float)omega, dotp.p, nullptr, r64_dot, K0_32.p);
            else if (const_k0)
                k_kfine_walsh<float, EPI_JACOBI, true, true, true><<<gt, 256, 0, st>>>(
                    x32, (float *)out, E32.p, D.nodeMask, kw32, (float)model.prm.cell_size, D.g, kt, r32, d32,
                    (float)omega, dotp.p, nullptr, r64_dot);
            else
                k_kfine_walsh<float, EPI_JACOBI, false, true, true><<<gt, 256, 0, st>>>(
                    x32, (float *)out, E32.p, D.nodeMask, kw32, (float)model.prm.cell_size, D.g, kt, r32, d32,
                    (float)omega, dotp.p, nullptr, r64_dot);
            finish_dot(dotp, (int)gt, vc_dot_slot);
            vc_dotted = true;
            continue;
        }
        if (adapted) {
            const int64_t nv = 3 * D.g.nodes();
            k_jacobi_adapted<T><<<grid_for(nv, 256), 256, 0, st>>>(xx[i], rr[i], diag_of<T>(D), out, w, nv);
        } else {
            vc_dotted = kbyu_level<T>(i, xx[i], out, EPI_JACOBI, rr[i], diag_of<T>(D), i == 0 ? vc_dot_slot : -1);
        }
        xx[i] = out;
    }
    TG_LAUNCH_CHECK("tg_solver:vcycle");
}

// z0 = M r0.  dot_slot >= 0: also leave z0.r0 in pw.scal[dot_slot] when a fused kernel can (returns
// true), so PCG skips its separate dot.  z32_ok (mixed): z may be left in FP32 in zout32 instead of
// z0 (*z_in32 tells; widening it to FP64 is exact).
bool Solver::vcycle(const double *r0, double *z0, int dot_slot, bool r32_ready, bool z32_ok, bool *z_in32) {
    vc_dot_slot = -1;
    vc_dotted = false;
    if (z_in32) *z_in32 = false;
    if (prec == TG_PREC_MIXED) {
        DevLevel &F = lv[0];
        const int64_t nv = 3 * F.g.nodes();
        if (!r32_ready) k_convert<double, float><<<grid_for(nv, 256), 256, 0, st>>>(r0, rin32, nv);
        if (dot_slot >= 0 && z32_ok && z_in32 && fine_tiled() && !adapted) {
            vc_dot_slot = dot_slot;
            r64_dot = r0;
            vcycle_t<float>(rin32, zout32);
            r64_dot = nullptr;
            vc_dot_slot = -1;
            *z_in32 = true;
            TG_LAUNCH_CHECK("tg_solver:vcycle");
            return vc_dotted;
        }
        vcycle_t<float>(rin32, zout32);
        if (dot_slot >= 0) {
            k_convert_dot<<<RED_BLOCKS, RED_THREADS, 0, st>>>(zout32, r0, z0, nv, pw.partial);
            k_reduce_final<<<1, RED_THREADS, 0, st>>>(pw.partial, RED_BLOCKS, pw.scal, dot_slot);
            TG_LAUNCH_CHECK("tg_solver:vcycle");
            return true;
        }
        k_convert<float, double><<<grid_for(nv, 256), 256, 0, st>>>(zout32, z0, nv);
        return false;
    }
    vc_dot_slot = dot_slot;
    vcycle_t<double>(r0, z0);
    vc_dot_slot = -1;
    return vc_dotted;
}

// Solving_PCG (TOP_GPU.m:602-667) with Solving_KbyU_MatrixFree / Solving_Vcycle as operators.
void Solver::pcg(const double *b_c, double *y_c, double tol, int maxIT, int warm, int *its_out, double *res_out) {
    const char *id = "tg_solver_pcg:State";
    if (!assembled) fail(TG_ERR_STATE, id, "assemble before solving");
    DevLevel &F = lv[0];
    TG_HIP(hipEventRecord(ev0, st), id);
    b.zero(st);
    k_unpack<double, double><<<grid_for(F.Nn, 256), 256, 0, st>>>(b_c, b, F.nodFull0, F.Nn);
    y.zero(st);
    if (warm) k_unpack<double, double><<<grid_for(F.Nn, 256), 256, 0, st>>>(y_c, y, F.nodFull0, F.Nn);
    TG_LAUNCH_CHECK(id);
    int its = 0;
    double res = 0;
    pw.run([&](const double *u, double *out, int slot) { return kbyu_level<double>(0, u, out, EPI_KU, nullptr, nullptr, slot); },
           [&](const double *r, double *z, int slot) {
               return vcycle(r, z, slot, pw.r32_fresh, pw.z32 != nullptr, &pw.z32_fresh);
           },
           b, y, tol, maxIT, warm != 0, &its, &res, "tg_solver_pcg:Numeric");
    k_pack<double, double><<<grid_for(F.Nn, 256), 256, 0, st>>>(y, y_c, F.nodFull0, F.Nn);
    TG_LAUNCH_CHECK(id);
    TG_HIP(hipEventRecord(ev1, st), id);
    TG_HIP(hipEventSynchronize(ev1), id);
    coarse.check_barrier(st);
    float ms = 0;
    TG_HIP(hipEventElapsedTime(&ms, ev0, ev1), id);
    last_ms = ms;
    last_its = its;
    if (its_out) *its_out = its;
    if (res_out) *res_out = res;
}

}  // namespace tg

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------
using namespace tg;

namespace {

Solver &S_of(tg_solver *s, const char *id) {
    if (!s || !s->s) fail(TG_ERR_ARGS, id, "null solver");
    return *s->s;
}

int level_index(const Solver &S, int32_t level, const char *id) {
    if (level < 1 || level > S.L) fail(TG_ERR_ARGS, id, "level %d out of range 1..%d", level, S.L);
    return level - 1;
}

}  // namespace

extern "C" {

int tg_solver_create(const tg_model *m, void *stream, tg_solver **out) {
    return guard("tg_solver_create:Input", [&] {
        if (!m || !out) fail(TG_ERR_ARGS, "tg_solver_create:Input", "null argument");
        auto *h = new tg_solver();
        try {
            h->s = new Solver(*m, (hipStream_t)stream);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void tg_solver_destroy(tg_solver *s) {
    if (!s) return;
    if (s->s) (void)hipStreamSynchronize(s->s->st);
    delete s->s;
    delete s;
}

int tg_solver_set_modulus(tg_solver *s, const double *d_E) {
    return guard("tg_solver_set_modulus:Input", [&] {
        Solver &S = S_of(s, "tg_solver_set_modulus:Input");
        if (!d_E) fail(TG_ERR_ARGS, "tg_solver_set_modulus:Input", "null modulus");
        S.set_modulus(d_E);
    });
}

int tg_solver_assemble(tg_solver *s) {
    return guard("tg_solver_assemble:State", [&] { S_of(s, "tg_solver_assemble:State").assemble(); });
}

int tg_solver_kbyu(tg_solver *s, int32_t level, const double *d_u, double *d_y) {
    return guard("tg_solver_kbyu:Input", [&] {
        const char *id = "tg_solver_kbyu:Input";
        Solver &S = S_of(s, id);
        const int i = level_index(S, level, id);
        if (!S.assembled) fail(TG_ERR_STATE, "tg_solver_kbyu:State", "assemble first");
        if (!d_u || !d_y) fail(TG_ERR_ARGS, id, "null vector");
        DevLevel &D = S.lv[i];
        D.x64.zero(S.st);
        k_unpack<double, double><<<grid_for(D.Nn, 256), 256, 0, S.st>>>(d_u, D.x64, D.nodFull0, D.Nn);
        S.kbyu_level<double>(i, D.x64, D.t64, EPI_KU, nullptr, nullptr);
        k_pack<double, double><<<grid_for(D.Nn, 256), 256, 0, S.st>>>(D.t64, d_y, D.nodFull0, D.Nn);
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
    });
}

int tg_solver_vcycle(tg_solver *s, const double *d_r, double *d_z) {
    return guard("tg_solver_vcycle:Input", [&] {
        const char *id = "tg_solver_vcycle:Input";
        Solver &S = S_of(s, id);
        if (!S.assembled) fail(TG_ERR_STATE, "tg_solver_vcycle:State", "assemble first");
        DevLevel &F = S.lv[0];
        S.pw.r.zero(S.st);
        k_unpack<double, double><<<grid_for(F.Nn, 256), 256, 0, S.st>>>(d_r, S.pw.r, F.nodFull0, F.Nn);
        S.vcycle(S.pw.r, S.pw.z);
        k_pack<double, double><<<grid_for(F.Nn, 256), 256, 0, S.st>>>(S.pw.z, d_z, F.nodFull0, F.Nn);
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
        S.coarse.check_barrier(S.st);
    });
}

int tg_solver_restrict(tg_solver *s, int32_t coarse_level, const double *d_rf, double *d_rc) {
    return guard("tg_solver_restrict:Input", [&] {
        const char *id = "tg_solver_restrict:Input";
        Solver &S = S_of(s, id);
        const int c = level_index(S, coarse_level, id);
        if (c < 1) fail(TG_ERR_ARGS, id, "coarse_level must be >= 2");
        DevLevel &C = S.lv[c], &F = S.lv[c - 1];
        F.t64.zero(S.st);
        k_unpack<double, double><<<grid_for(F.Nn, 256), 256, 0, S.st>>>(d_rf, F.t64, F.nodFull0, F.Nn);
        S.restrict_level<double>(c, F.t64, C.r64);
        k_pack<double, double><<<grid_for(C.Nn, 256), 256, 0, S.st>>>(C.r64, d_rc, C.nodFull0, C.Nn);
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
    });
}

int tg_solver_interpolate(tg_solver *s, int32_t coarse_level, const double *d_xc, double *d_xf) {
    return guard("tg_solver_interpolate:Input", [&] {
        const char *id = "tg_solver_interpolate:Input";
        Solver &S = S_of(s, id);
        const int c = level_index(S, coarse_level, id);
        if (c < 1) fail(TG_ERR_ARGS, id, "coarse_level must be >= 2");
        DevLevel &C = S.lv[c], &F = S.lv[c - 1];
        C.x64.zero(S.st);
        k_unpack<double, double><<<grid_for(C.Nn, 256), 256, 0, S.st>>>(d_xc, C.x64, C.nodFull0, C.Nn);
        S.prolong_level<double>(c, C.x64, F.t64, false);
        k_pack<double, double><<<grid_for(F.Nn, 256), 256, 0, S.st>>>(F.t64, d_xf, F.nodFull0, F.Nn);
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
    });
}

int tg_solver_coarse_solve(tg_solver *s, const double *d_r, double *d_x) {
    return guard("tg_solver_coarse_solve:Input", [&] {
        const char *id = "tg_solver_coarse_solve:Input";
        Solver &S = S_of(s, id);
        if (!S.assembled) fail(TG_ERR_STATE, "tg_solver_coarse_solve:State", "assemble first");
        DevLevel &C = S.lv[S.L - 1];
        C.r64.zero(S.st);
        k_unpack<double, double><<<grid_for(C.Nn, 256), 256, 0, S.st>>>(d_r, C.r64, C.nodFull0, C.Nn);
        S.coarse.solve<double>(C.r64, C.x64, C.nodeMask, S.st);
        k_pack<double, double><<<grid_for(C.Nn, 256), 256, 0, S.st>>>(C.x64, d_x, C.nodFull0, C.Nn);
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
        S.coarse.check_barrier(S.st);
    });
}

int tg_solver_pcg(tg_solver *s, const double *d_b, double *d_y, double tol, int32_t max_it, int32_t warm,
                  int32_t *its, double *resnorm) {
    return guard("tg_solver_pcg:Input", [&] {
        Solver &S = S_of(s, "tg_solver_pcg:Input");
        if (!d_b || !d_y || max_it < 1) fail(TG_ERR_ARGS, "tg_solver_pcg:Input", "bad arguments");
        int it = 0;
        double res = 0;
        S.pcg(d_b, d_y, tol, max_it, warm, &it, &res);
        if (its) *its = it;
        if (resnorm) *resnorm = res;
    });
}

int tg_solver_get_diag(tg_solver *s, int32_t level, double *d_out) {
    return guard("tg_solver_get_diag:Input", [&] {
        const char *id = "tg_solver_get_diag:Input";
        Solver &S = S_of(s, id);
        const int i = level_index(S, level, id);
        if (i >= S.L - 1) fail(TG_ERR_ARGS, id, "the coarsest level has no Jacobi diagonal");
        if (!S.assembled) fail(TG_ERR_STATE, "tg_solver_get_diag:State", "assemble first");
        DevLevel &D = S.lv[i];
        k_pack<double, double><<<grid_for(D.Nn, 256), 256, 0, S.st>>>(D.diag64, d_out, D.nodFull0, D.Nn);
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
    });
}

}  // extern "C"

namespace tg {
namespace {
// d_Ks(:,:,k) of the reference from the entry-major full-grid layout
__global__ void k_ks_to_reference(const double *__restrict__ Km, int64_t ldk, const int32_t *__restrict__ eleFull0,
                                  int64_t Ne, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Ne * 576) return;
    const int64_t k = t / 576, i = t % 576;
    out[t] = Km[i * ldk + eleFull0[k]];
}
}  // namespace
}  // namespace tg

extern "C" {

int tg_solver_get_ks(tg_solver *s, int32_t level, double *d_out) {
    return guard("tg_solver_get_ks:Input", [&] {
        const char *id = "tg_solver_get_ks:Input";
        Solver &S = S_of(s, id);
        const int i = level_index(S, level, id);
        if (i == 0 && !S.super) fail(TG_ERR_ARGS, id, "level 1 has no per-element matrices in standard mode");
        if (!S.assembled) fail(TG_ERR_STATE, "tg_solver_get_ks:State", "assemble first");
        DevLevel &D = S.lv[i];
        k_ks_to_reference<<<grid_for(D.Ne * 576, 256), 256, 0, S.st>>>(D.Ks64, D.ldk, D.eleFull0, D.Ne, d_out);
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
    });
}

int tg_solver_last_timing(const tg_solver *s, double *ms_total) {
    if (!s || !s->s || !ms_total) return TG_ERR_ARGS;
    *ms_total = s->s->last_ms;
    return TG_OK;
}

}  // extern "C"

extern "C" {

// Launch choices of the solver (TG_TUNE_* keys); no effect on results.
int tg_solver_set_tuning(tg_solver *s, int32_t key, int32_t value) {
    return guard("tg_solver_set_tuning:Input", [&] {
        const char *id = "tg_solver_set_tuning:Input";
        Solver &S = S_of(s, id);
        if (key == TG_TUNE_KFINE_KERNEL && (value == 0 || value == 1)) S.kfine_kernel = value;
        else if (key == TG_TUNE_KFINE_ZCHUNK && value >= 1 && value <= 4096) S.kfine_zchunk = value;
        else if (key == TG_TUNE_KELEM_PARTS && (value == 1 || value == 4 || value == 6 || value == 8 || value == 12 || value == 24))
            S.kelem_parts = value;
        else if (key == TG_TUNE_COARSE_GRAPH && (value == 0 || value == 1)) S.coarse.use_graph = value != 0;
        else if (key == TG_TUNE_COARSE_PERSISTENT && (value == 0 || value == 1)) S.coarse.use_persistent = value != 0;
        else if (key == TG_TUNE_COARSE_STREAMING && (value == 0 || value == 1)) S.coarse.pers_streaming = value != 0;
        else if (key == TG_TUNE_COARSE_POLLS && value >= 1) S.coarse.pers_polls = (unsigned)value;
        else if (key == TG_TUNE_KFINE_MFMA && value >= 0 && value <= 3) S.kfine_mfma = value;
        else if (key == TG_TUNE_TIME_KFINE && value >= 0 && value <= 2) {
            S.time_fine = value;
            S.tused = 0;
        } else fail(TG_ERR_ARGS, id, "unknown tuning key or value out of range");
    });
}

int tg_solver_set_jacobi_weight(tg_solver *s, double omega) {
    return guard("tg_solver_set_jacobi_weight:Input", [&] {
        const char *id = "tg_solver_set_jacobi_weight:Input";
        Solver &S = S_of(s, id);
        if (!(omega > 0.0 && omega < 2.0)) fail(TG_ERR_ARGS, id, "omega must lie in (0, 2), got %g", omega);
        S.omega = omega;
    });
}

int tg_solver_kfine_timing(tg_solver *s, int32_t precision, int32_t epilogue, int64_t *launches, double *ms_total) {
    return guard("tg_solver_kfine_timing:Input", [&] {
        const char *id = "tg_solver_kfine_timing:Input";
        Solver &S = S_of(s, id);
        if (!launches || !ms_total || (precision != 0 && precision != 1) || epilogue < 0 || epilogue > 3)
            fail(TG_ERR_ARGS, id, "bad arguments");
        TG_HIP(hipStreamSynchronize(S.st), id);
        const int key = 4 * precision + epilogue;
        int64_t cnt = 0;
        double tot = 0;
        for (size_t i = 0; i < S.tused; ++i) {
            if (S.tpool[i].key != key) continue;
            float ms = 0;
            TG_HIP(hipEventElapsedTime(&ms, S.tpool[i].ev[0], S.tpool[i].ev[1]), id);
            tot += ms;
            ++cnt;
        }
        *launches = cnt;
        *ms_total = tot;
    });
}

// Average duration of one K*u kernel launch at `level` (precision 0: FP64, 1: FP32 V-cycle copy),
// timed with HIP events around `reps` back-to-back launches on the solver's stream.
int tg_solver_time_kbyu(tg_solver *s, int32_t level, int32_t precision, int32_t reps, double *ms_per_launch) {
    return guard("tg_solver_time_kbyu:Input", [&] {
        const char *id = "tg_solver_time_kbyu:Input";
        Solver &S = S_of(s, id);
        const int i = level_index(S, level, id);
        if (!S.assembled) fail(TG_ERR_STATE, "tg_solver_time_kbyu:State", "assemble first");
        if (reps < 1 || !ms_per_launch) fail(TG_ERR_ARGS, id, "bad arguments");
        if (precision == 1 && S.prec != TG_PREC_MIXED) fail(TG_ERR_ARGS, id, "FP32 operators exist in mixed mode only");
        if (i == S.L - 1 && precision == 1) fail(TG_ERR_ARGS, id, "the coarsest level has no FP32 operator");
        DevLevel &D = S.lv[i];
        hipEvent_t a, b;
        TG_HIP(hipEventCreate(&a), id);
        TG_HIP(hipEventCreate(&b), id);
        if (precision == 1) {
            S.kbyu_level<float>(i, D.x32, D.t32, EPI_KU, nullptr, nullptr);
            TG_HIP(hipEventRecord(a, S.st), id);
            for (int k = 0; k < reps; ++k) S.kbyu_level<float>(i, D.x32, D.t32, EPI_KU, nullptr, nullptr);
        } else {
            S.kbyu_level<double>(i, D.x64, D.t64, EPI_KU, nullptr, nullptr);
            TG_HIP(hipEventRecord(a, S.st), id);
            for (int k = 0; k < reps; ++k) S.kbyu_level<double>(i, D.x64, D.t64, EPI_KU, nullptr, nullptr);
        }
        TG_HIP(hipEventRecord(b, S.st), id);
        TG_HIP(hipEventSynchronize(b), id);
        float ms = 0;
        TG_HIP(hipEventElapsedTime(&ms, a, b), id);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *ms_per_launch = ms / reps;
    });
}

}  // extern "C"
