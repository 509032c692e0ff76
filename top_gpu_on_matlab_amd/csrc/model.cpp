// Host model: voxel discretisation, mesh hierarchy, boundary conditions, BC-modified element
// matrices.  Builds every index map of TOP_GPU.m with direct grid arithmetic (no sorting);
// the maps are bit-identical to the reference's unique/intersect/find construction.
#include "model.h"

#include "kunit.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace tg {

// Vertex offsets (dy, dx, dz) of the eNodMat pattern (TOP_GPU.m:2083-2088): n1=(1,0,0),
// n2=(1,1,0), n3=(0,1,0), n4=(0,0,0), n5..n8 the same one layer up.
const int VY[8] = {1, 1, 0, 0, 1, 1, 0, 0};
const int VX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
const int VZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

static const double SXn[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
static const double SYn[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
static const double SZn[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// B^T D B * detJ of a unit-cube Hex8 at one natural point (TOP_GPU.m:2543-2588), column-major 24x24.
static void bdb_at(double xi, double eta, double zeta, double nu, double *K) {
    double B[6][24] = {};
    for (int a = 0; a < 8; ++a) {
        double dxi = 0.125 * SXn[a] * (1 + SYn[a] * eta) * (1 + SZn[a] * zeta);
        double deta = 0.125 * SYn[a] * (1 + SXn[a] * xi) * (1 + SZn[a] * zeta);
        double dzet = 0.125 * SZn[a] * (1 + SXn[a] * xi) * (1 + SYn[a] * eta);
        double dx = 2 * dxi, dy = 2 * deta, dz = 2 * dzet;
        int iu = 3 * a, iv = iu + 1, iw = iu + 2;
        B[0][iu] = dx;
        B[1][iv] = dy;
        B[2][iw] = dz;
        B[3][iv] = dz; B[3][iw] = dy;
        B[4][iu] = dz; B[4][iw] = dx;
        B[5][iu] = dy; B[5][iv] = dx;
    }
    double c = 1.0 / ((1 + nu) * (1 - 2 * nu));
    double D[6][6] = {};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D[i][j] = c * (i == j ? 1 - nu : nu);
    for (int i = 3; i < 6; ++i) D[i][i] = c * (1 - 2 * nu) / 2;
    double DB[6][24];
    for (int i = 0; i < 6; ++i)
        for (int k = 0; k < 24; ++k) {
            double s = 0;
            for (int j = 0; j < 6; ++j) s += D[i][j] * B[j][k];
            DB[i][k] = s;
        }
    for (int col = 0; col < 24; ++col)
        for (int row = 0; row < 24; ++row) {
            double s = 0;
            for (int i = 0; i < 6; ++i) s += B[i][row] * DB[i][col];
            K[row + 24 * col] = s * 0.125;
        }
}

// Ke = cellSize * (2x2x2 Gauss quadrature of B^T D B detJ, kunit.h): equal to the closed form of
// FEA_VoxelBasedElementStiffnessMatrix (TOP_GPU.m:1465-1503), the integrand being exactly
// integrated by the 2-point rule.
void element_stiffness(double nu, double cell, double *Ke) {
    const KUnit K = make_kunit(nu);
    for (int k = 0; k < 576; ++k) Ke[k] = K.v[k] * cell;
}

// computeK_ij (TOP_GPU.m:2589-2608): sub-voxel q = y + num*(x + num*z), sampled at the
// sub-voxel centre with y flipped, scaled by 8*cellSize/num^3.
void super_element_matrices(int num, double cell, double nu, double *out) {
    std::vector<double> span(num);
    for (int k = 0; k < num; ++k) span[k] = -1.0 + (k + 1 - 0.5) / num * 2.0;
    int eid = 0;
    for (int z = 0; z < num; ++z)
        for (int x = 0; x < num; ++x)
            for (int y = 0; y < num; ++y) {
                double *K = out + 576 * (size_t)eid;
                bdb_at(span[x], span[num - 1 - y], span[z], nu, K);
                for (int k = 0; k < 576; ++k) K[k] = K[k] / (num * num * num) * 8 * cell;
                ++eid;
            }
}

// iKEF = r^2*KEF0 + KEF1 of the nodal Helmholtz filter (TOP_GPU.m:889-920), 8x8 column-major.
void pde_filter_kernel(double radius, double ele_size, double *K) {
    double g = 1.0 / std::sqrt(3.0);
    double N[8][8], dS[24][8];
    for (int q = 0; q < 8; ++q) {
        double s = SXn[q] * g, t = SYn[q] * g, p = SZn[q] * g;
        for (int a = 0; a < 8; ++a) {
            N[q][a] = 0.125 * (1 + SXn[a] * s) * (1 + SYn[a] * t) * (1 + SZn[a] * p);
            dS[3 * q][a] = 0.125 * SXn[a] * (1 + SYn[a] * t) * (1 + SZn[a] * p);
            dS[3 * q + 1][a] = 0.125 * SYn[a] * (1 + SXn[a] * s) * (1 + SZn[a] * p);
            dS[3 * q + 2][a] = 0.125 * SZn[a] * (1 + SXn[a] * s) * (1 + SYn[a] * t);
        }
    }
    double r = (radius * ele_size) / 2 / std::sqrt(3.0);
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            double k0 = 0, k1 = 0;
            for (int i = 0; i < 24; ++i) k0 += dS[i][a] * dS[i][b];
            for (int q = 0; q < 8; ++q) k1 += N[q][a] * (1.0 / 8.0) * N[q][b];
            K[a + 8 * b] = r * r * k0 + k1;
        }
}

double ri_weight(int s, int a, int b, int c, int j) {
    double ty = (double)a / s, tx = (double)b / s, tz = (double)c / s;
    double wy = VY[j] ? ty : 1.0 - ty;
    double wx = VX[j] ? tx : 1.0 - tx;
    double wz = VZ[j] ? tz : 1.0 - tz;
    return wy * wx * wz;
}

// Node/element numbering of one level from its full-grid solid mask (TOP_GPU.m:2072-2116, 1388-1419).
static void discretize(HostLevel &L, bool finest) {
    const int64_t ny = L.resY, nx = L.resX, nz = L.resZ;
    const int64_t NY = ny + 1, NX = nx + 1;
    const int64_t nE = L.elemsFull(), nN = L.nodesFull();
    L.eleMapForward.assign(nE, 0);
    L.eleMapBack.clear();
    for (int64_t e = 0; e < nE; ++e)
        if (L.solid[e]) {
            L.eleMapBack.push_back((int32_t)(e + 1));
            L.eleMapForward[e] = (int32_t)L.eleMapBack.size();
        }
    L.Ne = (int64_t)L.eleMapBack.size();
    std::vector<uint8_t> active(nN, 0);
    std::vector<int32_t> cnt(nN, 0);
    for (int64_t k = 0; k < L.Ne; ++k) {
        int64_t e = L.eleMapBack[k] - 1;
        int64_t ey = e % ny, ex = (e / ny) % nx, ez = e / (ny * nx);
        for (int j = 0; j < 8; ++j) {
            int64_t n = (ey + VY[j]) + NY * ((ex + VX[j]) + NX * (ez + VZ[j]));
            active[n] = 1;
            ++cnt[n];
        }
    }
    L.nodMapForward.assign(nN, 0);
    L.nodFull.clear();
    for (int64_t n = 0; n < nN; ++n)
        if (active[n]) {
            L.nodFull.push_back((int32_t)(n + 1));
            L.nodMapForward[n] = (int32_t)L.nodFull.size();
        }
    L.Nn = (int64_t)L.nodFull.size();
    L.eNodMat.assign(L.Ne * 8, 0);
    for (int64_t k = 0; k < L.Ne; ++k) {
        int64_t e = L.eleMapBack[k] - 1;
        int64_t ey = e % ny, ex = (e / ny) % nx, ez = e / (ny * nx);
        for (int j = 0; j < 8; ++j) {
            int64_t n = (ey + VY[j]) + NY * ((ex + VX[j]) + NX * (ez + VZ[j]));
            L.eNodMat[k + L.Ne * j] = L.nodMapForward[n];
        }
    }
    L.nodesOnBoundary.clear();
    std::vector<uint8_t> onb(L.Nn + 1, 0);
    for (int64_t i = 0; i < L.Nn; ++i)
        if (cnt[L.nodFull[i] - 1] < 8) {
            L.nodesOnBoundary.push_back((int32_t)(i + 1));
            onb[i + 1] = 1;
        }
    L.elementsOnBoundary.clear();
    if (finest) {
        for (int64_t k = 0; k < L.Ne; ++k) {
            bool b = false;
            for (int j = 0; j < 8; ++j) b = b || onb[L.eNodMat[k + L.Ne * j]];
            if (b) L.elementsOnBoundary.push_back((int32_t)(k + 1));
        }
    }
    L.nodMapBack = L.nodFull;
}

static int matlab_round_pos(double x) { return (int)std::floor(x + 0.5); }

static void build(tg_model &m, const uint8_t *vox) {
    const tg_params &p = m.prm;
    int64_t numVoxels = 0;
    const int64_t nIn = (int64_t)m.nely * m.nelx * m.nelz;
    for (int64_t i = 0; i < nIn; ++i) numVoxels += vox[i] ? 1 : 0;
    if (numVoxels == 0) fail(TG_ERR_ARGS, "CreateVoxelFEAmodel:Input", "voxel model has no solid voxel");
    int numLevels = 0;
    // FEA_VoxelBasedDiscretization (TOP_GPU.m:2045-2054)
    while (numVoxels >= p.coarsest_resolution_control) {
        ++numLevels;
        numVoxels = (int64_t)std::floor((double)numVoxels / 8 + 0.5);
    }
    numLevels = std::max(3, numLevels);
    const int64_t f = 1LL << numLevels;
    auto adj = [&](int64_t n) { return (int32_t)(((n + f - 1) / f) * f); };
    HostLevel L1;
    L1.resX = adj(m.nelx);
    L1.resY = adj(m.nely);
    L1.resZ = adj(m.nelz);
    numLevels += 1;
    L1.solid.assign(L1.elemsFull(), 0);
    for (int64_t z = 0; z < m.nelz; ++z)
        for (int64_t x = 0; x < m.nelx; ++x)
            for (int64_t y = 0; y < m.nely; ++y)
                L1.solid[y + (int64_t)L1.resY * (x + (int64_t)L1.resX * z)] =
                    vox[y + (int64_t)m.nely * (x + (int64_t)m.nelx * z)] ? 1 : 0;
    discretize(L1, true);
    m.lv.clear();
    m.lv.push_back(std::move(L1));

    // Solving_BuildingMeshHierarchy (TOP_GPU.m:1276-1434)
    int nonDyadic = p.non_dyadic;
    if (nonDyadic == 1 && numLevels >= 4) numLevels -= 1;
    else nonDyadic = 0;
    m.numLevels = numLevels;
    m.nonDyadic = nonDyadic;
    const int64_t NY1 = m.lv[0].NY(), NX1 = m.lv[0].NX();
    for (int ii = 2; ii <= numLevels; ++ii) {
        const HostLevel &F = m.lv.back();
        const int s = (ii == 2 && nonDyadic == 1) ? 4 : 2;
        HostLevel C;
        C.span = s;
        C.resX = F.resX / s;
        C.resY = F.resY / s;
        C.resZ = F.resZ / s;
        const int64_t ny = C.resY, nx = C.resX, nz = C.resZ;
        const int64_t fny = F.resY, fnx = F.resX;
        const int64_t FNY = F.NY(), FNX = F.NX();
        const int s3 = s * s * s, nb = (s + 1) * (s + 1) * (s + 1);
        C.solid.assign(C.elemsFull(), 0);
        for (int64_t cz = 0; cz < nz; ++cz)
            for (int64_t cx = 0; cx < nx; ++cx)
                for (int64_t cy = 0; cy < ny; ++cy) {
                    bool any = false;
                    for (int c = 0; c < s && !any; ++c)
                        for (int b = 0; b < s && !any; ++b)
                            for (int a = 0; a < s && !any; ++a)
                                any = F.solid[(s * cy + a) + fny * ((s * cx + b) + fnx * (s * cz + c))] != 0;
                    C.solid[cy + ny * (cx + nx * cz)] = any ? 1 : 0;
                }
        // intermediate fine nodes: every fine node of the closed blocks of non-empty coarse cells
        std::vector<int32_t> cover(F.nodesFull(), 0);
        for (int64_t cz = 0; cz < nz; ++cz)
            for (int64_t cx = 0; cx < nx; ++cx)
                for (int64_t cy = 0; cy < ny; ++cy) {
                    if (!C.solid[cy + ny * (cx + nx * cz)]) continue;
                    for (int c = 0; c <= s; ++c)
                        for (int b = 0; b <= s; ++b)
                            for (int a = 0; a <= s; ++a)
                                ++cover[(s * cy + a) + FNY * ((s * cx + b) + FNX * (s * cz + c))];
                }
        C.fineInterId.assign(F.nodesFull(), 0);
        C.coeff.clear();
        for (int64_t n = 0; n < F.nodesFull(); ++n)
            if (cover[n]) {
                C.coeff.push_back((double)cover[n]);
                C.fineInterId[n] = (int32_t)C.coeff.size();
            }
        C.nInt = (int64_t)C.coeff.size();
        discretize(C, false);
        C.upMap.assign(C.Ne * s3, 0);
        C.transferMat.assign((int64_t)nb * C.Ne, 0);
        for (int64_t k = 0; k < C.Ne; ++k) {
            int64_t e = C.eleMapBack[k] - 1;
            int64_t cy = e % ny, cx = (e / ny) % nx, cz = e / (ny * nx);
            for (int c = 0; c < s; ++c)
                for (int b = 0; b < s; ++b)
                    for (int a = 0; a < s; ++a)
                        C.upMap[k + C.Ne * (a + s * b + s * s * c)] =
                            F.eleMapForward[(s * cy + a) + fny * ((s * cx + b) + fnx * (s * cz + c))];
            for (int c = 0; c <= s; ++c)
                for (int b = 0; b <= s; ++b)
                    for (int a = 0; a <= s; ++a)
                        C.transferMat[(a + (s + 1) * b + (s + 1) * (s + 1) * c) + (int64_t)nb * k] =
                            C.fineInterId[(s * cy + a) + FNY * ((s * cx + b) + FNX * (s * cz + c))];
        }
        C.solidNodeMap.assign(F.Nn, 0);
        for (int64_t j = 0; j < F.Nn; ++j) C.solidNodeMap[j] = C.fineInterId[F.nodFull[j] - 1];
        // reference nodMapBack: finest-grid node number of each coarse node (TOP_GPU.m:1404-1407)
        const int kk = (nonDyadic == 1) ? ii : ii - 1;
        const int64_t stride = 1LL << kk;
        for (int64_t j = 0; j < C.Nn; ++j) {
            int64_t n = C.nodFull[j] - 1;
            int64_t y = n % C.NY(), x = (n / C.NY()) % C.NX(), z = n / (C.NY() * C.NX());
            C.nodMapBack[j] = (int32_t)(1 + stride * y + NY1 * (stride * x + NX1 * stride * z));
        }
        m.lv.push_back(std::move(C));
    }
    m.Ke.assign(576, 0.0);
    element_stiffness(p.poisson_ratio, p.cell_size, m.Ke.data());
    if (p.super_element) {
        m.Imat.assign(576 * 64, 0.0);
        super_element_matrices(4, p.cell_size, p.poisson_ratio, m.Imat.data());
    }
}

// ---- boundary conditions ------------------------------------------------------------

static void builtin_case(tg_model &m, int opt) {
    const HostLevel &L = m.lv[0];
    const int64_t NY = L.NY(), NX = L.NX(), NZ = L.NZ();
    auto node = [&](int64_t y, int64_t x, int64_t z) { return y + NY * (x + NX * z); };
    std::vector<uint8_t> vol(NY * NX * NZ, 0);
    std::vector<double> fix, ld;
    auto collect = [&](std::vector<int64_t> &out) {
        out.clear();
        for (int64_t n = 0; n < (int64_t)vol.size(); ++n)
            if (vol[n]) out.push_back(L.nodMapForward[n]);
    };
    std::vector<int64_t> fnodes, lnodes;
    std::vector<double> force;
    if (opt == 1) {  // cantilever (TOP_GPU.m:507-516)
        for (int64_t z = 0; z <= m.nelz; ++z)
            for (int64_t y = 0; y <= m.nely; ++y) vol[node(y, 0, z)] = 1;
        collect(fnodes);
        std::fill(vol.begin(), vol.end(), 0);
        int zl = matlab_round_pos(m.nelz / 6.0 + 1);
        for (int64_t z = 0; z < zl; ++z)
            for (int64_t y = 0; y <= m.nely; ++y) vol[node(y, m.nelx, z)] = 1;
        collect(lnodes);
        double q = -1.0 / (double)lnodes.size();
        force = {0.0 / lnodes.size(), 0.0 / lnodes.size(), q};
    } else if (opt == 2) {  // MBB (TOP_GPU.m:517-529)
        vol[node(0, 0, 0)] = vol[node(0, m.nelx, 0)] = vol[node(m.nely, 0, 0)] = vol[node(m.nely, m.nelx, 0)] = 1;
        collect(fnodes);
        std::fill(vol.begin(), vol.end(), 0);
        vol[node(m.nely / 2, m.nelx / 2, m.nelz)] = 1;
        collect(lnodes);
        force = {0.0, 0.0, -1.0};
    } else {
        fail(TG_ERR_ARGS, "CreateVoxelFEAmodel:optCase", "unknown optCase %d", opt);
    }
    const int64_t nf = (int64_t)fnodes.size(), nl = (int64_t)lnodes.size();
    m.fixing.assign(nf * 4, 1.0);
    for (int64_t i = 0; i < nf; ++i) m.fixing[i] = (double)fnodes[i];
    m.nFix = nf;
    std::vector<double> lc(nl * 4, 0.0);
    for (int64_t i = 0; i < nl; ++i) {
        lc[i] = (double)lnodes[i];
        for (int c = 0; c < 3; ++c) lc[i + nl * (1 + c)] = force[c];
    }
    m.loads.assign(1, lc);
    m.finalized = false;
}

// first-occurrence unique of rows keyed by column 0, sorted by key
static std::vector<double> unique_rows(const std::vector<double> &t, int64_t n, int64_t &nout) {
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return t[a] < t[b]; });
    std::vector<int64_t> keep;
    for (int64_t k = 0; k < n; ++k)
        if (keep.empty() || t[idx[k]] != t[keep.back()]) keep.push_back(idx[k]);
    nout = (int64_t)keep.size();
    std::vector<double> out(nout * 4);
    for (int64_t r = 0; r < nout; ++r)
        for (int c = 0; c < 4; ++c) out[r + nout * c] = t[keep[r] + n * c];
    return out;
}

static void finalize(tg_model &m) {
    if (m.loads.empty() || m.nFix == 0)
        fail(TG_ERR_STATE, "FEA_ApplyBoundaryCondition:State", "no fixations or loads set");
    HostLevel &L = m.lv[0];
    const int64_t Nn = L.Nn, Ne = L.Ne;
    // FEA_ApplyBoundaryCondition (TOP_GPU.m:1511-1524)
    std::vector<double> fixing = m.fixing;
    int64_t nFix = m.nFix;
    for (auto &ld : m.loads) {
        int64_t nl = (int64_t)ld.size() / 4;
        std::vector<double> keys(ld.begin(), ld.begin() + nl);
        std::sort(keys.begin(), keys.end());
        std::vector<double> kept;
        std::vector<int64_t> rows;
        for (int64_t r = 0; r < nFix; ++r)
            if (!std::binary_search(keys.begin(), keys.end(), fixing[r])) rows.push_back(r);
        std::vector<double> t(rows.size() * 4);
        for (size_t r = 0; r < rows.size(); ++r)
            for (int c = 0; c < 4; ++c) t[r + rows.size() * c] = fixing[rows[r] + nFix * c];
        int64_t n2;
        fixing = unique_rows(t, (int64_t)rows.size(), n2);
        nFix = n2;
        int64_t nl2;
        ld = unique_rows(ld, nl, nl2);
        if (nl2 == 0) fail(TG_ERR_STATE, "FEA_ApplyBoundaryCondition:State", "No Loads!");
        if (nFix == 0) fail(TG_ERR_STATE, "FEA_ApplyBoundaryCondition:State", "No Fixations!");
    }
    m.fixing = fixing;
    m.nFix = nFix;
    const int64_t nCases = (int64_t)m.loads.size();
    m.F.assign(3 * Nn * nCases, 0.0);
    for (int64_t ic = 0; ic < nCases; ++ic) {
        const auto &ld = m.loads[ic];
        int64_t nl = (int64_t)ld.size() / 4;
        for (int64_t r = 0; r < nl; ++r) {
            int64_t n = (int64_t)ld[r] - 1;
            if (n < 0 || n >= Nn) fail(TG_ERR_ARGS, "FEA_ApplyBoundaryCondition:Input", "load node %lld out of range", (long long)(n + 1));
            for (int c = 0; c < 3; ++c) m.F[3 * n + c + 3 * Nn * ic] = ld[r + nl * (1 + c)];
        }
    }
    m.fixedDOF.assign(3 * Nn, 0);
    for (int64_t r = 0; r < nFix; ++r) {
        int64_t n = (int64_t)fixing[r] - 1;
        if (n < 0 || n >= Nn) fail(TG_ERR_ARGS, "FEA_ApplyBoundaryCondition:Input", "fixed node %lld out of range", (long long)(n + 1));
        for (int c = 0; c < 3; ++c)
            if (fixing[r + nFix * (1 + c)] == 1.0) m.fixedDOF[3 * n + c] = 1;
    }
    // Solving_SetupKeWithFixedDOFs (TOP_GPU.m:1565-1620 / 1647-1700)
    std::vector<int32_t> bpos(Nn + 1, 0);
    for (size_t k = 0; k < L.nodesOnBoundary.size(); ++k) bpos[L.nodesOnBoundary[k]] = (int32_t)(k + 1);
    std::vector<std::vector<int32_t>> lists(L.nodesOnBoundary.size());
    for (int32_t e1 : L.elementsOnBoundary)
        for (int j = 0; j < 8; ++j) {
            int32_t b = bpos[L.eNodMat[(e1 - 1) + Ne * j]];
            if (b) lists[b - 1].push_back(e1);
        }
    std::vector<int32_t> fnodes(nFix);
    for (int64_t r = 0; r < nFix; ++r) fnodes[r] = (int32_t)fixing[r];
    std::vector<int32_t> sortedF = fnodes;
    std::sort(sortedF.begin(), sortedF.end());
    sortedF.erase(std::unique(sortedF.begin(), sortedF.end()), sortedF.end());
    std::vector<int32_t> fixedStructIdx;  // boundary positions of fixed boundary nodes, ascending node id
    for (int32_t nb : L.nodesOnBoundary)
        if (std::binary_search(sortedF.begin(), sortedF.end(), nb)) fixedStructIdx.push_back(bpos[nb] - 1);
    std::vector<uint8_t> isFixE(Ne + 1, 0);
    for (int32_t k : fixedStructIdx)
        for (int32_t e1 : lists[k]) isFixE[e1] = 1;
    m.mapUniqueKes.assign(Ne, 0);
    m.eleWithFixed.clear();
    for (int64_t e = 1; e <= Ne; ++e)
        if (isFixE[e]) {
            m.eleWithFixed.push_back((int32_t)e);
            m.mapUniqueKes[e - 1] = (int32_t)m.eleWithFixed.size();
        }
    const int64_t nU = (int64_t)m.eleWithFixed.size();
    const bool se = m.prm.super_element != 0;
    const int64_t q = se ? 64 : 1;
    const std::vector<double> &base = se ? m.Imat : m.Ke;
    m.uniqueFree.assign(576 * q * nU, 0.0);
    for (int64_t u = 0; u < nU; ++u)
        std::copy(base.begin(), base.begin() + 576 * q, m.uniqueFree.begin() + 576 * q * u);
    m.uniqueFixed.assign(576 * nU, 0.0);
    m.nodeFixMask.assign(Nn, 0);
    for (size_t ii = 0; ii < fixedStructIdx.size(); ++ii) {
        const auto &arr = lists[fixedStructIdx[ii]];
        const int32_t bnode = L.nodesOnBoundary[fixedStructIdx[ii]];
        const int32_t iNode = (int32_t)fixing[ii];  // pairing as written (TOP_GPU.m:1597-1598)
        if (iNode != bnode)
            fail(TG_ERR_ARGS, "Solving_SetupKeWithFixedDOFs:Pairing",
                 "fixation %zu (node %d) pairs with boundary node %d: every fixed node must lie on the "
                 "domain boundary and fixations must be sorted (TOP_GPU.m:1587-1604)", ii + 1, iNode, bnode);
        const double nAdj = (double)arr.size();
        uint8_t mask = 0;
        for (int c = 0; c < 3; ++c)
            if (fixing[ii + nFix * (1 + c)] == 1.0) mask |= (uint8_t)(1 << c);
        m.nodeFixMask[iNode - 1] |= mask;
        for (int32_t e1 : arr) {
            int loc = -1;
            for (int j = 0; j < 8; ++j)
                if (L.eNodMat[(e1 - 1) + Ne * j] == iNode) loc = j;
            const int64_t u = m.mapUniqueKes[e1 - 1] - 1;
            for (int c = 0; c < 3; ++c) {
                if (!(mask & (1 << c))) continue;
                const int d = 3 * loc + c;
                for (int64_t qq = 0; qq < q; ++qq) {
                    double *K = m.uniqueFree.data() + 576 * (qq + q * u);
                    for (int k = 0; k < 24; ++k) {
                        K[d + 24 * k] = 0.0;
                        K[k + 24 * d] = 0.0;
                    }
                }
                m.uniqueFixed[d + 24 * d + 576 * u] = 1.0 / nAdj;
            }
        }
    }
    for (auto &v : m.uniqueFixed) v *= m.prm.youngs_modulus;
    m.finalized = true;
}

}  // namespace tg

using namespace tg;

extern "C" {

void tg_params_default(tg_params *p) {
    if (!p) return;
    p->youngs_modulus = 1.0;
    p->poisson_ratio = 0.3;
    p->cell_size = 1.0;
    p->weight_factor_jacobi = 0.6;
    p->coarsest_resolution_control = 20000;
    p->non_dyadic = 1;
    p->vcycle_type = 0;
    p->super_element = 0;
    p->precision = TG_PREC_FP64;
}

int tg_model_create(const uint8_t *voxels, int32_t nely, int32_t nelx, int32_t nelz, const tg_params *params,
                    tg_model **out) {
    return guard("tg_model_create:Input", [&] {
        if (!voxels || !out) fail(TG_ERR_ARGS, "tg_model_create:Input", "null argument");
        if (nelx < 3 || nely < 3 || nelz < 3)
            fail(TG_ERR_ARGS, "CreateVoxelFEAmodel:Input", "Inappropriate Input Model! (at least 3 layers per axis)");
        auto *m = new tg_model();
        if (params) m->prm = *params;
        else tg_params_default(&m->prm);
        m->nely = nely;
        m->nelx = nelx;
        m->nelz = nelz;
        try {
            build(*m, voxels);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

void tg_model_destroy(tg_model *m) { delete m; }

int tg_model_builtin_case(tg_model *m, int32_t opt_case) {
    return guard("tg_model_builtin_case:Input", [&] {
        if (!m) fail(TG_ERR_ARGS, "tg_model_builtin_case:Input", "null model");
        builtin_case(*m, opt_case);
    });
}

int tg_model_set_conditions(tg_model *m, const double *fixing, int64_t n_fix, const double *loads, int64_t n_load) {
    return guard("tg_model_set_conditions:Input", [&] {
        if (!m || (!fixing && n_fix) || (!loads && n_load)) fail(TG_ERR_ARGS, "tg_model_set_conditions:Input", "null argument");
        m->fixing.assign(fixing, fixing + 4 * n_fix);
        m->nFix = n_fix;
        m->loads.assign(1, std::vector<double>(loads, loads + 4 * n_load));
        m->finalized = false;
    });
}

int tg_model_add_loads(tg_model *m, const double *loads, int64_t n_load) {
    return guard("tg_model_add_loads:Input", [&] {
        if (!m || (!loads && n_load) || n_load < 0) fail(TG_ERR_ARGS, "tg_model_add_loads:Input", "null argument");
        m->loads.emplace_back(loads, loads + 4 * n_load);
        m->finalized = false;
    });
}

int tg_model_finalize(tg_model *m) {
    return guard("tg_model_finalize:Input", [&] {
        if (!m) fail(TG_ERR_ARGS, "tg_model_finalize:Input", "null model");
        finalize(*m);
    });
}

int32_t tg_model_num_levels(const tg_model *m) { return m ? m->numLevels : 0; }

int tg_model_level_info(const tg_model *m, int32_t level, tg_level_info *info) {
    return guard("tg_model_level_info:Input", [&] {
        if (!m || !info || level < 1 || level > m->numLevels) fail(TG_ERR_ARGS, "tg_model_level_info:Input", "bad level");
        const HostLevel &L = m->lv[level - 1];
        info->resX = L.resX;
        info->resY = L.resY;
        info->resZ = L.resZ;
        info->spanWidth = L.span;
        info->numElements = L.Ne;
        info->numNodes = L.Nn;
        info->numDOFs = 3 * L.Nn;
        info->intermediateNumNodes = L.nInt;
    });
}

static const std::vector<int32_t> *map_ptr(const tg_model *m, int32_t level, int32_t which) {
    if (!m || level < 1 || level > m->numLevels) return nullptr;
    const HostLevel &L = m->lv[level - 1];
    switch (which) {
        case TG_MAP_ENODMAT: return &L.eNodMat;
        case TG_MAP_ELEMAPBACK: return &L.eleMapBack;
        case TG_MAP_ELEMAPFORWARD: return &L.eleMapForward;
        case TG_MAP_NODMAPBACK: return &L.nodMapBack;
        case TG_MAP_NODMAPFORWARD: return &L.nodMapForward;
        case TG_MAP_NODESONBOUNDARY: return &L.nodesOnBoundary;
        case TG_MAP_ELEMENTSONBOUNDARY: return level == 1 ? &L.elementsOnBoundary : nullptr;
        case TG_MAP_ELEMENTUPWARDMAP: return level >= 2 ? &L.upMap : nullptr;
        case TG_MAP_TRANSFERMAT: return level >= 2 ? &L.transferMat : nullptr;
        case TG_MAP_SOLIDNODEMAP: return level >= 2 ? &L.solidNodeMap : nullptr;
        case TG_MAP_MAPUNIQUEKES: return (level == 1 && m->finalized) ? &m->mapUniqueKes : nullptr;
        case TG_MAP_ELEWITHFIXEDDOFS: return (level == 1 && m->finalized) ? &m->eleWithFixed : nullptr;
        default: return nullptr;
    }
}

int64_t tg_model_map_size(const tg_model *m, int32_t level, int32_t which) {
    const auto *v = map_ptr(m, level, which);
    return v ? (int64_t)v->size() : -1;
}

int tg_model_get_map(const tg_model *m, int32_t level, int32_t which, int32_t *out) {
    return guard("tg_model_get_map:Input", [&] {
        const auto *v = map_ptr(m, level, which);
        if (!v || !out) fail(TG_ERR_ARGS, "tg_model_get_map:Input", "map %d not available at level %d", which, level);
        std::copy(v->begin(), v->end(), out);
    });
}

static std::vector<double> table(const tg_model *m, int32_t level, int32_t which, bool &ok) {
    ok = true;
    if (!m || level < 1 || level > m->numLevels) { ok = false; return {}; }
    const HostLevel &L = m->lv[level - 1];
    switch (which) {
        case TG_TAB_TRANSFERCOEFF: if (level >= 2) return L.coeff; break;
        case TG_TAB_KE: return m->Ke;
        case TG_TAB_LOADS: if (m->finalized) return m->F; break;
        case TG_TAB_UNIQUEKESFREE: if (m->finalized) return m->uniqueFree; break;
        case TG_TAB_UNIQUEKESFIXED: if (m->finalized) return m->uniqueFixed; break;
        case TG_TAB_IMAT: if (m->prm.super_element) return m->Imat; break;
        case TG_TAB_FIXEDDOFS:
            if (m->finalized) return std::vector<double>(m->fixedDOF.begin(), m->fixedDOF.end());
            break;
        default: break;
    }
    ok = false;
    return {};
}

int64_t tg_model_table_size(const tg_model *m, int32_t level, int32_t which) {
    bool ok;
    auto t = table(m, level, which, ok);
    return ok ? (int64_t)t.size() : -1;
}

int tg_model_get_table(const tg_model *m, int32_t level, int32_t which, double *out) {
    return guard("tg_model_get_table:Input", [&] {
        bool ok;
        auto t = table(m, level, which, ok);
        if (!ok || !out) fail(TG_ERR_ARGS, "tg_model_get_table:Input", "table %d not available at level %d", which, level);
        std::copy(t.begin(), t.end(), out);
    });
}

}  // extern "C"
