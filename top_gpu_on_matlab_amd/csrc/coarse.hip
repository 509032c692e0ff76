// Coarsest-level direct solver: plane-blocked Cholesky on the device (see coarse.h).
#include "coarse.h"

namespace tg {
namespace {

__device__ inline int coord(int axis, int y, int x, int z) { return axis == 0 ? y : (axis == 1 ? x : z); }

// One lane per (node J, neighbour slot, DOF pair): K_c(J c, M d) = sum over the solid elements
// shared by J and M of Ks_e(3a_J + c, 3b_M + d)  (the entries sparse(iK, jK, sK) accumulates,
// TOP_GPU.m:1249-1254).  Lower blocks only (plane(M) <= plane(J)).
__global__ void k_coarse_assemble(const double *__restrict__ Ks, int64_t ldk, const uint8_t *__restrict__ eleMask,
                                  const uint8_t *__restrict__ nodeMask, const Grid g, int axis,
                                  const int32_t *__restrict__ lidx, const int64_t *__restrict__ off,
                                  const int64_t *__restrict__ adOff, const int64_t *__restrict__ asOff,
                                  double *__restrict__ Ad, double *__restrict__ As) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= g.nodes() * 243) return;
    const int64_t J = tid / 243;
    const int rem = (int)(tid % 243);
    const int k = rem / 9, cd = rem % 9, c = cd % 3, d = cd / 3;
    if (!(nodeMask[J] & NODE_ACTIVE)) return;
    const int jy = (int)(J % g.NY);
    const int64_t t = J / g.NY;
    const int jx = (int)(t % g.NX), jz = (int)(t / g.NX);
    const int dy = k % 3 - 1, dx = (k / 3) % 3 - 1, dz = k / 9 - 1;
    const int my = jy + dy, mx = jx + dx, mz = jz + dz;
    if ((unsigned)my >= (unsigned)g.NY || (unsigned)mx >= (unsigned)g.NX || (unsigned)mz >= (unsigned)g.NZ) return;
    const int64_t M = my + (int64_t)g.NY * (mx + (int64_t)g.NX * mz);
    if (!(nodeMask[M] & NODE_ACTIVE)) return;
    const int pJ = coord(axis, jy, jx, jz), pM = coord(axis, my, mx, mz);
    if (pM > pJ) return;
    double val = 0;
    for (int oz = -1; oz <= 0; ++oz)
        for (int ox = -1; ox <= 0; ++ox)
            for (int oy = -1; oy <= 0; ++oy) {
                const int ey = jy + oy, ex = jx + ox, ez = jz + oz;
                if ((unsigned)ey >= (unsigned)g.EY || (unsigned)ex >= (unsigned)g.EX || (unsigned)ez >= (unsigned)g.EZ)
                    continue;
                const int ty = dy - oy, tx = dx - ox, tz = dz - oz;
                if ((unsigned)ty > 1u || (unsigned)tx > 1u || (unsigned)tz > 1u) continue;
                const int64_t e = ey + (int64_t)g.EY * (ex + (int64_t)g.EX * ez);
                if (!eleMask[e]) continue;
                const int a = vtx_of(-oy, -ox, -oz), b = vtx_of(ty, tx, tz);
                val += Ks[(int64_t)((3 * a + c) + 24 * (3 * b + d)) * ldk + e];
            }
    const int64_t np = off[pJ + 1] - off[pJ];
    const int64_t row = 3 * (int64_t)lidx[J] + c, col = 3 * (int64_t)lidx[M] + d;
    if (pM == pJ) Ad[adOff[pJ] + row + np * col] = val;
    else As[asOff[pJ] + row + np * col] = val;
}

// ---- plane-block Cholesky: blocked right-looking, panels of CH_NB columns --------------------
// Per panel: the kb x kb diagonal block is factorised in LDS by one workgroup, which also forms
// W = inv(L_kk); the rows under it are X = P W^T (k_chol_panel); the trailing matrix is updated by
// k_chol_update.  Column-major, lower triangle; the upper triangle is not referenced.
constexpr int CH_NB = 64;
constexpr int CH_PR = 32;  // panel rows per workgroup

// info: set to col0 + j + 1 (LAPACK convention, first failure only) when pivot j is not positive;
// the factorisation then continues with a unit pivot so every value stays finite.
// Steps of CH_RB columns: every lane factorises the CH_RB x CH_RB diagonal sub-block itself (from LDS,
// in registers), the rows below are scaled (one lane per row), then a rank-CH_RB trailing update;
// the inverse is formed in row blocks of CH_IB with one lane per (row in block, column).  Two barriers
// per step and one per inverse row block.
constexpr int CH_RB = 4;  // columns per factorisation step
constexpr int CH_IB = 4;  // rows per inverse block (CH_IB lanes per column: 4 x 64 = 256)
// Batched over blockIdx.y: problem y works on A + y * dA, W + y * CH_NB^2 and info[y * dinfo] (the two
// elimination chains' planes of one step share a launch).
__global__ __launch_bounds__(256) void k_chol_diag(double *__restrict__ A, int64_t lda, int kb,
                                                   double *__restrict__ W, int *__restrict__ info, int col0,
                                                   int64_t dA, int dinfo) {
    __shared__ double a[CH_NB][CH_NB + 1];  // a[i][j] = A(i, j), factorised in place
    __shared__ double w[CH_NB][CH_NB + 1];  // inv(L_kk)
    __shared__ double rdg[CH_NB];           // 1 / L(j, j)
    const int t = threadIdx.x;
    A += blockIdx.y * dA;
    W += blockIdx.y * (CH_NB * CH_NB);
    info += blockIdx.y * dinfo;
    for (int idx = t; idx < kb * kb; idx += 256) {
        const int i = idx % kb, j = idx / kb;
        a[i][j] = A[i + (int64_t)j * lda];
    }
    // rows/columns kb..CH_NB-1 of the padded block: identity, so every step can work on full blocks
    for (int idx = t; idx < CH_NB * CH_NB; idx += 256) {
        const int i = idx % CH_NB, j = idx / CH_NB;
        if (i >= kb || j >= kb) a[i][j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    int bad = 0;
    for (int j0 = 0; j0 < CH_NB; j0 += CH_RB) {
        // the diagonal sub-block, factorised redundantly by every lane
        // (one division per column; the rest of the step multiplies by the reciprocal pivot)
        double L[CH_RB][CH_RB], rp[CH_RB];
#pragma unroll
        for (int r = 0; r < CH_RB; ++r)
#pragma unroll
            for (int c = 0; c <= r; ++c) L[r][c] = a[j0 + r][j0 + c];
#pragma unroll
        for (int c = 0; c < CH_RB; ++c) {
            double d = L[c][c];
#pragma unroll
            for (int k = 0; k < c; ++k) d -= L[c][k] * L[c][k];
            if (!(d > 0.0)) {
                if (!bad) bad = j0 + c + 1;
                d = 1.0;
            }
            L[c][c] = sqrt(d);
            rp[c] = 1.0 / L[c][c];
#pragma unroll
            for (int r = c + 1; r < CH_RB; ++r) {
                double v = L[r][c];
#pragma unroll
                for (int k = 0; k < c; ++k) v -= L[r][k] * L[c][k];
                L[r][c] = v * rp[c];
            }
        }
        // rows below: x L^T = a(i, j0 .. j0+3)
        const int i = j0 + CH_RB + t;
        if (i < CH_NB) {
            double x[CH_RB];
#pragma unroll
            for (int c = 0; c < CH_RB; ++c) {
                double v = a[i][j0 + c];
#pragma unroll
                for (int k = 0; k < c; ++k) v -= x[k] * L[c][k];
                x[c] = v * rp[c];
            }
#pragma unroll
            for (int c = 0; c < CH_RB; ++c) a[i][j0 + c] = x[c];
        }
        __syncthreads();  // every lane has read the sub-block; the scaled columns are complete
        if (t == 0) {
#pragma unroll
            for (int r = 0; r < CH_RB; ++r) {
#pragma unroll
                for (int c = 0; c <= r; ++c) a[j0 + r][j0 + c] = L[r][c];
                rdg[j0 + r] = rp[r];
            }
        }
        // trailing update of the lower triangle past the step: lane (row ri, column phase cph)
        const int ri = t % CH_NB, cph = t / CH_NB;
        if (ri >= j0 + CH_RB) {
            double ar[CH_RB];
#pragma unroll
            for (int c = 0; c < CH_RB; ++c) ar[c] = a[ri][j0 + c];
            for (int l = j0 + CH_RB + cph; l <= ri; l += 4) {
                double v = a[ri][l];
#pragma unroll
                for (int c = 0; c < CH_RB; ++c) v -= ar[c] * a[l][j0 + c];
                a[ri][l] = v;
            }
        }
        __syncthreads();
    }
    // W = inv(L): row block i0 .. i0+3, lane (r = t % 4, column c = t / 4); the four lanes of a column
    // are adjacent and pass the block's new rows along by shuffles
    const int r = t % CH_IB, c = t / CH_IB;
    for (int i0 = 0; i0 < CH_NB; i0 += CH_IB) {
        const int i = i0 + r;
        double sacc = 0.0, sacc2 = 0.0;  // sum over the finished rows k < i0 of L(i, k) W(k, c)
        int k = c;
        for (; k + 1 < i0; k += 2) {
            sacc += a[i][k] * w[k][c];
            sacc2 += a[i][k + 1] * w[k + 1][c];
        }
        if (k < i0) sacc += a[i][k] * w[k][c];
        sacc += sacc2;
        double wv = 0.0;
#pragma unroll
        for (int rr = 0; rr < CH_IB; ++rr) {
            // row i0 + rr: lane rr finishes it with the block rows above it (held by lanes < rr)
            double add = 0.0;
#pragma unroll
            for (int q = 0; q < rr; ++q) {
                const double wq = __shfl(wv, (t & ~(CH_IB - 1)) + q, 64);
                if (r == rr) add += a[i][i0 + q] * wq;
            }
            if (r == rr) {
                const double rii = rdg[i];
                wv = c > i ? 0.0 : (c == i ? rii : -(sacc + add) * rii);
            }
        }
        w[i][c] = wv;
        __syncthreads();
    }
    for (int idx = t; idx < kb * kb; idx += 256) {
        const int ii = idx % kb, j = idx / kb;
        if (ii >= j) A[ii + (int64_t)j * lda] = a[ii][j];
        W[ii + CH_NB * j] = w[ii][j];
    }
    if (t == 0 && bad && bad <= kb && *info == 0) *info = col0 + bad;
}

// Rows under the diagonal block: X = P W^T for the m x kb panel P (column-major, lda), in place;
// CH_PR rows per workgroup, lane (row r, column group cg) forms 8 outputs.
__global__ __launch_bounds__(256) void k_chol_panel(double *__restrict__ P, int64_t lda, int m, int kb,
                                                    const double *__restrict__ W, int64_t dA) {
    __shared__ double w[CH_NB][CH_NB + 1];
    __shared__ double pr[CH_PR][CH_NB + 1];
    const int t = threadIdx.x;
    P += blockIdx.y * dA;  // batched like k_chol_diag
    W += blockIdx.y * (CH_NB * CH_NB);
    const int row0 = blockIdx.x * CH_PR;
    const int rows = min(CH_PR, m - row0);
    for (int idx = t; idx < kb * kb; idx += 256) {
        const int i = idx % kb, j = idx / kb;
        w[i][j] = W[i + CH_NB * j];
    }
    for (int idx = t; idx < rows * kb; idx += 256) {
        const int r = idx % rows, l = idx / rows;
        pr[r][l] = P[(row0 + r) + (int64_t)l * lda];
    }
    __syncthreads();
    const int r = t % CH_PR, cg = t / CH_PR;  // 8 groups of 8 columns
    if (r < rows) {
#pragma unroll
        for (int jj = 0; jj < CH_NB / 8; ++jj) {
            const int j = cg * (CH_NB / 8) + jj;
            if (j < kb) {
                double acc = 0.0;
                for (int l = 0; l <= j; ++l) acc += pr[r][l] * w[j][l];
                P[(row0 + r) + (int64_t)j * lda] = acc;
            }
        }
    }
}

// Trailing update C -= P P^T of the lower triangle (C m x m, P m x kb, both column-major with lda),
// one 64 x 64 tile of C per workgroup (tiles on and below the diagonal, column by column), each lane
// 4 x 4 entries; batched over blockIdx.y like k_chol_diag (problem y at P + y dA, C + y dA).
__global__ __launch_bounds__(256) void k_chol_update(const double *__restrict__ P, double *__restrict__ C, int64_t lda,
                                                     int m, int kb, int64_t dA) {
    __shared__ double pi[CH_NB][CH_NB + 1];  // pi[r][k] = P(I0 + r, k)
    __shared__ double pj[CH_NB][CH_NB + 1];  // pj[r][k] = P(J0 + r, k)
    P += blockIdx.y * dA;
    C += blockIdx.y * dA;
    const int nT = (m + CH_NB - 1) / CH_NB;
    int tile = blockIdx.x, J = 0;
    while (tile >= nT - J) {
        tile -= nT - J;
        ++J;
    }
    const int I = J + tile;
    const int I0 = I * CH_NB, J0 = J * CH_NB;
    const int t = threadIdx.x;
    for (int idx = t; idx < CH_NB * kb; idx += 256) {
        const int r = idx % CH_NB, k = idx / CH_NB;
        pi[r][k] = I0 + r < m ? P[(I0 + r) + (int64_t)k * lda] : 0.0;
        pj[r][k] = J0 + r < m ? P[(J0 + r) + (int64_t)k * lda] : 0.0;
    }
    __syncthreads();
    const int ty = t % 16, tx = t / 16;  // rows I0 + 4 ty + i, columns J0 + 4 tx + j
    double acc[4][4] = {};
    for (int k = 0; k < kb; ++k) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = pi[4 * ty + i][k];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = pj[4 * tx + j][k];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = J0 + 4 * tx + j;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = I0 + 4 * ty + i;
            if (row < m && col < m && row >= col) C[row + (int64_t)col * lda] -= acc[i][j];
        }
    }
}

__global__ void k_identity(double *__restrict__ X, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    X[i] = (i % n == i / n) ? 1.0 : 0.0;
}

// Row-major copy of a column-major n x n block: Tr[r * n + c] = A[r + n * c].
__global__ void k_transpose_block(const double *__restrict__ A, double *__restrict__ Tr, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    const int64_t r = i / n, c = i - r * n;
    Tr[i] = A[r + n * c];
}

// Batched triangular products over all planes, out_p = Tri_p * in_p with Tri_p row-major n_p x n_p,
// lower (columns 0..r) or upper (columns r..n_p-1) triangular.  One wavefront per pair of rows
// (r, n_p - 1 - r) of a plane (pair Q, plane pairPlane[Q], r = Q - pairOff[p]; the middle row of an
// odd n_p alone): the pair holds n_p + 1 entries, so every wave streams the same amount.  The pair's
// two row segments are walked as one run of n_p + 1 entries, 8 x 64 per round with every load issued
// before the first use (clamped indices).
__global__ __launch_bounds__(256) void k_tri_pairs(const double *__restrict__ Tri, const int64_t *__restrict__ adOff,
                                                   const int64_t *__restrict__ off, const int32_t *__restrict__ pairPlane,
                                                   const int64_t *__restrict__ pairOff, int64_t npairs, int upper,
                                                   const double *__restrict__ in, double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t Q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (Q >= npairs) return;
    const int p = pairPlane[Q];
    const int64_t o = off[p], n = off[p + 1] - o;
    const int64_t r1 = Q - pairOff[p], r2 = n - 1 - r1;
    const int64_t len1 = upper ? n - r1 : r1 + 1;
    const int64_t len2 = r2 == r1 ? 0 : (upper ? n - r2 : r2 + 1);
    const int64_t c1 = upper ? r1 : 0, c2 = upper ? r2 : 0;
    const double *row1 = Tri + adOff[p] + r1 * n + c1, *x1 = in + o + c1;
    const double *row2 = Tri + adOff[p] + r2 * n + c2 - len1, *x2 = in + o + c2 - len1;  // indexed by j >= len1
    const int64_t tot = len1 + len2;
    double s1[8] = {0, 0, 0, 0, 0, 0, 0, 0}, s2[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t j0 = lane; j0 < tot; j0 += 8 * 64) {
        double a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t j = min(j0 + 64 * k, tot - 1);
            const bool first = j < len1;
            a[k] = first ? row1[j] : row2[j];
            b[k] = first ? x1[j] : x2[j];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t j = j0 + 64 * k;
            const double v = j < tot ? a[k] * b[k] : 0.0;
            if (j < len1) s1[k] += v;
            else s2[k] += v;
        }
    }
    double t1 = ((s1[0] + s1[1]) + (s1[2] + s1[3])) + ((s1[4] + s1[5]) + (s1[6] + s1[7]));
    double t2 = ((s2[0] + s2[1]) + (s2[2] + s2[3])) + ((s2[4] + s2[5]) + (s2[6] + s2[7]));
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
        t1 += __shfl_down(t1, o2, 64);
        t2 += __shfl_down(t2, o2, 64);
    }
    if (lane == 0) {
        out[o + r1] = t1;
        if (len2) out[o + r2] = t2;
    }
}

// One step of a block recurrence, one wavefront per row:  y[r] = c[r] - sum_j M[r * ld + j] v(j), with
// v = [v1 (m1 entries); v2 (m - m1 entries)] and M row-major n x m, rows zero-padded to
// ld = nchunks * NK * 64 (y may alias c).  Two independent problems share a launch (blocks
// [0, nbA) take problem A, the rest problem B).  The block stages v in LDS chunk by chunk; each lane
// issues its NK matrix loads of a chunk before the first use (no guards: the padding is zero), so a
// chunk costs one memory round trip.
struct ChainProb {
    const double *M;
    int64_t n, m, m1, ld;
    int nchunks;
    const double *c, *v1, *v2;
    double *y;
};

template <int NK>
__global__ __launch_bounds__(256) void k_chain_rows(const ChainProb A, const ChainProb B, unsigned nbA) {
    __shared__ double sv[NK * 64];
    const bool isA = blockIdx.x < nbA;
    const ChainProb &q = isA ? A : B;
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)(isA ? blockIdx.x : blockIdx.x - nbA) * 4 + (threadIdx.x >> 6);
    const double *row = q.M + (r < q.n ? r : 0) * q.ld + lane;
    double t = 0;
    for (int k0 = 0; k0 < q.nchunks; ++k0) {
        double a[NK];
#pragma unroll
        for (int k = 0; k < NK; ++k) a[k] = row[(int64_t)k0 * NK * 64 + 64 * k];
        __syncthreads();
        for (int j = threadIdx.x; j < NK * 64; j += 256) {
            const int64_t jj = (int64_t)k0 * NK * 64 + j;
            sv[j] = jj < q.m1 ? q.v1[jj] : (jj < q.m ? q.v2[jj - q.m1] : 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NK; ++k) t += a[k] * sv[lane + 64 * k];
    }
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
    if (lane == 0 && r < q.n) q.y[r] = q.c[r] - t;
}

// Padded row length (multiple of 64, or of 1024 beyond 1024) and the chunking used for it.
inline int64_t chain_ld(int64_t m) { return m <= 1024 ? ((m + 63) / 64) * 64 : ((m + 1023) / 1024) * 1024; }
inline int chain_nk(int64_t ld) { return (int)(ld <= 1024 ? ld / 64 : 16); }

ChainProb chain_prob(const double *M, int64_t n, int64_t m, int64_t m1, int64_t ld, const double *c, const double *v1,
                     const double *v2, double *y) {
    ChainProb q;
    q.M = M;
    q.n = n;
    q.m = m;
    q.m1 = m1;
    q.ld = ld;
    q.nchunks = (int)(ld <= 1024 ? 1 : ld / 1024);
    q.c = c;
    q.v1 = v1;
    q.v2 = v2;
    q.y = y;
    return q;
}

void launch_chain_nk(int nk, const ChainProb &A, const ChainProb &B, unsigned nbA, unsigned nbB, hipStream_t s) {
    switch (nk) {
#define TG_CHAIN_CASE(K) \
    case K: k_chain_rows<K><<<nbA + nbB, 256, 0, s>>>(A, B, nbA); break;
        TG_CHAIN_CASE(1) TG_CHAIN_CASE(2) TG_CHAIN_CASE(3) TG_CHAIN_CASE(4) TG_CHAIN_CASE(5) TG_CHAIN_CASE(6)
        TG_CHAIN_CASE(7) TG_CHAIN_CASE(8) TG_CHAIN_CASE(9) TG_CHAIN_CASE(10) TG_CHAIN_CASE(11) TG_CHAIN_CASE(12)
        TG_CHAIN_CASE(13) TG_CHAIN_CASE(14) TG_CHAIN_CASE(15) TG_CHAIN_CASE(16)
#undef TG_CHAIN_CASE
        default: fail(TG_ERR_ARGS, "tg_solver:coarse", "bad chain chunking");
    }
}

// One launch for both problems when they use the same chunking, otherwise one each (n == 0: absent).
void launch_chain(const ChainProb &A, const ChainProb &B, hipStream_t s) {
    const unsigned nbA = A.n ? grid_for(A.n, 4) : 0, nbB = B.n ? grid_for(B.n, 4) : 0;
    const int nkA = A.n ? chain_nk(A.ld) : 0, nkB = B.n ? chain_nk(B.ld) : 0;
    if (nbA && nbB && nkA == nkB) {
        launch_chain_nk(nkA, A, B, nbA, nbB, s);
    } else {
        if (nbA) launch_chain_nk(nkA, A, A, nbA, 0, s);
        if (nbB) launch_chain_nk(nkB, B, B, nbB, 0, s);
    }
}

template <typename T>
__global__ void k_pack_planes(const T *__restrict__ f, double *__restrict__ b, const uint8_t *__restrict__ mask,
                              const int32_t *__restrict__ lidx, const int64_t *__restrict__ off, const Grid g, int axis) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes() || !(mask[n] & NODE_ACTIVE)) return;
    const int y = (int)(n % g.NY);
    const int64_t t = n / g.NY;
    const int x = (int)(t % g.NX), z = (int)(t / g.NX);
    const int64_t base = off[coord(axis, y, x, z)] + 3 * (int64_t)lidx[n];
    b[base] = (double)f[3 * n];
    b[base + 1] = (double)f[3 * n + 1];
    b[base + 2] = (double)f[3 * n + 2];
}

template <typename T>
__global__ void k_unpack_planes(const double *__restrict__ b, T *__restrict__ f, const uint8_t *__restrict__ mask,
                                const int32_t *__restrict__ lidx, const int64_t *__restrict__ off, const Grid g, int axis) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes()) return;
    if (!(mask[n] & NODE_ACTIVE)) {
        f[3 * n] = f[3 * n + 1] = f[3 * n + 2] = (T)0;
        return;
    }
    const int y = (int)(n % g.NY);
    const int64_t t = n / g.NY;
    const int x = (int)(t % g.NX), z = (int)(t / g.NX);
    const int64_t base = off[coord(axis, y, x, z)] + 3 * (int64_t)lidx[n];
    f[3 * n] = (T)b[base];
    f[3 * n + 1] = (T)b[base + 1];
    f[3 * n + 2] = (T)b[base + 2];
}

void rb(rocblas_status s, const char *what) {
    if (s != rocblas_status_success)
        fail(TG_ERR_HIP, "tg_solver:rocblas", "%s failed with rocblas_status %d", what, (int)s);
}

}  // namespace

CoarseSolver::~CoarseSolver() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (h) rocblas_destroy_handle(h);

}

void CoarseSolver::setup(const HostLevel &L, const Grid &grid, hipStream_t st) {
    g = grid;
    const int64_t cnt[3] = {g.NY, g.NX, g.NZ};
    axis = 0;
    for (int a = 1; a < 3; ++a)
        if (cnt[a] > cnt[axis]) axis = a;
    P = (int)cnt[axis];
    // local numbering inside each plane: remaining axes in (y, x, z) order, y fastest
    std::vector<int32_t> lid(g.nodes(), -1);
    std::vector<int64_t> count(P, 0);
    std::vector<uint8_t> active(g.nodes(), 0);
    for (int64_t k = 0; k < L.Nn; ++k) active[L.nodFull[k] - 1] = 1;
    for (int64_t z = 0; z < g.NZ; ++z)
        for (int64_t x = 0; x < g.NX; ++x)
            for (int64_t y = 0; y < g.NY; ++y) {
                const int64_t n = y + g.NY * (x + g.NX * z);
                if (!active[n]) continue;
                const int p = (int)(axis == 0 ? y : (axis == 1 ? x : z));
                lid[n] = (int32_t)count[p]++;
            }
    nd.assign(P, 0);
    off.assign(P + 1, 0);
    for (int p = 0; p < P; ++p) {
        nd[p] = 3 * count[p];
        off[p + 1] = off[p] + nd[p];
    }
    ntot = off[P];
    m = P / 2;
    adOff.assign(P + 1, 0);
    asOff.assign(P + 1, 0);
    fOff.assign(P + 1, 0);
    bOff.assign(P + 1, 0);
    fLd.assign(P, 0);
    bLd.assign(P, 0);
    for (int p = 0; p < P; ++p) {
        const int64_t prev = p > 0 ? nd[p - 1] : 0, next = p + 1 < P ? nd[p + 1] : 0;
        adOff[p + 1] = adOff[p] + nd[p] * nd[p];
        asOff[p + 1] = asOff[p] + nd[p] * prev;
        // forward operator columns: the plane it is eliminated against (both neighbours for m)
        const int64_t fc = p < m ? prev : (p > m ? next : prev + next);
        // backward operator columns: the plane it is back-substituted from (none for m)
        const int64_t bc = p < m ? next : (p > m ? prev : 0);
        fLd[p] = fc ? chain_ld(fc) : 0;
        bLd[p] = bc ? chain_ld(bc) : 0;
        fOff[p + 1] = fOff[p] + nd[p] * fLd[p];
        bOff[p + 1] = bOff[p] + nd[p] * bLd[p];
    }
    // row pairs (r, n_p - 1 - r) of the batched triangular products
    std::vector<int64_t> pOff(P + 1, 0);
    for (int p = 0; p < P; ++p) pOff[p + 1] = pOff[p] + (nd[p] + 1) / 2;
    npairs = pOff[P];
    std::vector<int32_t> pp(std::max<int64_t>(npairs, 1), 0);
    for (int p = 0; p < P; ++p)
        for (int64_t Q = pOff[p]; Q < pOff[p + 1]; ++Q) pp[Q] = p;
    lidx.upload(lid, st);
    d_off.upload(off, st);
    d_adOff.upload(adOff, st);
    d_asOff.upload(asOff, st);
    pairPlane.upload(pp, st);
    d_pairOff.upload(pOff, st);
    Ad.alloc(adOff[P]);
    Linv.alloc(adOff[P]);
    LT.alloc(adOff[P]);
    As.alloc(std::max<int64_t>(asOff[P], 1));
    Fop.alloc(std::max<int64_t>(fOff[P], 1));
    Bop.alloc(std::max<int64_t>(bOff[P], 1));
    Fop.zero(st);  // row padding stays zero
    Bop.zero(st);
    bb.alloc(ntot);
    yy.alloc(ntot);
    xx.alloc(ntot);
    info.alloc(P);
    chW.alloc(2 * CH_NB * CH_NB);
    if (!h) rb(rocblas_create_handle(&h), "rocblas_create_handle");
    TG_HIP(hipStreamSynchronize(st), "tg_solver:coarse");
}

void CoarseSolver::factor(const double *Ks, int64_t ldk, const uint8_t *eleMask, const uint8_t *nodeMask,
                          hipStream_t st) {
    const char *id = "tg_solver:coarse";
    rb(rocblas_set_stream(h, st), "rocblas_set_stream");
    Ad.zero(st);
    if (asOff[P]) As.zero(st);
    info.zero(st);
    const int64_t nth = g.nodes() * 243;
    k_coarse_assemble<<<grid_for(nth, 256), 256, 0, st>>>(Ks, ldk, eleMask, nodeMask, g, axis, lidx, d_off, d_adOff,
                                                          d_asOff, Ad, As);
    TG_LAUNCH_CHECK(id);
    const double one = 1.0, mone = -1.0, zero = 0.0;
    // As_p (n_p x n_{p-1}, column-major) holds A_{p,p-1}.  Chain A turns As_{p+1} into
    // L_{p+1,p} = A_{p+1,p} L_pp^-T (right trsm); chain B turns As_p into X_p = L_pp^-1 A_{p,p-1}
    // = L_{p-1,p}^T (left trsm).
    // Plane factor (blocked Cholesky, kernels above) and its inverse inv(L) = L \ I by one triangular solve.  Every
    // call runs on the solver stream: with the two chains on two streams (two rocBLAS handles) the
    // factors came out different from run to run on ROCm 7.2 (wrong in the low to leading digits on
    // about one assembly in eight), however the streams were ordered; on one stream they are
    // bitwise reproducible.
    // Factor plane p, or planes p and q (same size) together: the LDS kernels take both in one launch
    // (blockIdx.y); the rocBLAS calls stay per plane.
    auto factor_planes = [&](int p, int q) {
        const rocblas_int np = (rocblas_int)nd[p];
        const int nb = q >= 0 ? 2 : 1;
        const int64_t dA = q >= 0 ? adOff[q] - adOff[p] : 0;
        const int dinfo = q >= 0 ? q - p : 0;
        for (int k0 = 0; k0 < np; k0 += CH_NB) {
            const int kb = std::min<int>(CH_NB, np - k0), mr = np - k0 - kb;
            double *D = Ad.p + adOff[p] + k0 + (int64_t)k0 * np;
            k_chol_diag<<<dim3(1, nb), 256, 0, st>>>(D, np, kb, chW.p, info.p + p, k0, dA, dinfo);
            if (mr > 0)
                k_chol_panel<<<dim3((mr + CH_PR - 1) / CH_PR, nb), 256, 0, st>>>(D + kb, np, mr, kb, chW.p, dA);
            TG_LAUNCH_CHECK(id);
            if (mr > 0) {
                const int nT = (mr + CH_NB - 1) / CH_NB;
                k_chol_update<<<dim3(nT * (nT + 1) / 2, nb), 256, 0, st>>>(D + kb, D + kb + (int64_t)kb * np, np, mr,
                                                                        kb, dA);
                TG_LAUNCH_CHECK(id);
            }
        }
        for (int y = 0; y < nb; ++y) {
            const int pp = y ? q : p;
            k_identity<<<grid_for((int64_t)np * np, 256), 256, 0, st>>>(Linv.p + adOff[pp], np);
            TG_LAUNCH_CHECK(id);
            rb(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                             rocblas_diagonal_non_unit, np, np, &one, Ad.p + adOff[pp], np, Linv.p + adOff[pp], np),
               "rocblas_dtrsm inverse");
        }
    };
    // chains A (p = 0 .. m-1) and B (q = P-1 .. m+1) advance together, one plane each per step
    const int lenA = m, lenB = P - 1 - m;
    for (int k = 0; k < std::max(lenA, lenB); ++k) {
        const int p = k < lenA && nd[k] > 0 ? k : -1;
        const int q = k < lenB && nd[P - 1 - k] > 0 ? P - 1 - k : -1;
        if (p > 0 && nd[p - 1] > 0)  // S_p -= L_{p,p-1} L_{p,p-1}^T
            rb(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, (rocblas_int)nd[p], (rocblas_int)nd[p - 1],
                             &mone, As.p + asOff[p], (rocblas_int)nd[p], &one, Ad.p + adOff[p], (rocblas_int)nd[p]),
               "rocblas_dsyrk");
        if (q >= 0 && q + 1 < P && nd[q + 1] > 0)  // S_q -= X_{q+1}^T X_{q+1}
            rb(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, (rocblas_int)nd[q],
                             (rocblas_int)nd[q + 1], &mone, As.p + asOff[q + 1], (rocblas_int)nd[q + 1], &one,
                             Ad.p + adOff[q], (rocblas_int)nd[q]),
               "rocblas_dsyrk");
        if (p >= 0 && q >= 0 && nd[p] == nd[q]) {
            factor_planes(p, q);
        } else {
            if (p >= 0) factor_planes(p, -1);
            if (q >= 0) factor_planes(q, -1);
        }
        if (p >= 0 && nd[p + 1] > 0)  // L_{p+1,p} = A_{p+1,p} L_pp^-T
            rb(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                             rocblas_diagonal_non_unit, (rocblas_int)nd[p + 1], (rocblas_int)nd[p], &one,
                             Ad.p + adOff[p], (rocblas_int)nd[p], As.p + asOff[p + 1], (rocblas_int)nd[p + 1]),
               "rocblas_dtrsm A");
        if (q >= 0 && nd[q - 1] > 0)  // X_q = L_qq^-1 A_{q,q-1}
            rb(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                             rocblas_diagonal_non_unit, (rocblas_int)nd[q], (rocblas_int)nd[q - 1], &one,
                             Ad.p + adOff[q], (rocblas_int)nd[q], As.p + asOff[q], (rocblas_int)nd[q]),
               "rocblas_dtrsm B");
    }
    if (nd[m] > 0) {  // separator
        const rocblas_int nm = (rocblas_int)nd[m];
        if (m > 0 && nd[m - 1] > 0)
            rb(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, nm, (rocblas_int)nd[m - 1], &mone,
                             As.p + asOff[m], nm, &one, Ad.p + adOff[m], nm), "rocblas_dsyrk");
        if (m + 1 < P && nd[m + 1] > 0)
            rb(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, nm, (rocblas_int)nd[m + 1], &mone,
                             As.p + asOff[m + 1], (rocblas_int)nd[m + 1], &one, Ad.p + adOff[m], nm), "rocblas_dsyrk");
        factor_planes(m, -1);
    }
    // Solve operators, row-major with zero-padded rows (column-major results of transposed products):
    //   LT_p = inv(L_pp) (row-major copy); inv(L_pp)^T is the column-major Linv itself.
    //   forward  A: inv(L_pp) L_{p,p-1}       = (As_p^T Linv_p^T)^T      (n_p x n_{p-1})
    //            B: inv(L_pp) L_{p,p+1}       = (X_{p+1} Linv_p^T)^T     (n_p x n_{p+1})
    //            m: [inv(L_mm) L_{m,m-1} | inv(L_mm) L_{m,m+1}]
    //   backward A: inv(L_pp)^T L_{p+1,p}^T   = (As_{p+1} Linv_p)^T      (n_p x n_{p+1})
    //            B: inv(L_pp)^T L_{p-1,p}^T   = (X_p^T Linv_p)^T         (n_p x n_{p-1})
    for (int p = 0; p < P; ++p) {
        const rocblas_int np = (rocblas_int)nd[p];
        if (np == 0) continue;
        const int64_t prev = p > 0 ? nd[p - 1] : 0, next = p + 1 < P ? nd[p + 1] : 0;
        const double *Lp = Linv.p + adOff[p];
        k_transpose_block<<<grid_for((int64_t)np * np, 256), 256, 0, st>>>(Lp, LT.p + adOff[p], np);
        TG_LAUNCH_CHECK(id);
        double *F = Fop.p + fOff[p], *B = Bop.p + bOff[p];
        const rocblas_int fl = (rocblas_int)fLd[p], bl = (rocblas_int)bLd[p];
        if (p <= m && prev > 0)  // A-side coupling (chain A, or the separator's first block)
            rb(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_transpose, (rocblas_int)prev, np, np,
                             &one, As.p + asOff[p], np, Lp, np, &zero, F, fl), "rocblas_dgemm F A");
        if (p >= m && next > 0)  // B-side coupling (chain B, or the separator's second block)
            rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)next, np, np,
                             &one, As.p + asOff[p + 1], (rocblas_int)next, Lp, np, &zero, F + (p == m ? prev : 0),
                             fl), "rocblas_dgemm F B");
        if (p < m && next > 0)
            rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)next, np, np, &one,
                             As.p + asOff[p + 1], (rocblas_int)next, Lp, np, &zero, B, bl), "rocblas_dgemm B A");
        if (p > m && prev > 0)
            rb(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)prev, np, np, &one,
                             As.p + asOff[p], np, Lp, np, &zero, B, bl), "rocblas_dgemm B B");
    }
    std::vector<rocblas_int> hinfo(P);
    TG_HIP(hipMemcpyAsync(hinfo.data(), info.p, sizeof(rocblas_int) * P, hipMemcpyDeviceToHost, st), id);
    TG_HIP(hipStreamSynchronize(st), id);
    for (int p = 0; p < P; ++p)
        if (hinfo[p] != 0)
            fail(TG_ERR_NUMERIC, "tg_solver:coarse",
                 "coarsest-level factorisation failed on plane block %d (info %d): matrix not positive definite",
                 p, (int)hinfo[p]);
    factored = true;
}

// The solve's fixed sequence on the internal plane-ordered vectors (bb -> xx):
//   yy = blockdiag(LT) bb; forward chains A (p = 1 .. m-1) and B (p = P-2 .. m+1) step by step together;
//   the separator; xx = blockdiag(Linv^T) yy; backward chains A (p = m-1 .. 0) and B (p = m+1 .. P-1).
void CoarseSolver::enqueue_core(hipStream_t s) const {
    const char *id = "tg_solver:coarse";
    const unsigned gr = grid_for(npairs, 4);
    const ChainProb none = chain_prob(nullptr, 0, 0, 0, 0, nullptr, nullptr, nullptr, nullptr);
    auto fwd = [&](int p, int q) -> ChainProb {  // y_p -= Fop_p y_q
        if (p < 0 || p >= P || nd[p] == 0 || q < 0 || q >= P || nd[q] == 0) return none;
        return chain_prob(Fop.p + fOff[p], nd[p], nd[q], nd[q], fLd[p], yy.p + off[p], yy.p + off[q], nullptr,
                          yy.p + off[p]);
    };
    auto bwd = [&](int p, int q) -> ChainProb {  // x_p -= Bop_p x_q
        if (p < 0 || p >= P || nd[p] == 0 || q < 0 || q >= P || nd[q] == 0) return none;
        return chain_prob(Bop.p + bOff[p], nd[p], nd[q], nd[q], bLd[p], xx.p + off[p], xx.p + off[q], nullptr,
                          xx.p + off[p]);
    };
    k_tri_pairs<<<gr, 256, 0, s>>>(LT, d_adOff, d_off, pairPlane, d_pairOff, npairs, 0, bb, yy);
    const int lenA = m, lenB = P - 1 - m;  // planes per chain
    for (int k = 1; k < std::max(lenA, lenB); ++k)
        launch_chain(k < lenA ? fwd(k, k - 1) : none, k < lenB ? fwd(P - 1 - k, P - k) : none, s);
    if (nd[m] > 0) {  // separator: y_m -= Fop_m [y_{m-1}; y_{m+1}]
        const int64_t na = (m > 0) ? nd[m - 1] : 0, nb = (m + 1 < P) ? nd[m + 1] : 0;
        if (na + nb > 0)
            launch_chain(chain_prob(Fop.p + fOff[m], nd[m], na + nb, na, fLd[m], yy.p + off[m],
                                    na ? yy.p + off[m - 1] : yy.p, nb ? yy.p + off[m + 1] : yy.p, yy.p + off[m]),
                         none, s);
    }
    k_tri_pairs<<<gr, 256, 0, s>>>(Linv, d_adOff, d_off, pairPlane, d_pairOff, npairs, 1, yy, xx);
    for (int k = 1; k <= std::max(lenA, lenB); ++k)
        launch_chain(k <= lenA ? bwd(m - k, m - k + 1) : none, k <= lenB ? bwd(m + k, m + k - 1) : none, s);
    TG_LAUNCH_CHECK(id);
}

template <typename T>
void CoarseSolver::solve(const T *r, T *x, const uint8_t *nodeMask, hipStream_t st) {
    const char *id = "tg_solver:coarse";
    if (!factored) fail(TG_ERR_STATE, id, "coarsest level not factorised (call assemble first)");
    k_pack_planes<T><<<grid_for(g.nodes(), 256), 256, 0, st>>>(r, bb, nodeMask, lidx, d_off, g, axis);
    TG_LAUNCH_CHECK(id);
    if (!use_graph) {
        enqueue_core(st);
    } else {
        if (!exec) {
        // the sequence reads only buffers owned here, so one capture serves every later solve
        hipStream_t cap;
        TG_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), id);
        TG_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed), id);
        enqueue_core(cap);
        hipGraph_t graph;
        TG_HIP(hipStreamEndCapture(cap, &graph), id);
        TG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), id);
        TG_HIP(hipGraphDestroy(graph), id);
        TG_HIP(hipStreamDestroy(cap), id);
    }
    TG_HIP(hipGraphLaunch(exec, st), id);
    }
    k_unpack_planes<T><<<grid_for(g.nodes(), 256), 256, 0, st>>>(xx, x, nodeMask, lidx, d_off, g, axis);
    TG_LAUNCH_CHECK(id);
}

template void CoarseSolver::solve<double>(const double *, double *, const uint8_t *, hipStream_t);
template void CoarseSolver::solve<float>(const float *, float *, const uint8_t *, hipStream_t);

}  // namespace tg
