// Coarsest-level direct solver: twisted block-Schur-complement chain on the device (see coarse.h).
#include "coarse.h"

#include <cstring>

namespace tg {
namespace {

__device__ inline int coord(int axis, int y, int x, int z) { return axis == 0 ? y : (axis == 1 ? x : z); }

// In-plane neighbour index (0..8) of a node offset: the two offsets along the axes other than `axis`,
// in (y, x, z) order.
__device__ inline int inplane9(int axis, int dy, int dx, int dz) {
    const int a = axis == 0 ? dx : dy, b = axis == 2 ? dx : dz;
    return (a + 1) + 3 * (b + 1);
}

constexpr int CB = 64;  // pivot block of the Gauss-Jordan sweep (plane blocks are padded to multiples)
constexpr int NCPL = 27;  // couplings per DOF row to one neighbouring plane (3 x 3 nodes x 3 DOFs)

// Plane blocks before assembly: zero, with ones on the diagonal of the padding rows (n_p <= i < npad_p),
// so the padded block is [A_pp 0; 0 I] and its inverse [S^-1 0; 0 I].
__global__ void k_plane_init(double *__restrict__ S, const int64_t *__restrict__ sOff, const int64_t *__restrict__ npad,
                             const int64_t *__restrict__ nd) {
    const int p = blockIdx.y;
    const int64_t np = npad[p], t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np * np) return;
    const int64_t i = t / np, j = t - i * np;
    S[sOff[p] + t] = (i == j && i >= nd[p]) ? 1.0 : 0.0;
}

// One lane per (node J, neighbour slot, DOF pair): K_c(J c, M d) = sum over the solid elements shared by
// J and M of Ks_e(3a_J + c, 3b_M + d) (the entries sparse(iK, jK, sK) accumulates, TOP_GPU.m:1249-1254).
// Same-plane entries go to the plane's dense block (row-major, padded); entries coupling J to the
// previous / next plane go to the compact coupling rows (slot = in-plane neighbour x 3 + d, with the
// neighbour's local DOF as column; absent neighbours get value 0 and column 0).  Only one triangle of
// K_c is formed, as chol(K_c, 'lower') reads only one: entries above the diagonal in plane order are
// summed from the transposed element entries Ks_e(3b + d, 3a + c), over the same elements in the same
// (ascending) order as their mirror image, so the plane blocks are exactly symmetric and the coupling
// to the next plane is exactly the transpose of the next plane's coupling to this one.
__global__ void k_coarse_assemble(const double *__restrict__ Ks, int64_t ldk, const uint8_t *__restrict__ eleMask,
                                  const uint8_t *__restrict__ nodeMask, const Grid g, int axis,
                                  const int32_t *__restrict__ lidx, const int64_t *__restrict__ off,
                                  const int64_t *__restrict__ sOff, const int64_t *__restrict__ npad,
                                  double *__restrict__ S, double *__restrict__ dnV, int32_t *__restrict__ dnC,
                                  double *__restrict__ upV, int32_t *__restrict__ upC) {
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
    const int pJ = coord(axis, jy, jx, jz), dp = coord(axis, dy, dx, dz);
    const int64_t row = off[pJ] + 3 * (int64_t)lidx[J] + c;
    const int slot = 3 * inplane9(axis, dy, dx, dz) + d;
    const int my = jy + dy, mx = jx + dx, mz = jz + dz;
    const bool inside = (unsigned)my < (unsigned)g.NY && (unsigned)mx < (unsigned)g.NX && (unsigned)mz < (unsigned)g.NZ;
    const int64_t M = inside ? my + (int64_t)g.NY * (mx + (int64_t)g.NX * mz) : 0;
    if (!inside || !(nodeMask[M] & NODE_ACTIVE)) {
        if (dp == -1) { dnV[row * NCPL + slot] = 0.0; dnC[row * NCPL + slot] = 0; }
        if (dp == 1) { upV[row * NCPL + slot] = 0.0; upC[row * NCPL + slot] = 0; }
        return;
    }
    const int64_t col = 3 * (int64_t)lidx[M] + d;
    const bool tr = dp == 1 || (dp == 0 && col > row - off[pJ]);
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
                val += Ks[(int64_t)(tr ? (3 * b + d) + 24 * (3 * a + c) : (3 * a + c) + 24 * (3 * b + d)) * ldk + e];
            }
    if (dp == 0) {
        S[sOff[pJ] + (row - off[pJ]) * npad[pJ] + col] = val;
    } else if (dp == -1) {
        dnV[row * NCPL + slot] = val;
        dnC[row * NCPL + slot] = (int32_t)col;
    } else {
        upV[row * NCPL + slot] = val;
        upC[row * NCPL + slot] = (int32_t)col;
    }
}

// ---- Schur complement update:  S(i, j) -= sum_k V(i, k) N(C(i, k), j)  (A_{p,p+-1} N, up to two terms) ----
struct SchurTerm {
    const double *V;
    const int32_t *C;
    const double *N;
    int64_t ldN;
};
struct SchurProb {
    double *S;
    int64_t np, n;
    int nt;
    SchurTerm t[2];
};

__global__ __launch_bounds__(256) void k_schur(const SchurProb A, const SchurProb B) {
    const SchurProb &q = blockIdx.z ? B : A;
    const int64_t i = blockIdx.y, j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n || j >= q.n) return;
    double acc = 0.0;
    for (int s = 0; s < q.nt; ++s) {
        const double *V = q.t[s].V + i * NCPL;
        const int32_t *C = q.t[s].C + i * NCPL;
        const double *N = q.t[s].N + j;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int k = 0; k < NCPL; k += 3) {
            a0 += V[k] * N[(int64_t)C[k] * q.t[s].ldN];
            a1 += V[k + 1] * N[(int64_t)C[k + 1] * q.t[s].ldN];
            a2 += V[k + 2] * N[(int64_t)C[k + 2] * q.t[s].ldN];
        }
        acc += (a0 + a1) + a2;
    }
    q.S[i * q.np + j] -= acc;
}

// ---- blocked Gauss-Jordan sweep: S <- S^-1 in place, pivot blocks K = 0 .. nb-1 -------------------------
// Per pivot block K (with D = S_KK^-1, computed from the current S):
//   S_IJ -= S_IK D S_KJ (I, J != K),  S_IK <- S_IK D,  S_KJ <- -D S_KJ,  S_KK <- D,
// which after the last block leaves the inverse (S is SPD, so every pivot block is the SPD Schur
// complement of the blocks already swept and no pivoting is needed).  k_sweep_panel forms T_I = S_IK D
// and copies the old pivot row Rw_I = S_KI; k_sweep_update then writes every block from T, Rw and D, so
// no block is read after another workgroup has overwritten it.
struct SweepProb {
    double *S;      // npad x npad, row-major
    int64_t np;     // npad
    int nb;         // npad / CB
    double *T;      // npad x CB   (T_I = S_IK D)
    double *Rw;     // CB x npad   (old pivot row block)
    double *D;      // CB x CB
    int *info;      // first non-positive pivot (1-based, 0 = none)
};

// 1 / x for a positive normal x: the hardware reciprocal estimate refined by two Newton-Raphson steps
// (p <- p + p (1 - x p)), each of which squares the relative error.
__device__ __forceinline__ double recip_nr(double x) {
    double p = __builtin_amdgcn_rcp(x);
    double e = fma(-x, p, 1.0);
    p = fma(p, e, p);
    e = fma(-x, p, 1.0);
    return fma(p, e, p);
}

// 64 x 64 SPD inverse by Gauss-Jordan in registers: thread t holds the 4 x 4 block at rows 4 (t / 16),
// columns 4 (t % 16); each pivot's row and column pass through LDS (double-buffered, one barrier per
// pivot).  Pivots go four at a time so that the pivot's place in a thread's block is static.  Every
// entry takes the rank-1 update, and the threads holding row k or column k then overwrite those
// entries (row k: -a_kj / a_kk, column k: a_ik / a_kk, pivot: 1 / a_kk).  Returns 1 + the first
// non-positive pivot index (0 if none; such a pivot is replaced by 1).
__device__ int gj_inverse64(double (&r)[4][4], double (*rowb)[CB], double (*colb)[CB]) {
    const int t = threadIdx.x, ti = t >> 4, tj = t & 15, bi = 4 * ti, bj = 4 * tj;
    int bad = 0;
    for (int k4 = 0; k4 < CB / 4; ++k4) {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = 4 * k4 + kk, b = kk & 1;
            if (ti == k4) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) rowb[b][bj + jj] = r[kk][jj];
            }
            if (tj == k4) {
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) colb[b][bi + ii] = r[ii][kk];
            }
            __syncthreads();
            double piv = colb[b][k];
            if (!(piv > 0.0)) {
                if (!bad) bad = k + 1;
                piv = 1.0;
            }
            const double p = recip_nr(piv);
            double ap[4], kb[4];
#pragma unroll
            for (int ii = 0; ii < 4; ++ii) ap[ii] = colb[b][bi + ii] * p;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) kb[jj] = rowb[b][bj + jj];
#pragma unroll
            for (int ii = 0; ii < 4; ++ii)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) r[ii][jj] = fma(-ap[ii], kb[jj], r[ii][jj]);
            if (tj == k4) {
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) r[ii][kk] = ap[ii];
            }
            if (ti == k4) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) r[kk][jj] = -p * kb[jj];
                if (tj == k4) r[kk][kk] = p;
            }
        }
    }
    return bad;
}

__global__ __launch_bounds__(256) void k_sweep_panel(const SweepProb A, const SweepProb B, int K) {
    __shared__ double dl[CB][CB + 1];   // D
    __shared__ double sl[CB][CB + 1];   // S_IK
    __shared__ double rowb[2][CB], colb[2][CB];
    const SweepProb &q = blockIdx.y ? B : A;
    const int I = blockIdx.x;
    if (K >= q.nb || I >= q.nb) return;
    const int t = threadIdx.x, bi = 4 * (t / 16), bj = 4 * (t % 16);
    const int64_t np = q.np, k0 = (int64_t)K * CB, i0 = (int64_t)I * CB;
    double r[4][4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) r[ii][jj] = q.S[(k0 + bi + ii) * np + k0 + bj + jj];
    const int bad = gj_inverse64(r, rowb, colb);
    if (I == K) {
#pragma unroll
        for (int ii = 0; ii < 4; ++ii)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) q.D[(bi + ii) * CB + bj + jj] = r[ii][jj];
        if (t == 0 && bad && *q.info == 0) *q.info = (int)k0 + bad;
        return;
    }
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            dl[bi + ii][bj + jj] = r[ii][jj];
            sl[bi + ii][bj + jj] = q.S[(i0 + bi + ii) * np + k0 + bj + jj];
            q.Rw[(int64_t)(bi + ii) * np + i0 + bj + jj] = q.S[(k0 + bi + ii) * np + i0 + bj + jj];
        }
    __syncthreads();
    double acc[4][4] = {};
    for (int k = 0; k < CB; ++k) {
        double a[4], b[4];
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) a[ii] = sl[bi + ii][k];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) b[jj] = dl[k][bj + jj];
#pragma unroll
        for (int ii = 0; ii < 4; ++ii)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) acc[ii][jj] += a[ii] * b[jj];
    }
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) q.T[(i0 + bi + ii) * CB + bj + jj] = acc[ii][jj];
}

// Every 64 x 64 block (I, J) of the plane block from T, Rw and D.  The products run on the matrix cores
// (MFMA f64 16x16x4: lane l holds A(l % 16, l / 16), B(l / 16, l % 16), result register r at
// (4 r + l / 16, l % 16)); wave w computes the 32 x 32 quadrant (w / 2, w % 2) of the block.
typedef double tg_cdbl4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_sweep_update(const SweepProb A, const SweepProb B, int K) {
    __shared__ double al[CB][CB + 1];
    __shared__ double bl[CB][CB + 1];
    const SweepProb &q = blockIdx.y ? B : A;
    if (K >= q.nb) return;
    const int I = blockIdx.x / q.nb, J = blockIdx.x % q.nb;
    if (I >= q.nb) return;
    const int t = threadIdx.x;
    const int64_t np = q.np, i0 = (int64_t)I * CB, j0 = (int64_t)J * CB;
    if (J == K) {  // S_IK <- T_I, S_KK <- D
        for (int idx = t; idx < CB * CB; idx += 256) {
            const int ii = idx / CB, jj = idx % CB;
            q.S[(i0 + ii) * np + j0 + jj] = I == K ? q.D[ii * CB + jj] : q.T[(i0 + ii) * CB + jj];
        }
        return;
    }
    // S_IJ -= T_I Rw_J  (I != K),  S_KJ = -D Rw_J
    for (int idx = t; idx < CB * CB; idx += 256) {
        const int ii = idx / CB, jj = idx % CB;
        al[ii][jj] = I == K ? q.D[ii * CB + jj] : q.T[(i0 + ii) * CB + jj];
        bl[ii][jj] = q.Rw[(int64_t)ii * np + j0 + jj];
    }
    __syncthreads();
    const int w = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
    const int R0 = 32 * (w >> 1), C0 = 32 * (w & 1);
    tg_cdbl4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = tg_cdbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int kk = 0; kk < CB / 4; ++kk) {
        const double x0 = al[R0 + li][4 * kk + lk], x1 = al[R0 + 16 + li][4 * kk + lk];
        const double y0 = bl[4 * kk + lk][C0 + li], y1 = bl[4 * kk + lk][C0 + 16 + li];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, y0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, y1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, y0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, y1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double *dst = q.S + (i0 + R0 + 16 * a + 4 * r + lk) * np + j0 + C0 + 16 * b + li;
                *dst = I == K ? -acc[a][b][r] : *dst - acc[a][b][r];
            }
}

// ---- solve operators:  Out(i, col0 + 3 jn + c) = sum_k S(i, C(3 jn, k)) V(3 jn + c, k) ----------------
// (S^-1 times a coupling block whose column 3 jn + c holds the 27 couplings of the neighbouring plane's
// DOF 3 jn + c; the three DOFs of one node share their column pattern.)
struct DSProb {
    const double *S;
    int64_t np, nrows, ncn;  // S row stride, rows, column nodes
    const double *V;
    const int32_t *C;
    double *Out;
    int64_t ld, col0;
};

// One thread per output column 3 jn + c and run of DS_ROWS rows: the column's 27 couplings and their
// columns stay in registers over the run, the rows' gathered entries come from the cache (the lanes of
// a wavefront share the row), and lanes write consecutive columns.  (A thread per row and column node
// re-read the couplings for every row: 167 us per launch at 128 x 256 x 128.)
constexpr int DS_ROWS = 16;
__global__ __launch_bounds__(256) void k_dense_sparse(const DSProb P0, const DSProb P1, const DSProb P2,
                                                      const DSProb P3) {
    const DSProb &q = blockIdx.y == 0 ? P0 : (blockIdx.y == 1 ? P1 : (blockIdx.y == 2 ? P2 : P3));
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncol = 3 * q.ncn, nruns = (q.nrows + DS_ROWS - 1) / DS_ROWS;
    if (t >= ncol * nruns) return;
    const int64_t jc = t % ncol, i0 = (t / ncol) * DS_ROWS;
    const int64_t jn = jc / 3;
    const int32_t *C = q.C + 3 * jn * NCPL;  // the three DOFs of node jn share their column pattern
    const double *V = q.V + jc * NCPL;
    int32_t ck[NCPL];
    double vk[NCPL];
#pragma unroll
    for (int k = 0; k < NCPL; ++k) {
        ck[k] = C[k];
        vk[k] = V[k];
    }
    const int64_t i1 = i0 + DS_ROWS < q.nrows ? i0 + DS_ROWS : q.nrows;
    for (int64_t i = i0; i < i1; ++i) {
        const double *srow = q.S + i * q.np;
        double o = 0.0;
#pragma unroll
        for (int k = 0; k < NCPL; ++k) o += srow[ck[k]] * vk[k];
        q.Out[i * q.ld + q.col0 + jc] = o;
    }
}

// c = blockdiag(S_p^-1) b over all planes: one wavefront per DOF row.
__global__ __launch_bounds__(256) void k_plane_gemv(const double *__restrict__ S, const int64_t *__restrict__ sOff,
                                                    const int64_t *__restrict__ npad, const int64_t *__restrict__ off,
                                                    const int32_t *__restrict__ rowPlane, int64_t ntot,
                                                    const double *__restrict__ b, double *__restrict__ c) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= ntot) return;
    const int p = rowPlane[r];
    const int64_t o = off[p], n = off[p + 1] - o;
    const double *row = S + sOff[p] + (r - o) * npad[p];
    const double *x = b + o;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t j = lane;
    for (; j + 192 < n; j += 256) {
        const double a0 = row[j], a1 = row[j + 64], a2 = row[j + 128], a3 = row[j + 192];
        const double x0 = x[j], x1 = x[j + 64], x2 = x[j + 128], x3 = x[j + 192];
        s[0] += a0 * x0;
        s[1] += a1 * x1;
        s[2] += a2 * x2;
        s[3] += a3 * x3;
    }
    for (; j < n; j += 64) s[0] += row[j] * x[j];
    double tsum = (s[0] + s[1]) + (s[2] + s[3]);
    for (int o2 = 32; o2 > 0; o2 >>= 1) tsum += __shfl_down(tsum, o2, 64);
    if (lane == 0) c[r] = tsum;
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

// ---- the solve as one persistent launch -------------------------------------------------------------------
// Launched step by step, the solve's 2 P - 1 dependent steps each start from cold: a step's matrix rows
// are fetched only once its launch runs.  Here every step is a stage of one launch whose workgroups
// stay resident and pass the recurrence vectors to each other directly.  A stage holds one or two row
// problems  y = c + M1 v1 - M2 [v2a; v2b],  one wavefront per row (the two chains of a step split the
// workgroups in proportion to their rows).  The forward stages fold the block-diagonal product into
// the recurrence, z_p = S_p^-1 b_p - F_p z_{p-1}, and write zz; the separator and the backward stages,
// x_p = z_p - B_p x_{p+1}, write xx.
// Dataflow instead of barriers: before the launch (by the kernel that packs the right-hand side: the
// V-cycle's restriction gather, else k_pack_planes) zz and xx are filled with a sentinel bit pattern (a signalling NaN that arithmetic never produces), every
// entry is written once per launch, and a
// workgroup that needs an earlier stage's vector polls exactly the entries it stages until they are no
// longer the sentinel.  Stores and these loads are device-scope atomic accesses, coherent across the
// XCDs' separate L2 caches.  (A grid barrier between stages measured 2.6 us with per-group counters and
// 8 us with one counter, on top of the exchange itself.)  Each wavefront runs ahead as far as its inputs
// allow: the loads of a stage that do not depend on earlier stages (its M1 row with the lane's entries
// of the right-hand-side plane, the first chunk of its M2 row) are issued before it polls the previous
// stage's input, so a stage costs one memory round trip.  A poll that waits
// longer than a bounded number of tries sets an error flag and gives up, so the launch always ends;
// the host checks the flag.  G is at most the number of workgroups the device holds at once, so every
// producer is resident while consumers wait.
struct PRow {
    const double *M1, *M2;  // row-major, rows zero-padded to ld1 / ld2 (nullptr: no such term)
    int64_t ld1, ld2;
    const double *v1;         // n1 entries (a plane of the right-hand side, written before the launch)
    const double *v2a, *v2b;  // n2a, n2b entries (written by earlier stages of the launch)
    const double *c;          // nullptr: none; else y[r] = c[r] + ... (written by an earlier stage)
    double *y;
    int n, n1, n2a, n2b;
    int yoff;  // >= 0: the result is also the solution, full-grid DOF gmap[yoff + r] (else -1)
};

// The solution on the full grid (one of x64, x32), through the plane-order map; inactive DOFs zeroed.
struct PersOut {
    double *x64;
    float *x32;
    const int32_t *gmap, *gzero;
    int nzero;
};
struct PStage {
    PRow q[2];
    int nq;  // 1 or 2
    int gA;  // workgroups of q[0] when nq == 2
};

constexpr unsigned long long PERS_EMPTY = COARSE_EMPTY;

// The table's pointers are generic; these name them as global memory, so that the loads are global
// (not flat) accesses.
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) unsigned long long gull;
__device__ __forceinline__ gcdouble *gmem(const double *p) { return (gcdouble *)p; }
__device__ __forceinline__ unsigned long long coh_bits(const double *p) {
    return __hip_atomic_load((gull *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void coh_store(double *p, double v) {
    __hip_atomic_store((gull *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
// The value at p once an earlier stage has written it (bounded wait; the error flag records a give-up).
// A poll waits for its own load, so polls are about a memory round trip (~1 us) apart.  Every 64 polls
// the error flag is read too: once any wait of the launch has given up, every other one ends at once.
__device__ __forceinline__ double await_value(const double *p, unsigned long long bits, int *err, unsigned max_polls) {
    unsigned polls = 0;
    while (bits == PERS_EMPTY) {
        __builtin_amdgcn_s_sleep(1);
        bits = coh_bits(p);
        ++polls;
        if (polls > max_polls ||
            ((polls & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bits = 0;
        }
    }
    return __longlong_as_double((long long)bits);
}

// Stages n doubles (zero past the given lengths) into LDS: all of a thread's loads are issued before
// any is examined, so a ready vector costs one memory round trip per 2048 entries.  AWAIT: the entries
// are written by earlier stages of this launch (polled until written).
template <bool AWAIT>
__device__ __forceinline__ void stage_vec(double *dst, int n, const double *a, int na, const double *b, int nb,
                                          int *err, unsigned max_polls) {
    for (int j0 = 0; j0 < n; j0 += 8 * 256) {
        unsigned long long v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int j = j0 + 256 * k + (int)threadIdx.x;
            unsigned long long x = 0;
            if (j < na) x = AWAIT ? coh_bits(a + j) : (unsigned long long)__double_as_longlong(gmem(a)[j]);
            else if (j < na + nb) x = AWAIT ? coh_bits(b + (j - na)) : (unsigned long long)__double_as_longlong(gmem(b)[j - na]);
            v[k] = x;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int j = j0 + 256 * k + (int)threadIdx.x;
            if (j < n) {
                double x = __longlong_as_double((long long)v[k]);
                if (AWAIT && v[k] == PERS_EMPTY) x = await_value(j < na ? a + j : b + (j - na), v[k], err, max_polls);
                dst[j] = x;
            }
        }
    }
}

template <int NK, bool NT>
__global__ __launch_bounds__(256, 2) void k_coarse_persistent(const PStage *__restrict__ stages, int nstages, int L2,
                                                              int *err, unsigned max_polls, const PersOut o) {
    extern __shared__ double sv[];  // two buffers of L2: [v2a; v2b] of stage s in buffer s & 1, zero past the ends
    constexpr int CH = 64 * NK;
    const int g = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = g * 256 + (int)threadIdx.x; i < o.nzero; i += (int)gridDim.x * 256) {  // inactive DOFs
        if (o.x64) o.x64[o.gzero[i]] = 0.0;
        else o.x32[o.gzero[i]] = 0.0f;
    }
    auto chunk_len = [&](int64_t ld) { return (int)(((ld + CH - 1) / CH) * CH); };
    // NT: the factor rows are read once per solve; streaming (non-temporal) loads keep them from
    // displacing the finest level's data in the caches
    auto load_chunk = [&](const double *row, int64_t ld, int64_t c0, double (&a)[NK]) {
#pragma unroll
        for (int k = 0; k < NK; ++k)
            a[k] = (c0 + 64 * k < ld) ? (NT ? __builtin_nontemporal_load(gmem(row) + c0 + 64 * k + lane)
                                            : gmem(row)[c0 + 64 * k + lane])
                                      : 0.0;
    };
    // A stage in flight: this wavefront's problem and row (one row per wavefront), whether the
    // workgroup has rows at all, and the stage's operands that depend on no earlier stage, in
    // registers: the M1 row with this lane's entries of v1 (ld1 <= CH: one chunk) and the first M2 chunk.
    struct Cur {
        int qi, r, gi;  // gi: lane 0's full-grid DOF of the row's solution entry (yoff >= 0)
        bool any;
        double m1[NK], v1[NK], m2[NK];
    };
    // Part A of stage s: its loads, issued without waiting for any of them.
    auto part_a = [&](int s, Cur &c) {
        const PStage &S = stages[s];
        c.qi = (S.nq == 2 && g >= S.gA) ? 1 : 0;
        c.r = 4 * (c.qi ? g - S.gA : g) + w;
        const PRow &q = S.q[c.qi];
        c.any = c.r - w < q.n;
        c.gi = (q.yoff >= 0 && lane == 0 && c.r < q.n) ? o.gmap[q.yoff + c.r] : 0;
        if (c.r < q.n) {
            if (q.M1) {
                load_chunk(q.M1 + (int64_t)c.r * q.ld1, q.ld1, 0, c.m1);
#pragma unroll
                for (int k = 0; k < NK; ++k) c.v1[k] = 64 * k + lane < q.n1 ? gmem(q.v1)[64 * k + lane] : 0.0;
            }
            if (q.M2) load_chunk(q.M2 + (int64_t)c.r * q.ld2, q.ld2, 0, c.m2);
        }
    };
    // Part B: the earlier stages' vectors, the products and the store.
    auto part_b = [&](int s, Cur &c) {
        const PRow &q = stages[s].q[c.qi];
        const int r = c.r;
        double *sv2 = sv + (s & 1) * L2;
        double cval = 0.0;
        unsigned long long cbits = 0;
        if (q.c && lane == 0 && r < q.n) cbits = coh_bits(q.c + r);  // in the same round trip as the staging
        if (c.any && q.M2) stage_vec<true>(sv2, chunk_len(q.ld2), q.v2a, q.n2a, q.v2b, q.n2b, err, max_polls);
        if (q.c && lane == 0 && r < q.n) cval = await_value(q.c + r, cbits, err, max_polls);
        lds_barrier();  // unconditional (also in workgroups without rows in the stage)
        if (r < q.n) {
            double t1 = 0.0, t2 = 0.0;
            if (q.M1) {
#pragma unroll
                for (int k = 0; k < NK; ++k) t1 = fma(c.m1[k], c.v1[k], t1);
            }
            if (q.M2) {
#pragma unroll
                for (int k = 0; k < NK; ++k) t2 = fma(c.m2[k], sv2[64 * k + lane], t2);
                const double *row = q.M2 + (int64_t)r * q.ld2;
                for (int64_t c0 = CH; c0 < q.ld2; c0 += CH) {
                    double a[NK];
                    load_chunk(row, q.ld2, c0, a);
#pragma unroll
                    for (int k = 0; k < NK; ++k) t2 = fma(a[k], sv2[c0 + 64 * k + lane], t2);
                }
            }
            double t = t1 - t2;
            for (int sh = 32; sh > 0; sh >>= 1) t += __shfl_xor(t, sh, 64);
            if (lane == 0) {
                const double val = cval + t;
                coh_store(q.y + r, val);
                if (q.yoff >= 0) {
                    if (o.x64) o.x64[c.gi] = val;
                    else o.x32[c.gi] = (float)val;
                }
            }
        }
    };
    // Two stages in flight: part A of stage s + 1 (its loads) is issued before part B of stage s waits
    // for its input; two register sets in turn.  LDS: part B of stage s writes buffer s & 1 before its
    // barrier and reads it after; the next write of that buffer (stage s + 2) follows the barrier of
    // stage s + 1, which every wave reaches after its reads of stage s.
    Cur ca, cb;
    part_a(0, ca);
    for (int s = 0; s < nstages; s += 2) {
        if (s + 1 < nstages) part_a(s + 1, cb);
        part_b(s, ca);
        if (s + 1 == nstages) break;
        if (s + 2 < nstages) part_a(s + 2, ca);
        part_b(s + 1, cb);
    }
}

template <typename T>
__global__ void k_pack_planes(const T *__restrict__ f, double *__restrict__ b, const uint8_t *__restrict__ mask,
                              const int32_t *__restrict__ lidx, const int64_t *__restrict__ off, const Grid g, int axis,
                              double *__restrict__ e1 = nullptr, double *__restrict__ e2 = nullptr) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes() || !(mask[n] & NODE_ACTIVE)) return;
    const int y = (int)(n % g.NY);
    const int64_t t = n / g.NY;
    const int x = (int)(t % g.NX), z = (int)(t / g.NX);
    const int64_t base = off[coord(axis, y, x, z)] + 3 * (int64_t)lidx[n];
    b[base] = (double)f[3 * n];
    b[base + 1] = (double)f[3 * n + 1];
    b[base + 2] = (double)f[3 * n + 2];
    if (e1) {  // the persistent solve's vectors start out "not written" (every plane slot is an active node's)
        const double empty = __longlong_as_double((long long)PERS_EMPTY);
        for (int c = 0; c < 3; ++c) {
            e1[base + c] = empty;
            e2[base + c] = empty;
        }
    }
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

}  // namespace

CoarseSolver::~CoarseSolver() {
    if (exec) (void)hipGraphExecDestroy(exec);
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
    npad.assign(P, 0);
    off.assign(P + 1, 0);
    sOff.assign(P + 1, 0);
    int64_t maxpad = CB;
    for (int p = 0; p < P; ++p) {
        nd[p] = 3 * count[p];
        npad[p] = ((nd[p] + CB - 1) / CB) * CB;
        maxpad = std::max(maxpad, npad[p]);
        off[p + 1] = off[p] + nd[p];
        sOff[p + 1] = sOff[p] + npad[p] * npad[p];
    }
    ntot = off[P];
    m = P / 2;
    fOff.assign(P + 1, 0);
    bOff.assign(P + 1, 0);
    fLd.assign(P, 0);
    bLd.assign(P, 0);
    for (int p = 0; p < P; ++p) {
        const int64_t prev = p > 0 ? nd[p - 1] : 0, next = p + 1 < P ? nd[p + 1] : 0;
        // forward operator columns: the plane it is eliminated against (both neighbours for m)
        const int64_t fc = p < m ? prev : (p > m ? next : prev + next);
        // backward operator columns: the plane it is back-substituted from (none for m)
        const int64_t bc = p < m ? next : (p > m ? prev : 0);
        fLd[p] = fc ? chain_ld(fc) : 0;
        bLd[p] = bc ? chain_ld(bc) : 0;
        fOff[p + 1] = fOff[p] + nd[p] * fLd[p];
        bOff[p + 1] = bOff[p] + nd[p] * bLd[p];
    }
    std::vector<int32_t> rp(std::max<int64_t>(ntot, 1), 0);
    for (int p = 0; p < P; ++p)
        for (int64_t r = off[p]; r < off[p + 1]; ++r) rp[r] = p;
    // plane-order DOF -> full-grid DOF, and the inactive nodes' DOFs (32-bit: the persistent solve's
    // output map; it is used only where 3 x the nodes fit)
    std::vector<int32_t> gm, gz;  // kept until the stream synchronisation below (asynchronous uploads)
    if (3 * g.nodes() < (1ll << 31)) {
        gm.assign(std::max<int64_t>(ntot, 1), 0);
        for (int64_t n = 0; n < g.nodes(); ++n) {
            const int64_t y = n % g.NY, x = (n / g.NY) % g.NX, z = n / ((int64_t)g.NY * g.NX);
            const int p = (int)(axis == 0 ? y : (axis == 1 ? x : z));
            for (int c = 0; c < 3; ++c) {
                if (lid[n] >= 0) gm[off[p] + 3 * (int64_t)lid[n] + c] = (int32_t)(3 * n + c);
                else gz.push_back((int32_t)(3 * n + c));
            }
        }
        gmap.upload(gm, st);
        nzero = (int64_t)gz.size();
        if (nzero) gzero.upload(gz, st);
    } else {
        gmap.release();
        gzero.release();
        nzero = 0;
    }
    lidx.upload(lid, st);
    d_off.upload(off, st);
    d_sOff.upload(sOff, st);
    d_npad.upload(npad, st);
    d_nd.upload(nd, st);
    rowPlane.upload(rp, st);
    Sinv.alloc(std::max<int64_t>(sOff[P], 1));
    dnV.alloc(std::max<int64_t>(ntot * NCPL, 1));
    upV.alloc(std::max<int64_t>(ntot * NCPL, 1));
    dnC.alloc(std::max<int64_t>(ntot * NCPL, 1));
    upC.alloc(std::max<int64_t>(ntot * NCPL, 1));
    Fop.alloc(std::max<int64_t>(fOff[P], 1));
    Bop.alloc(std::max<int64_t>(bOff[P], 1));
    Fop.zero(st);  // row padding stays zero
    Bop.zero(st);
    T.alloc(2 * maxpad * CB);
    Rw.alloc(2 * maxpad * CB);
    Dk.alloc(2 * CB * CB);
    bb.alloc(std::max<int64_t>(ntot, 1));
    yy.alloc(std::max<int64_t>(ntot, 1));
    xx.alloc(std::max<int64_t>(ntot, 1));
    info.alloc(P);
    pers_built = false;  // the persistent plan points into the buffers above
    TG_HIP(hipStreamSynchronize(st), "tg_solver:coarse");
}

void CoarseSolver::factor(const double *Ks, int64_t ldk, const uint8_t *eleMask, const uint8_t *nodeMask,
                          hipStream_t st) {
    const char *id = "tg_solver:coarse";
    factored = false;
    int64_t maxpad = CB, maxn = 1;
    for (int p = 0; p < P; ++p) {
        maxpad = std::max(maxpad, npad[p]);
        maxn = std::max(maxn, nd[p]);
    }
    k_plane_init<<<dim3(grid_for(maxpad * maxpad, 256), P), 256, 0, st>>>(Sinv, d_sOff, d_npad, d_nd);
    info.zero(st);
    k_coarse_assemble<<<grid_for(g.nodes() * 243, 256), 256, 0, st>>>(Ks, ldk, eleMask, nodeMask, g, axis, lidx, d_off,
                                                                      d_sOff, d_npad, Sinv, dnV, dnC, upV, upC);
    TG_LAUNCH_CHECK(id);
    auto S_of = [&](int p) { return Sinv.p + sOff[p]; };
    auto has = [&](int p) { return p >= 0 && p < P && nd[p] > 0; };
    const SchurProb noneS{};
    // S_p -= A_{p,p-1} B_{p-1} (chain A, dn rows of p) and / or A_{p,p+1} B_{p+1} (chain B, up rows of p)
    auto schur = [&](int p, bool withPrev, bool withNext) {
        SchurProb q{};
        q.S = S_of(p);
        q.np = npad[p];
        q.n = nd[p];
        if (withPrev && has(p - 1))
            q.t[q.nt++] = SchurTerm{dnV.p + off[p] * NCPL, dnC.p + off[p] * NCPL, Bop.p + bOff[p - 1], bLd[p - 1]};
        if (withNext && has(p + 1))
            q.t[q.nt++] = SchurTerm{upV.p + off[p] * NCPL, upC.p + off[p] * NCPL, Bop.p + bOff[p + 1], bLd[p + 1]};
        return q;
    };
    auto launch_schur = [&](const SchurProb &A, const SchurProb &B) {
        if (!A.nt && !B.nt) return;
        const int64_t n = std::max(A.nt ? A.n : 0, B.nt ? B.n : 0);
        k_schur<<<dim3(grid_for(n, 256), (unsigned)n, 2), 256, 0, st>>>(A.nt ? A : noneS, B.nt ? B : noneS);
        TG_LAUNCH_CHECK(id);
    };
    // S_p <- S_p^-1 for up to two planes together
    auto sweep = [&](int p, int q) {
        SweepProb pr[2] = {};
        int nbmax = 0;
        const int pl[2] = {p, q};
        for (int y = 0; y < 2; ++y) {
            if (!has(pl[y])) continue;
            pr[y] = SweepProb{S_of(pl[y]), npad[pl[y]], (int)(npad[pl[y]] / CB), T.p + y * maxpad * CB,
                              Rw.p + y * maxpad * CB, Dk.p + y * CB * CB, info.p + pl[y]};
            nbmax = std::max(nbmax, pr[y].nb);
        }
        for (int K = 0; K < nbmax; ++K) {
            k_sweep_panel<<<dim3(nbmax, 2), 256, 0, st>>>(pr[0], pr[1], K);
            k_sweep_update<<<dim3(nbmax * nbmax, 2), 256, 0, st>>>(pr[0], pr[1], K);
        }
        TG_LAUNCH_CHECK(id);
    };
    // solve operators: Out = S_p^-1 A_{p,r} for neighbour plane r, from r's coupling rows towards p
    std::vector<DSProb> ops;
    auto op = [&](int p, int r, double *out, int64_t ld, int64_t col0) {
        if (!has(p) || !has(r)) return;
        const bool rIsNext = r == p + 1;  // A_{p,p+1}(., j) = A_{p+1,p}(j, .) = dn rows of p+1; else up rows of p-1
        DSProb d;
        d.S = S_of(p);
        d.np = npad[p];
        d.nrows = nd[p];
        d.ncn = nd[r] / 3;
        d.V = (rIsNext ? dnV.p : upV.p) + off[r] * NCPL;
        d.C = (rIsNext ? dnC.p : upC.p) + off[r] * NCPL;
        d.Out = out;
        d.ld = ld;
        d.col0 = col0;
        ops.push_back(d);
    };
    auto flush_ops = [&]() {
        for (size_t i0 = 0; i0 < ops.size(); i0 += 4) {
            DSProb q[4] = {};
            int64_t work = 0;
            const int cnt = (int)std::min<size_t>(4, ops.size() - i0);
            for (int y = 0; y < cnt; ++y) {
                q[y] = ops[i0 + y];
                work = std::max(work, 3 * q[y].ncn * ((q[y].nrows + DS_ROWS - 1) / DS_ROWS));
            }
            if (work) k_dense_sparse<<<dim3(grid_for(work, 256), cnt), 256, 0, st>>>(q[0], q[1], q[2], q[3]);
        }
        TG_LAUNCH_CHECK(id);
        ops.clear();
    };
    // chains A (p = 0 .. m-1) and B (q = P-1 .. m+1) advance together, one plane each per step
    const int lenA = m, lenB = P - 1 - m;
    for (int k = 0; k < std::max(lenA, lenB); ++k) {
        const int p = k < lenA && has(k) ? k : -1;
        const int q = k < lenB && has(P - 1 - k) ? P - 1 - k : -1;
        launch_schur(p > 0 ? schur(p, true, false) : noneS, q >= 0 && q < P - 1 ? schur(q, false, true) : noneS);
        sweep(p, q);
        if (p >= 0) {
            op(p, p + 1, Bop.p + bOff[p], bLd[p], 0);  // B_p = S_p^-1 A_{p,p+1}
            op(p, p - 1, Fop.p + fOff[p], fLd[p], 0);  // F_p = S_p^-1 A_{p,p-1}
        }
        if (q >= 0) {
            op(q, q - 1, Bop.p + bOff[q], bLd[q], 0);  // B_q = S_q^-1 A_{q,q-1}
            op(q, q + 1, Fop.p + fOff[q], fLd[q], 0);  // F_q = S_q^-1 A_{q,q+1}
        }
        flush_ops();
    }
    if (has(m)) {  // separator
        launch_schur(schur(m, true, true), noneS);
        sweep(m, -1);
        const int64_t na = has(m - 1) ? nd[m - 1] : 0;
        op(m, m - 1, Fop.p + fOff[m], fLd[m], 0);
        op(m, m + 1, Fop.p + fOff[m], fLd[m], na);
        flush_ops();
    }
    std::vector<int> hinfo(P);
    TG_HIP(hipMemcpyAsync(hinfo.data(), info.p, sizeof(int) * P, hipMemcpyDeviceToHost, st), id);
    TG_HIP(hipStreamSynchronize(st), id);
    for (int p = 0; p < P; ++p)
        if (hinfo[p] != 0)
            fail(TG_ERR_NUMERIC, "tg_solver:coarse",
                 "coarsest-level factorisation failed on plane block %d (pivot %d): matrix not positive definite",
                 p, hinfo[p]);
    factored = true;
}

// The solve's fixed sequence on the internal plane-ordered vectors (bb -> yy, in place on yy):
//   yy = blockdiag(S_p^-1) bb; forward chains A (p = 1 .. m-1) and B (p = P-2 .. m+1) step by step together;
//   the separator; backward chains A (p = m-1 .. 0) and B (p = m+1 .. P-1).
void CoarseSolver::enqueue_core(hipStream_t s) const {
    const char *id = "tg_solver:coarse";
    const ChainProb none = chain_prob(nullptr, 0, 0, 0, 0, nullptr, nullptr, nullptr, nullptr);
    auto ok = [&](int p) { return p >= 0 && p < P && nd[p] > 0; };
    auto fwd = [&](int p, int q) -> ChainProb {  // z_p = c_p - F_p z_q
        if (!ok(p) || !ok(q)) return none;
        return chain_prob(Fop.p + fOff[p], nd[p], nd[q], nd[q], fLd[p], yy.p + off[p], yy.p + off[q], nullptr,
                          yy.p + off[p]);
    };
    auto bwd = [&](int p, int q) -> ChainProb {  // x_p = z_p - B_p x_q
        if (!ok(p) || !ok(q)) return none;
        return chain_prob(Bop.p + bOff[p], nd[p], nd[q], nd[q], bLd[p], yy.p + off[p], yy.p + off[q], nullptr,
                          yy.p + off[p]);
    };
    if (ntot) k_plane_gemv<<<grid_for(ntot, 4), 256, 0, s>>>(Sinv, d_sOff, d_npad, d_off, rowPlane, ntot, bb, yy);
    const int lenA = m, lenB = P - 1 - m;  // planes per chain
    for (int k = 1; k < std::max(lenA, lenB); ++k)
        launch_chain(k < lenA ? fwd(k, k - 1) : none, k < lenB ? fwd(P - 1 - k, P - k) : none, s);
    if (ok(m)) {  // separator: x_m = c_m - F_m [z_{m-1}; z_{m+1}]
        const int64_t na = ok(m - 1) ? nd[m - 1] : 0, nb = ok(m + 1) ? nd[m + 1] : 0;
        if (na + nb > 0)
            launch_chain(chain_prob(Fop.p + fOff[m], nd[m], na + nb, na, fLd[m], yy.p + off[m],
                                    na ? yy.p + off[m - 1] : yy.p, nb ? yy.p + off[m + 1] : yy.p, yy.p + off[m]),
                         none, s);
    }
    for (int k = 1; k <= std::max(lenA, lenB); ++k)
        launch_chain(k <= lenA ? bwd(m - k, m - k + 1) : none, k <= lenB ? bwd(m + k, m + k - 1) : none, s);
    TG_LAUNCH_CHECK(id);
}

// The stage table of the persistent solve: the same recurrences as enqueue_core, with the block-diagonal
// product folded into the forward stages (stage 0 holds the chains' first planes, z = S^-1 b alone).
void CoarseSolver::build_persistent(hipStream_t st) {
    const char *id = "tg_solver:coarse";
    pers_built = true;
    pers_ok = false;
    auto ok = [&](int p) { return p >= 0 && p < P && nd[p] > 0; };
    int64_t maxpad = 0, maxld1 = 0, maxld2 = 0;
    for (int p = 0; p < P; ++p) maxpad = std::max(maxpad, npad[p]);
    const int nk = (int)((maxpad + 63) / 64);
    if (nk < 1 || nk > 16) return;  // planes wider than 1024 DOFs: step by step
    if (!gmap.p) return;            // no 32-bit output map (more than 2^31 / 3 nodes)
    PRow none{};
    // forward results (and the chains' inputs) in zz = yy, the separator's and backward results in xx
    const double *zz = yy.p;
    double *xxp = xx.p;
    auto row = [&](int p, double *y, bool diag, const double *M2, int64_t ld2, const double *va, int qa,
                   const double *vb, int qb, const double *c) {
        PRow q{};
        q.n = (int)nd[p];
        q.y = y + off[p];
        q.yoff = y == xxp ? (int)off[p] : -1;  // the separator's and backward results are the solution
        if (diag) {
            q.M1 = Sinv.p + sOff[p];
            q.ld1 = npad[p];
            q.v1 = bb.p + off[p];
            q.n1 = (int)nd[p];
            maxld1 = std::max(maxld1, q.ld1);
        }
        if (M2 && (qa >= 0 || qb >= 0)) {
            q.M2 = M2;
            q.ld2 = ld2;
            q.v2a = qa >= 0 ? va + off[qa] : va;
            q.n2a = qa >= 0 ? (int)nd[qa] : 0;
            q.v2b = qb >= 0 ? vb + off[qb] : vb;
            q.n2b = qb >= 0 ? (int)nd[qb] : 0;
            maxld2 = std::max(maxld2, q.ld2);
        }
        if (c) q.c = c + off[p];
        return q;
    };
    std::vector<PStage> tab;
    auto add = [&](const PRow &a, const PRow &b) {
        PStage S{};
        const bool ha = a.n > 0, hb = b.n > 0;
        if (!ha && !hb) return;
        S.q[0] = ha ? a : b;
        S.q[1] = ha && hb ? b : none;
        S.nq = ha && hb ? 2 : 1;
        tab.push_back(S);
    };
    const int lenA = m, lenB = P - 1 - m;
    for (int k = 0; k < std::max(lenA, lenB); ++k) {  // forward: z_p = S_p^-1 b_p - F_p z_{p-1|p+1}
        const int p = k, q = P - 1 - k;
        PRow a = none, b = none;
        if (k < lenA && ok(p))
            a = row(p, yy.p, true, Fop.p + fOff[p], fLd[p], zz, ok(p - 1) && k > 0 ? p - 1 : -1, zz, -1, nullptr);
        if (k < lenB && ok(q))
            b = row(q, yy.p, true, Fop.p + fOff[q], fLd[q], zz, ok(q + 1) && k > 0 ? q + 1 : -1, zz, -1, nullptr);
        add(a, b);
    }
    if (ok(m))  // separator: x_m = S_m^-1 b_m - F_m [z_{m-1}; z_{m+1}]
        add(row(m, xxp, true, Fop.p + fOff[m], fLd[m], zz, ok(m - 1) ? m - 1 : -1, zz, ok(m + 1) ? m + 1 : -1,
                nullptr),
            none);
    for (int k = 1; k <= std::max(lenA, lenB); ++k) {  // backward: x_p = z_p - B_p x_{p+1|p-1}
        const int p = m - k, q = m + k;
        PRow a = none, b = none;
        if (k <= lenA && ok(p)) a = row(p, xxp, false, Bop.p + bOff[p], bLd[p], xxp, ok(p + 1) ? p + 1 : -1, xxp, -1, zz);
        if (k <= lenB && ok(q)) b = row(q, xxp, false, Bop.p + bOff[q], bLd[q], xxp, ok(q - 1) ? q - 1 : -1, xxp, -1, zz);
        add(a, b);
    }
    if (tab.empty()) return;
    const int64_t ch = 64 * (int64_t)nk;
    const int64_t L1 = ((std::max<int64_t>(maxld1, 1) + ch - 1) / ch) * ch;
    if (L1 > ch) return;  // the kernel keeps an M1 row in one chunk of registers
    const int64_t L2 = ((std::max<int64_t>(maxld2, 1) + ch - 1) / ch) * ch;
    const size_t lds = 2 * (size_t)L2 * sizeof(double);  // two stage buffers
    if (lds > 64 * 1024) return;
    int dev = 0, cus = 0, per_cu = 0;
    TG_HIP(hipGetDevice(&dev), id);
    TG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), id);
    switch (nk) {
#define TG_PERS_CASE(K)                                                                                   \
    case K:                                                                                               \
        TG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_coarse_persistent<K, true>, 256, lds), id); \
        break;
        TG_PERS_CASE(1) TG_PERS_CASE(2) TG_PERS_CASE(3) TG_PERS_CASE(4) TG_PERS_CASE(5) TG_PERS_CASE(6)
        TG_PERS_CASE(7) TG_PERS_CASE(8) TG_PERS_CASE(9) TG_PERS_CASE(10) TG_PERS_CASE(11) TG_PERS_CASE(12)
        TG_PERS_CASE(13) TG_PERS_CASE(14) TG_PERS_CASE(15) TG_PERS_CASE(16)
#undef TG_PERS_CASE
    }
    if (per_cu < 1 || cus < 1) return;
    const int G = std::min(per_cu, 2) * cus;
    for (PStage &S : tab) {  // one row per wavefront, or the step-by-step sequence
        if (S.nq == 2) {
            const int64_t na = S.q[0].n, nb = S.q[1].n;
            S.gA = (int)std::min<int64_t>(G - 1, std::max<int64_t>(1, (G * na + (na + nb) / 2) / (na + nb)));
            if (na > 4 * (int64_t)S.gA || nb > 4 * (int64_t)(G - S.gA)) return;
        } else if (S.q[0].n > 4 * (int64_t)G) {
            return;
        }
    }
    std::vector<unsigned char> bytes(tab.size() * sizeof(PStage));
    std::memcpy(bytes.data(), tab.data(), bytes.size());
    pers_table.upload(bytes, st);
    pers_err.alloc(1);
    pers_err.zero(st);
    TG_HIP(hipStreamSynchronize(st), id);
    pers_pending = false;
    pers_nk = nk;
    pers_groups = G;
    pers_stages = (int)tab.size();
    pers_L1 = (int)L1;
    pers_L2 = (int)L2;
    pers_ok = true;
}

void CoarseSolver::enqueue_persistent(hipStream_t st, double *x64, float *x32) {
    const char *id = "tg_solver:coarse";
    const size_t lds = 2 * (size_t)pers_L2 * sizeof(double);
    const PStage *tab = (const PStage *)pers_table.p;
    const PersOut o{x64, x32, gmap.p, gzero.p, (int)nzero};
    switch (pers_nk) {
#define TG_PERS_CASE(K)                                                                                   \
    case K:                                                                                               \
        if (pers_streaming)                                                                                \
            k_coarse_persistent<K, true><<<pers_groups, 256, lds, st>>>(tab, pers_stages, pers_L2, pers_err.p, \
                                                                        pers_polls, o);                     \
        else                                                                                              \
            k_coarse_persistent<K, false><<<pers_groups, 256, lds, st>>>(tab, pers_stages, pers_L2, pers_err.p, \
                                                                         pers_polls, o);                    \
        break;
        TG_PERS_CASE(1) TG_PERS_CASE(2) TG_PERS_CASE(3) TG_PERS_CASE(4) TG_PERS_CASE(5) TG_PERS_CASE(6)
        TG_PERS_CASE(7) TG_PERS_CASE(8) TG_PERS_CASE(9) TG_PERS_CASE(10) TG_PERS_CASE(11) TG_PERS_CASE(12)
        TG_PERS_CASE(13) TG_PERS_CASE(14) TG_PERS_CASE(15) TG_PERS_CASE(16)
#undef TG_PERS_CASE
        default: fail(TG_ERR_ARGS, id, "bad persistent chunking");
    }
    TG_LAUNCH_CHECK(id);
    pers_pending = true;
}

void CoarseSolver::check_barrier(hipStream_t st) {
    if (!pers_pending) return;
    const char *id = "tg_solver:coarse";
    int e = 0;
    TG_HIP(hipMemcpyAsync(&e, pers_err.p, sizeof(int), hipMemcpyDeviceToHost, st), id);
    TG_HIP(hipStreamSynchronize(st), id);
    pers_pending = false;
    if (e) {
        pers_err.zero(st);
        TG_HIP(hipStreamSynchronize(st), id);
        fail(TG_ERR_HIP, id, "coarsest-level solve: a stage of the persistent launch timed out waiting for its input");
    }
}

bool CoarseSolver::plane_pack(PlanePack &pk, hipStream_t st) {
    if (!factored || !use_persistent || use_graph) return false;
    if (!pers_built) build_persistent(st);
    if (!pers_ok) return false;
    pk = PlanePack{bb.p, yy.p, xx.p, lidx.p, d_off.p, axis, PERS_EMPTY};
    return true;
}

template <typename T>
void CoarseSolver::solve(const T *r, T *x, const uint8_t *nodeMask, hipStream_t st, bool packed) {
    const char *id = "tg_solver:coarse";
    if (!factored) fail(TG_ERR_STATE, id, "coarsest level not factorised (call assemble first)");
    if (use_persistent && !use_graph && !pers_built) build_persistent(st);
    const bool pers = use_persistent && !use_graph && pers_ok;
    if (packed && !pers) fail(TG_ERR_STATE, id, "right-hand side packed for a persistent solve that is not set up");
    if (!packed) {
        k_pack_planes<T><<<grid_for(g.nodes(), 256), 256, 0, st>>>(r, bb, nodeMask, lidx, d_off, g, axis,
                                                                   pers ? yy.p : nullptr, pers ? xx.p : nullptr);
        TG_LAUNCH_CHECK(id);
    }
    if (pers) {
        // the launch writes the solution on the full grid itself (no unpack)
        if constexpr (sizeof(T) == 8) enqueue_persistent(st, (double *)x, nullptr);
        else enqueue_persistent(st, nullptr, (float *)x);
        return;
    } else if (!use_graph) {
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
    k_unpack_planes<T><<<grid_for(g.nodes(), 256), 256, 0, st>>>(yy, x, nodeMask, lidx, d_off, g, axis);
    TG_LAUNCH_CHECK(id);
}

template void CoarseSolver::solve<double>(const double *, double *, const uint8_t *, hipStream_t, bool);
template void CoarseSolver::solve<float>(const float *, float *, const uint8_t *, hipStream_t, bool);

}  // namespace tg
