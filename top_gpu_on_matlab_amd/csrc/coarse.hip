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

__global__ void k_lower_only(const double *__restrict__ S, double *__restrict__ L, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    const int64_t r = i % n, c = i / n;
    L[i] = (r >= c) ? S[i] : 0.0;
}

// Row-major copy of a column-major n x n block: Tr[r * n + c] = A[r + n * c].
__global__ void k_transpose_block(const double *__restrict__ A, double *__restrict__ Tr, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    const int64_t r = i / n, c = i - r * n;
    Tr[i] = A[r + n * c];
}

// Batched triangular products over all planes, one wavefront per row (row R of the stacked
// vector, plane rowPlane[R]):  out_p = Tri_p * in_p, Tri_p row-major n_p x n_p, lower (columns
// 0..r) or upper (columns r..n_p-1) triangular.
__global__ __launch_bounds__(256) void k_tri_rows(const double *__restrict__ Tri, const int64_t *__restrict__ adOff,
                                                  const int64_t *__restrict__ off, const int32_t *__restrict__ rowPlane,
                                                  int64_t ntot, int upper, const double *__restrict__ in,
                                                  double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t R = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (R >= ntot) return;
    const int p = rowPlane[R];
    const int64_t o = off[p], n = off[p + 1] - o, r = R - o;
    const double *row = Tri + adOff[p] + r * n;
    const double *x = in + o;
    const int64_t c0 = upper ? r : 0, c1 = upper ? n : r + 1;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t c = c0 + lane; c < c1; c += 8 * 64) {
        // loads are unconditional (clamped index) so all eight are in flight before the first use
        double a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t cc = min(c + 64 * k, c1 - 1);
            a[k] = row[cc];
            b[k] = x[cc];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += (c + 64 * k < c1) ? a[k] * b[k] : 0.0;
    }
    double t = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    for (int o2 = 32; o2 > 0; o2 >>= 1) t += __shfl_down(t, o2, 64);
    if (lane == 0) out[R] = t;
}

// One step of a block recurrence, one wavefront per row:  y[r] = c[r] - sum_j M[r * ldm + j] v[j]
// (M row-major n x m, rows zero-padded to ldm = nchunks * NK * 64; y may alias c).  The block
// stages v in LDS chunk by chunk; each lane issues its NK matrix loads of a chunk before the
// first use (no guards: the padding is zero), so a chunk costs one memory round trip.
template <int NK>
__global__ __launch_bounds__(256) void k_chain_rows(const double *__restrict__ M, int64_t n, int64_t m, int64_t ldm,
                                                    int nchunks, const double *c, const double *__restrict__ v,
                                                    double *y) {
    __shared__ double sv[NK * 64];
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const double *row = M + (r < n ? r : 0) * ldm + lane;
    double t = 0;
    for (int q = 0; q < nchunks; ++q) {
        double a[NK];
#pragma unroll
        for (int k = 0; k < NK; ++k) a[k] = row[(int64_t)q * NK * 64 + 64 * k];
        __syncthreads();
        for (int j = threadIdx.x; j < NK * 64; j += 256) {
            const int64_t jj = (int64_t)q * NK * 64 + j;
            sv[j] = jj < m ? v[jj] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NK; ++k) t += a[k] * sv[lane + 64 * k];
    }
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
    if (lane == 0 && r < n) y[r] = c[r] - t;
}

// Padded row length (multiple of 64, or of 1024 beyond 1024) and the chunking used for it.
inline int64_t chain_ld(int64_t m) { return m <= 1024 ? ((m + 63) / 64) * 64 : ((m + 1023) / 1024) * 1024; }

void launch_chain(const double *M, int64_t n, int64_t m, const double *c, const double *v, double *y, hipStream_t s) {
    const int64_t ld = chain_ld(m);
    const unsigned gr = grid_for(n, 4);
    const int nk = (int)(ld <= 1024 ? ld / 64 : 16), nch = (int)(ld <= 1024 ? 1 : ld / 1024);
    switch (nk) {
#define TG_CHAIN_CASE(K) \
    case K: k_chain_rows<K><<<gr, 256, 0, s>>>(M, n, m, ld, nch, c, v, y); break;
        TG_CHAIN_CASE(1) TG_CHAIN_CASE(2) TG_CHAIN_CASE(3) TG_CHAIN_CASE(4) TG_CHAIN_CASE(5) TG_CHAIN_CASE(6)
        TG_CHAIN_CASE(7) TG_CHAIN_CASE(8) TG_CHAIN_CASE(9) TG_CHAIN_CASE(10) TG_CHAIN_CASE(11) TG_CHAIN_CASE(12)
        TG_CHAIN_CASE(13) TG_CHAIN_CASE(14) TG_CHAIN_CASE(15) TG_CHAIN_CASE(16)
#undef TG_CHAIN_CASE
        default: fail(TG_ERR_ARGS, "tg_solver:coarse", "bad chain chunking");
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
    adOff.assign(P + 1, 0);
    asOff.assign(P + 1, 0);
    mOff.assign(P + 1, 0);
    hOff.assign(P + 1, 0);
    for (int p = 0; p < P; ++p) {
        const int64_t prev = p > 0 ? nd[p - 1] : 0, next = p + 1 < P ? nd[p + 1] : 0;
        adOff[p + 1] = adOff[p] + nd[p] * nd[p];
        asOff[p + 1] = asOff[p] + nd[p] * prev;
        mOff[p + 1] = mOff[p] + nd[p] * (prev ? chain_ld(prev) : 0);
        hOff[p + 1] = hOff[p] + nd[p] * (next ? chain_ld(next) : 0);
    }
    std::vector<int32_t> rp(std::max<int64_t>(ntot, 1), 0);
    for (int p = 0; p < P; ++p)
        for (int64_t R = off[p]; R < off[p + 1]; ++R) rp[R] = p;
    lidx.upload(lid, st);
    d_off.upload(off, st);
    d_adOff.upload(adOff, st);
    d_asOff.upload(asOff, st);
    rowPlane.upload(rp, st);
    Ad.alloc(adOff[P]);
    Linv.alloc(adOff[P]);
    LT.alloc(adOff[P]);
    As.alloc(std::max<int64_t>(asOff[P], 1));
    Mf.alloc(std::max<int64_t>(mOff[P], 1));
    Hm.alloc(std::max<int64_t>(hOff[P], 1));
    Mf.zero(st);  // row padding stays zero
    Hm.zero(st);
    bb.alloc(ntot);
    yy.alloc(ntot);
    xx.alloc(ntot);
    info.alloc(2 * P);
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
    for (int p = 0; p < P; ++p) {
        const rocblas_int np = (rocblas_int)nd[p];
        if (np == 0) continue;
        double *S = Ad.p + adOff[p];
        if (p > 0 && nd[p - 1] > 0)
            rb(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, np, (rocblas_int)nd[p - 1], &mone,
                             As.p + asOff[p], np, &one, S, np), "rocblas_dsyrk");
        rb(rocsolver_dpotrf(h, rocblas_fill_lower, np, S, np, info.p + p), "rocsolver_dpotrf");
        k_lower_only<<<grid_for((int64_t)np * np, 256), 256, 0, st>>>(S, Linv.p + adOff[p], np);
        TG_LAUNCH_CHECK(id);
        rb(rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, np, Linv.p + adOff[p], np, info.p + P + p),
           "rocsolver_dtrtri");
        if (p + 1 < P && nd[p + 1] > 0)
            rb(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                             rocblas_diagonal_non_unit, (rocblas_int)nd[p + 1], np, &one, S, np, As.p + asOff[p + 1],
                             (rocblas_int)nd[p + 1]), "rocblas_dtrsm");
    }
    // Solve operators (all row-major): LT_p = inv(L_pp) and UT_p = inv(L_pp)^T (the column-major
    // Linv itself), Mf_p = inv(L_pp) L_{p,p-1} (n_p x n_{p-1}, from (As_p^T Linv_p^T) column-major)
    // and Hm_p = inv(L_pp)^T L_{p+1,p}^T (n_p x n_{p+1}, from L_{p+1,p} inv(L_pp) column-major), the
    // last two with rows zero-padded to chain_ld.
    for (int p = 0; p < P; ++p) {
        const rocblas_int np = (rocblas_int)nd[p];
        if (np == 0) continue;
        const int64_t prev = p > 0 ? nd[p - 1] : 0, next = p + 1 < P ? nd[p + 1] : 0;
        k_transpose_block<<<grid_for((int64_t)np * np, 256), 256, 0, st>>>(Linv.p + adOff[p], LT.p + adOff[p], np);
        TG_LAUNCH_CHECK(id);
        if (prev > 0)
            rb(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_transpose, (rocblas_int)prev, np, np,
                             &one, As.p + asOff[p], np, Linv.p + adOff[p], np, &zero, Mf.p + mOff[p],
                             (rocblas_int)chain_ld(prev)), "rocblas_dgemm Mf");
        if (next > 0)
            rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)next, np, np, &one,
                             As.p + asOff[p + 1], (rocblas_int)next, Linv.p + adOff[p], np, &zero,
                             Hm.p + hOff[p], (rocblas_int)chain_ld(next)), "rocblas_dgemm Hm");
    }
    std::vector<rocblas_int> hinfo(2 * P);
    TG_HIP(hipMemcpyAsync(hinfo.data(), info.p, sizeof(rocblas_int) * 2 * P, hipMemcpyDeviceToHost, st), id);
    TG_HIP(hipStreamSynchronize(st), id);
    for (int p = 0; p < 2 * P; ++p)
        if (hinfo[p] != 0)
            fail(TG_ERR_NUMERIC, "tg_solver:coarse",
                 "coarsest-level factorisation failed on plane block %d (info %d): matrix not positive definite",
                 p % P, (int)hinfo[p]);
    factored = true;
}

// The solve's fixed sequence on the internal plane-ordered vectors (bb -> xx):
//   forward:  yy = blockdiag(LT) bb;  yy_p -= Mf_p yy_{p-1}  (p = 1 .. P-1, in order)
//   backward: xx = blockdiag(UT) yy;  xx_p -= Nb_p xx_{p+1}  (p = P-2 .. 0, in order)
void CoarseSolver::enqueue_core(hipStream_t s) const {
    const char *id = "tg_solver:coarse";
    const unsigned gr = grid_for(ntot, 4);
    k_tri_rows<<<gr, 256, 0, s>>>(LT, d_adOff, d_off, rowPlane, ntot, 0, bb, yy);
    for (int p = 1; p < P; ++p) {
        if (nd[p] == 0 || nd[p - 1] == 0) continue;
        launch_chain(Mf.p + mOff[p], nd[p], nd[p - 1], yy.p + off[p], yy.p + off[p - 1], yy.p + off[p], s);
    }
    k_tri_rows<<<gr, 256, 0, s>>>(Linv, d_adOff, d_off, rowPlane, ntot, 1, yy, xx);
    for (int p = P - 2; p >= 0; --p) {
        if (nd[p] == 0 || nd[p + 1] == 0) continue;
        launch_chain(Hm.p + hOff[p], nd[p], nd[p + 1], xx.p + off[p], xx.p + off[p + 1], xx.p + off[p], s);
    }
    TG_LAUNCH_CHECK(id);
}

template <typename T>
void CoarseSolver::solve(const T *r, T *x, const uint8_t *nodeMask, hipStream_t st) {
    const char *id = "tg_solver:coarse";
    if (!factored) fail(TG_ERR_STATE, id, "coarsest level not factorised (call assemble first)");
    k_pack_planes<T><<<grid_for(g.nodes(), 256), 256, 0, st>>>(r, bb, nodeMask, lidx, d_off, g, axis);
    TG_LAUNCH_CHECK(id);
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
    k_unpack_planes<T><<<grid_for(g.nodes(), 256), 256, 0, st>>>(xx, x, nodeMask, lidx, d_off, g, axis);
    TG_LAUNCH_CHECK(id);
}

template void CoarseSolver::solve<double>(const double *, double *, const uint8_t *, hipStream_t);
template void CoarseSolver::solve<float>(const float *, float *, const uint8_t *, hipStream_t);

}  // namespace tg
