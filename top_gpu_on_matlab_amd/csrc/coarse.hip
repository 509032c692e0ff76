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

// W (row-major, n x (n + m)) = [A | -B] with A (n x n) read transposed or not, B (n x m).
//   transposed == 0: A(r, c) = Lc[r + n c],       B(r, c') = Bc[r + n c']        (B col-major n x m)
//   transposed == 1: A(r, c) = Lc[c + n r],       B(r, c') = Bc[c' + m r]        (B = H^T, H col-major m x n)
__global__ void k_pack_op(const double *__restrict__ Lc, const double *__restrict__ Bc, double *__restrict__ W,
                          int64_t n, int64_t m, int transposed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = n + m;
    if (i >= n * w) return;
    const int64_t r = i / w, c = i % w;
    double v;
    if (c < n) v = transposed ? Lc[c + n * r] : Lc[r + n * c];
    else {
        const int64_t cc = c - n;
        v = -(transposed ? Bc[cc + m * r] : Bc[r + n * cc]);
    }
    W[i] = v;
}

// out[r] = sum_c W[r, c] * in(c), in = [in1 (n1) ; in2 (n2)], one wavefront per row.  Eight
// independent loads per lane are in flight per step (eight partial sums, combined in a fixed order).
__global__ __launch_bounds__(256) void k_gemv_rows(const double *__restrict__ W, int64_t rows, int64_t n1, int64_t n2,
                                                   const double *__restrict__ in1, const double *__restrict__ in2,
                                                   double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int64_t w = n1 + n2;
    const double *row = W + r * w;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t c = lane;
    for (; c + 7 * 64 < w; c += 8 * 64) {
        double a[8], x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t cc = c + 64 * k;
            a[k] = row[cc];
            x[k] = cc < n1 ? in1[cc] : in2[cc - n1];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += a[k] * x[k];
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const int64_t cc = c + 64 * k;
        if (cc < w) s[k] += row[cc] * (cc < n1 ? in1[cc] : in2[cc - n1]);
    }
    double t = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
    if (lane == 0) out[r] = t;
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
    wOff.assign(P + 1, 0);
    vOff.assign(P + 1, 0);
    for (int p = 0; p < P; ++p) {
        const int64_t prev = p > 0 ? nd[p - 1] : 0, next = p + 1 < P ? nd[p + 1] : 0;
        adOff[p + 1] = adOff[p] + nd[p] * nd[p];
        asOff[p + 1] = asOff[p] + nd[p] * prev;
        wOff[p + 1] = wOff[p] + nd[p] * (nd[p] + prev);
        vOff[p + 1] = vOff[p] + nd[p] * (nd[p] + next);
    }
    lidx.upload(lid, st);
    d_off.upload(off, st);
    d_adOff.upload(adOff, st);
    d_asOff.upload(asOff, st);
    Ad.alloc(adOff[P]);
    Linv.alloc(adOff[P]);
    As.alloc(std::max<int64_t>(asOff[P], 1));
    Fm.alloc(std::max<int64_t>(asOff[P], 1));
    Hm.alloc(std::max<int64_t>(asOff[P], 1));
    W.alloc(wOff[P]);
    V.alloc(vOff[P]);
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
    for (int p = 0; p < P; ++p) {
        const rocblas_int np = (rocblas_int)nd[p];
        if (np == 0) continue;
        const int64_t prev = p > 0 ? nd[p - 1] : 0, next = p + 1 < P ? nd[p + 1] : 0;
        if (prev > 0)
            rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, np, (rocblas_int)prev, np, &one,
                             Linv.p + adOff[p], np, As.p + asOff[p], np, &zero, Fm.p + asOff[p], np), "rocblas_dgemm F");
        if (next > 0)
            rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)next, np, np, &one,
                             As.p + asOff[p + 1], (rocblas_int)next, Linv.p + adOff[p], np, &zero,
                             Hm.p + asOff[p + 1], (rocblas_int)next), "rocblas_dgemm H");
        k_pack_op<<<grid_for(np * (np + prev), 256), 256, 0, st>>>(Linv.p + adOff[p], Fm.p + asOff[p],
                                                                   W.p + wOff[p], np, prev, 0);
        TG_LAUNCH_CHECK(id);
        k_pack_op<<<grid_for(np * (np + next), 256), 256, 0, st>>>(Linv.p + adOff[p],
                                                                   next > 0 ? Hm.p + asOff[p + 1] : Hm.p,
                                                                   V.p + vOff[p], np, next, 1);
        TG_LAUNCH_CHECK(id);
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

template <typename T>
void CoarseSolver::solve(const T *r, T *x, const uint8_t *nodeMask, hipStream_t st) {
    const char *id = "tg_solver:coarse";
    if (!factored) fail(TG_ERR_STATE, id, "coarsest level not factorised (call assemble first)");
    k_pack_planes<T><<<grid_for(g.nodes(), 256), 256, 0, st>>>(r, bb, nodeMask, lidx, d_off, g, axis);
    TG_LAUNCH_CHECK(id);
    for (int p = 0; p < P; ++p) {
        const int64_t prev = p > 0 ? nd[p - 1] : 0;
        if (nd[p] == 0) continue;
        k_gemv_rows<<<grid_for(nd[p], 4), 256, 0, st>>>(W.p + wOff[p], nd[p], nd[p], prev, bb.p + off[p],
                                                         prev ? yy.p + off[p - 1] : yy.p, yy.p + off[p]);
    }
    for (int p = P - 1; p >= 0; --p) {
        const int64_t next = p + 1 < P ? nd[p + 1] : 0;
        if (nd[p] == 0) continue;
        k_gemv_rows<<<grid_for(nd[p], 4), 256, 0, st>>>(V.p + vOff[p], nd[p], nd[p], next, yy.p + off[p],
                                                         next ? xx.p + off[p + 1] : xx.p, xx.p + off[p]);
    }
    TG_LAUNCH_CHECK(id);
    k_unpack_planes<T><<<grid_for(g.nodes(), 256), 256, 0, st>>>(xx, x, nodeMask, lidx, d_off, g, axis);
    TG_LAUNCH_CHECK(id);
}

template void CoarseSolver::solve<double>(const double *, double *, const uint8_t *, hipStream_t);
template void CoarseSolver::solve<float>(const float *, float *, const uint8_t *, hipStream_t);

}  // namespace tg
