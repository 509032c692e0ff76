// Device kernels of the structured-grid solver.
//
// Layout: every level keeps its vectors on the level's full padded grid
// (NY = resY+1, NX = resX+1, NZ = resZ+1 nodes, y fastest), DOF-interleaved: v[3n + c].
// Nodes outside the mesh hold 0.  A per-node mask byte carries bit c = "DOF c fixed"
// (finest level) and bit 3 = "node belongs to the mesh".
//
// All kernels are node-owned gathers: every output value is written by exactly one lane
// and summed in a fixed order, so results are bitwise reproducible.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kunit.h"

namespace tg {

constexpr uint8_t NODE_ACTIVE = 8;

// eNodMat vertex pattern (dy, dx, dz) and its inverse.
__host__ __device__ constexpr int vtx_y(int b) { return (b == 0 || b == 1 || b == 4 || b == 5) ? 1 : 0; }
__host__ __device__ constexpr int vtx_x(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__host__ __device__ constexpr int vtx_z(int b) { return b >= 4 ? 1 : 0; }
__host__ __device__ constexpr int vtx_of(int ty, int tx, int tz) { return 4 * tz + (ty ? (tx ? 1 : 0) : (tx ? 2 : 3)); }
// neighbour slot of relative node offset (dy, dx, dz) in {-1,0,1}^3
__host__ __device__ constexpr int nb_of(int dy, int dx, int dz) { return (dy + 1) + 3 * (dx + 1) + 9 * (dz + 1); }

struct Grid {
    int32_t NY, NX, NZ;  // nodes
    int32_t EY, EX, EZ;  // elements
    __host__ __device__ int64_t nodes() const { return (int64_t)NY * NX * NZ; }
    __host__ __device__ int64_t elems() const { return (int64_t)EY * EX * EZ; }
};

enum { EPI_KU = 0, EPI_RESID = 1, EPI_JACOBI = 2, EPI_RESID_JAC0 = 3 };

// Load the 27 neighbour nodes' values (fixed DOFs zeroed when MASKED).
template <typename T, bool MASKED>
__device__ inline void load_neighbours(T un[27][3], const T *__restrict__ u, const uint8_t *__restrict__ mask,
                                       const Grid g, int iy, int ix, int iz, int64_t n) {
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
                const int k = nb_of(dy, dx, dz);
                const bool ok = (unsigned)(iy + dy) < (unsigned)g.NY && (unsigned)(ix + dx) < (unsigned)g.NX &&
                                (unsigned)(iz + dz) < (unsigned)g.NZ;
                const int64_t m = n + dy + (int64_t)g.NY * (dx + (int64_t)g.NX * dz);
                T v0 = 0, v1 = 0, v2 = 0;
                if (ok) {
                    v0 = u[3 * m];
                    v1 = u[3 * m + 1];
                    v2 = u[3 * m + 2];
                    if (MASKED) {
                        const uint8_t mm = mask[m];
                        if (mm & 1) v0 = 0;
                        if (mm & 2) v1 = 0;
                        if (mm & 4) v2 = 0;
                    }
                }
                un[k][0] = v0;
                un[k][1] = v1;
                un[k][2] = v2;
            }
}

template <typename T, int EPI>
__device__ inline void epilogue(T *__restrict__ out, int64_t n, const T acc[3], const T *__restrict__ r,
                                const T *__restrict__ diag, const T *__restrict__ xold, T omega) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T v = acc[c];
        if (EPI == EPI_RESID) v = r[3 * n + c] - v;
        if (EPI == EPI_JACOBI) v = xold[3 * n + c] + (omega * (r[3 * n + c] - v)) / diag[3 * n + c];
        out[3 * n + c] = v;
    }
}

// Workgroup barrier that orders LDS accesses only: outstanding global loads stay in flight across
// it (__syncthreads() also waits for them), which lets a z-march prefetch the next plane.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Finest level, standard voxels: y = sum_e E_e * K0 * u_e with fixed DOFs as identity rows and
// zero columns (the BC-modified element matrices of the reference, TOP_GPU.m:704-720, 1729-1741).
// CK: K0 = cs * KUNIT_NU03 with the unit table folded into compile-time constants (no loads);
// otherwise K0 is read from memory (column-major 24x24, cs = 1).  Void and out-of-grid elements
// carry E = 0 and contribute exactly zero, so the loop over the 8 elements is branch-free.
template <typename T, int EPI, bool CK>
__global__ __launch_bounds__(256) void k_kfine(const T *__restrict__ u, T *__restrict__ out,
                                               const T *__restrict__ E, const uint8_t *__restrict__ mask,
                                               const T *__restrict__ K0, T cs, const Grid g, const T *__restrict__ r,
                                               const T *__restrict__ diag, T omega) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes()) return;
    const int iy = (int)(n % g.NY);
    const int64_t t = n / g.NY;
    const int ix = (int)(t % g.NX), iz = (int)(t / g.NX);
    const uint8_t mk = mask[n];
    if (!(mk & NODE_ACTIVE)) {
        out[3 * n] = 0;
        out[3 * n + 1] = 0;
        out[3 * n + 2] = 0;
        return;
    }
    T un[27][3];
    load_neighbours<T, true>(un, u, mask, g, iy, ix, iz, n);
    T acc[3] = {0, 0, 0};
#pragma unroll
    for (int oz = -1; oz <= 0; ++oz)
#pragma unroll
        for (int ox = -1; ox <= 0; ++ox)
#pragma unroll
            for (int oy = -1; oy <= 0; ++oy) {
                const int ey = iy + oy, ex = ix + ox, ez = iz + oz;
                const bool ok = (unsigned)ey < (unsigned)g.EY && (unsigned)ex < (unsigned)g.EX &&
                                (unsigned)ez < (unsigned)g.EZ;
                const T Ee = ok ? E[ey + (int64_t)g.EY * (ex + (int64_t)g.EX * ez)] : (T)0;
                const int a = vtx_of(-oy, -ox, -oz);
                T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const int k = nb_of(oy + vtx_y(b), ox + vtx_x(b), oz + vtx_z(b));
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const T uv = un[k][d];
                        const int i0 = (3 * b + d) + 24 * (3 * a + 0), i1 = i0 + 24, i2 = i0 + 48;
                        if (CK) {
                            s0 += uv * (T)KUNIT_NU03.v[i0];
                            s1 += uv * (T)KUNIT_NU03.v[i1];
                            s2 += uv * (T)KUNIT_NU03.v[i2];
                        } else {
                            s0 += uv * K0[i0];
                            s1 += uv * K0[i1];
                            s2 += uv * K0[i2];
                        }
                    }
                }
                acc[0] += s0 * Ee;
                acc[1] += s1 * Ee;
                acc[2] += s2 * Ee;
            }
    if (CK) {
        acc[0] *= cs;
        acc[1] *= cs;
        acc[2] *= cs;
    }
    if (mk & 1) acc[0] = u[3 * n];
    if (mk & 2) acc[1] = u[3 * n + 1];
    if (mk & 4) acc[2] = u[3 * n + 2];
    epilogue<T, EPI>(out, n, acc, r, diag, u, omega);
}

// Per-element matrices in entry-major ("SoA") layout: Km[i * ldk + e], i = r + 24c, for the coarse
// levels (Galerkin Ks, MASKED = false) and the super-element finest level (K_mat = I_mat * E_sub
// with BC masks, MASKED = true); y_e = u_e^T * K_e (TOP_GPU.m:757-760).
// Two stages.  Stage 1: y_e = u_e^T K_e for every solid
// element, PARTS threads per element (24 / PARTS output columns each; consecutive threads take
// consecutive elements, so every matrix entry row is streamed coalesced), written entry-major
// (f[j * ldk + e]).  Stage 2: each node sums its (up to 8) element contributions in the fixed
// element order (z, x, y), then applies identity rows for fixed DOFs (MASKED) and the epilogue.
template <typename T, bool MASKED, int PARTS>
__global__ __launch_bounds__(256) void k_kelem_vec(const T *__restrict__ u, T *__restrict__ f,
                                                   const T *__restrict__ Km, int64_t ldk,
                                                   const uint8_t *__restrict__ eleMask,
                                                   const uint8_t *__restrict__ mask, const Grid g) {
    constexpr int KE_PARTS = PARTS;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int part = (int)(gid / ldk);
    const int64_t e = gid - (int64_t)part * ldk;
    if (part >= KE_PARTS || e >= g.elems() || !eleMask[e]) return;
    const int ey = (int)(e % g.EY);
    const int64_t t = e / g.EY;
    const int ex = (int)(t % g.EX), ez = (int)(t / g.EX);
    T ue[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int64_t n = (ey + vtx_y(b)) + (int64_t)g.NY * ((ex + vtx_x(b)) + (int64_t)g.NX * (ez + vtx_z(b)));
        const uint8_t mm = MASKED ? mask[n] : (uint8_t)0;
#pragma unroll
        for (int d = 0; d < 3; ++d) ue[3 * b + d] = (mm & (1 << d)) ? (T)0 : u[3 * n + d];
    }
    const T *K = Km + e;
#pragma unroll
    for (int jj = 0; jj < 24 / KE_PARTS; ++jj) {
        const int j = part * (24 / KE_PARTS) + jj;
        T s = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i) s += ue[i] * K[(int64_t)(i + 24 * j) * ldk];
        f[(int64_t)j * ldk + e] = s;
    }
}

// The FP32 V-cycle's copy of the per-element matrices, packed symmetric: entry (i, j), i <= j, of K_e
// at Kp[(i + j (j + 1) / 2) * ldk + e], 300 of the 576 entries.  The element matrices are symmetric
// (Galerkin products P' K P of symmetric BC-modified matrices, TOP_GPU.m:1729-1741; I_mat * E_sub on
// the super-element level), so the upper triangle stands for both.  One thread per element reads
// each stored entry once and adds it into both outputs it touches; outputs as in k_kelem_vec.
constexpr int KSYM_N = 300;
__host__ __device__ constexpr int ksym_index(int i, int j) { return i + j * (j + 1) / 2; }  // i <= j

template <typename T, bool MASKED>
__global__ __launch_bounds__(256) void k_kelem_sym(const T *__restrict__ u, T *__restrict__ f,
                                                   const T *__restrict__ Kp, int64_t ldk,
                                                   const uint8_t *__restrict__ eleMask,
                                                   const uint8_t *__restrict__ mask, const Grid g) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.elems() || !eleMask[e]) return;
    const int ey = (int)(e % g.EY);
    const int64_t t = e / g.EY;
    const int ex = (int)(t % g.EX), ez = (int)(t / g.EX);
    T ue[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int64_t n = (ey + vtx_y(b)) + (int64_t)g.NY * ((ex + vtx_x(b)) + (int64_t)g.NX * (ez + vtx_z(b)));
        const uint8_t mm = MASKED ? mask[n] : (uint8_t)0;
#pragma unroll
        for (int d = 0; d < 3; ++d) ue[3 * b + d] = (mm & (1 << d)) ? (T)0 : u[3 * n + d];
    }
    const T *K = Kp + e;
    T acc[24];
#pragma unroll
    for (int j = 0; j < 24; ++j) acc[j] = 0;
#pragma unroll
    for (int j = 0; j < 24; ++j)
#pragma unroll
        for (int i = 0; i <= j; ++i) {
            const T k = K[(int64_t)ksym_index(i, j) * ldk];
            acc[j] += k * ue[i];
            if (i != j) acc[i] += k * ue[j];
        }
#pragma unroll
    for (int j = 0; j < 24; ++j) f[(int64_t)j * ldk + e] = acc[j];
}

// Kp (packed upper triangle, T) from K (all 576 entries, column-major per element, S), entry-major.
template <typename S, typename T>
__global__ void k_pack_sym(const S *__restrict__ K, T *__restrict__ Kp, int64_t ldk) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= KSYM_N * ldk) return;
    const int p = (int)(idx / ldk);
    const int64_t e = idx - (int64_t)p * ldk;
    int j = 0;
    while (ksym_index(0, j + 1) <= p) ++j;
    const int i = p - ksym_index(0, j);
    Kp[idx] = (T)K[(int64_t)(i + 24 * j) * ldk + e];
}

// The coarsest level's right-hand side in plane order, written by the gather that forms it (when set):
// active node n of plane p -> b[off[p] + 3 lidx[n] + c] in FP64, and the persistent solve's recurrence
// vectors e1, e2 set to the "not written" sentinel at the same slots (every slot is an active node's).
struct PlanePack {
    double *b = nullptr, *e1 = nullptr, *e2 = nullptr;
    const int32_t *lidx = nullptr;
    const int64_t *off = nullptr;
    int axis = 0;
    unsigned long long empty = 0;
};

template <typename T, int EPI, bool MASKED>
__global__ __launch_bounds__(256) void k_kelem_gather(const T *__restrict__ f, int64_t ldk, T *__restrict__ out,
                                                      const T *__restrict__ u, const uint8_t *__restrict__ eleMask,
                                                      const uint8_t *__restrict__ mask, const Grid g,
                                                      const T *__restrict__ r, const T *__restrict__ diag, T omega,
                                                      const PlanePack pk = PlanePack{}) {
    // one thread per node of node plane blockIdx.y; in-plane and element indices in 32 bits (the host
    // checks that they fit), neighbour elements by added strides
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int nYX = g.NY * g.NX;
    if (p >= nYX) return;
    const int iz = (int)blockIdx.y;
    const int ix = p / g.NY, iy = p - ix * g.NY;
    const int64_t n = p + (int64_t)nYX * iz;
    const uint8_t mk = mask[n];
    if (!(mk & NODE_ACTIVE)) {
        out[3 * n] = 0;
        out[3 * n + 1] = 0;
        out[3 * n + 2] = 0;
        return;
    }
    T acc[3] = {0, 0, 0};
    const int eYX = g.EY * g.EX;
    const int e00 = iy + g.EY * ix + eYX * iz;  // element (iy, ix, iz) (meaningful only where in the grid)
#pragma unroll
    for (int oz = -1; oz <= 0; ++oz)
#pragma unroll
        for (int ox = -1; ox <= 0; ++ox)
#pragma unroll
            for (int oy = -1; oy <= 0; ++oy) {
                const int ey = iy + oy, ex = ix + ox, ez = iz + oz;
                const bool ok = (unsigned)ey < (unsigned)g.EY && (unsigned)ex < (unsigned)g.EX &&
                                (unsigned)ez < (unsigned)g.EZ;
                const int e = e00 + oy + (ox ? -g.EY : 0) + (oz ? -eYX : 0);
                if (ok && eleMask[e]) {
                    const int a = vtx_of(-oy, -ox, -oz);
                    acc[0] += f[(int64_t)(3 * a) * ldk + e];
                    acc[1] += f[(int64_t)(3 * a + 1) * ldk + e];
                    acc[2] += f[(int64_t)(3 * a + 2) * ldk + e];
                }
            }
    if (MASKED) {
        if (mk & 1) acc[0] = u[3 * n];
        if (mk & 2) acc[1] = u[3 * n + 1];
        if (mk & 4) acc[2] = u[3 * n + 2];
    }
    epilogue<T, EPI>(out, n, acc, r, diag, u, omega);
    if (EPI == EPI_KU && pk.b) {
        const int p = pk.axis == 0 ? iy : (pk.axis == 1 ? ix : iz);
        const int64_t base = pk.off[p] + 3 * (int64_t)pk.lidx[n];
        const double empty = __longlong_as_double((long long)pk.empty);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            pk.b[base + c] = (double)acc[c];  // the value stored above (EPI_KU)
            pk.e1[base + c] = empty;
            pk.e2[base + c] = empty;
        }
    }
}

// Trilinear weight of coarse vertex j at fine offset (a, b, c) of a span-S block; exact in binary.
template <typename T, int S>
__device__ inline T ri_w(int a, int b, int c, int j) {
    const T ty = (T)a / S, tx = (T)b / S, tz = (T)c / S;
    const T wy = vtx_y(j) ? ty : (T)1 - ty;
    const T wx = vtx_x(j) ? tx : (T)1 - tx;
    const T wz = vtx_z(j) ? tz : (T)1 - tz;
    return wy * wx * wz;
}

// Workgroups go to the 8 XCDs of the MI355X round-robin by blockIdx.x (workgroup 8 k + x is the k-th
// one XCD x runs).  This renumbering gives XCD x a contiguous range of [0, n), so neighbouring tiles,
// which share halo data, meet in the same L2.  A bijection of [0, n) for any n.
constexpr int N_XCDS = 8;
__device__ __forceinline__ int xcd_contiguous(int b, int n) {
    const int q = n / N_XCDS, r = n % N_XCDS, x = b % N_XCDS, k = b / N_XCDS;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Restriction r_c = R r_f (Solving_RestrictResidual, TOP_GPU.m:817-847): for every coarse node J, the
// sum over the non-empty coarse elements around J of RI(:,J)' * (r_f ./ coeff) on the element's
// (S+1)^3 fine nodes; coeff is transferMatCoeffi per fine grid node, a count of at most 8 non-empty
// coarse elements (1 where unused), stored as one byte and converted exactly to T.
// Element-centric and LDS-tiled, the trilinear weights applied in separable order.  A
// workgroup owns RT_E x RT_E coarse elements (y, x) over RT_Z coarse layers and marches through the
// fine node planes of that box: each plane of v = r_f ./ coeff is staged in LDS (the next one is
// prefetched into registers), every lane forms, for its (element, vertex column (jy, jx), component),
// the 2D trilinear sum over the element's (S+1)^2 fine nodes of the plane, and folds it into the
// vertex's two z weights.  When a coarse layer is complete the lane writes its part of the element's
// 8-vertex contributions C_e (entry-major f[(3 j + c) * ldk + e]); k_kelem_gather then sums them per
// coarse node over the non-empty elements, which is r_c = P^T r_f.
constexpr int RT_E = 8, RT_Z = 2;
constexpr int RT_THREADS = RT_E * RT_E * 12;  // one lane per (element, vertex column, component)
template <typename T, int S>
__global__ __launch_bounds__(RT_THREADS) void k_restrict_tiled(const T *__restrict__ rf, const uint8_t *__restrict__ coeff,
                                                               const Grid gc, const Grid gf, T *__restrict__ f,
                                                               int64_t ldk) {
    constexpr int FN = RT_E * S + 1;                      // fine nodes per tile side
    constexpr int NV = 3 * FN * FN;                       // staged values per plane
    constexpr int NI = (NV + RT_THREADS - 1) / RT_THREADS;  // staged values per lane
    __shared__ T sV[NV];
    const int t = threadIdx.x;
    const int tilesY = (gc.EY + RT_E - 1) / RT_E, tilesX = (gc.EX + RT_E - 1) / RT_E;
    int bid = xcd_contiguous((int)blockIdx.x, (int)gridDim.x);
    const int tz = bid / (tilesY * tilesX);
    bid -= tz * tilesY * tilesX;
    const int e0y = (bid % tilesY) * RT_E, e0x = (bid / tilesY) * RT_E, e0z = tz * RT_Z;
    const int ez1 = min(e0z + RT_Z, gc.EZ);
    const int64_t NYX = (int64_t)gf.NY * gf.NX;
    // this lane's output: component, vertex column (jy, jx), element (ly, lx) of the tile
    const int oc = t % 3, q = t / 3, jy = q & 1, jx = (q >> 1) & 1, el = q >> 2;
    const int ly = el % RT_E, lx = el / RT_E;
    const int ey = e0y + ly, ex = e0x + lx;
    const bool own = ey < gc.EY && ex < gc.EX;
    // the next plane's raw values (the division waits for them, so it is done at the LDS store, a
    // plane later, and the loads stay in flight meanwhile); out-of-grid slots read node 0 and store 0
    T preR[NI];
    uint8_t preC[NI];
    bool preOk[NI];
    auto fetch = [&](int iz) {
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int i = min(t + RT_THREADS * k, NV - 1);
            const int node = i / 3, c = i - 3 * node;
            const int gy = S * e0y + node % FN, gx = S * e0x + node / FN;
            const bool ok = gy < gf.NY && gx < gf.NX && iz < gf.NZ;
            const int64_t n = ok ? gy + (int64_t)gf.NY * gx + NYX * iz : 0;
            preR[k] = rf[3 * n + c];
            preC[k] = coeff[n];
            preOk[k] = ok;
        }
    };
    const T *pv = sV + oc * FN * FN + S * ly + FN * S * lx;
    T acc0 = 0, acc1 = 0;
    const int iz0 = S * e0z, iz1 = S * ez1;
    fetch(iz0);
    for (int iz = iz0; iz <= iz1; ++iz) {
        lds_barrier();
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int i = t + RT_THREADS * k;
            if (i < NV) sV[(i % 3) * FN * FN + i / 3] = preOk[k] ? preR[k] / (T)preC[k] : (T)0;
        }
        lds_barrier();
        if (iz < iz1) fetch(iz + 1);
        T val = 0;
#pragma unroll
        for (int b = 0; b <= S; ++b) {
            T sy = 0;
#pragma unroll
            for (int a = 0; a <= S; ++a) sy += (jy ? (T)a / S : (T)(S - a) / S) * pv[a + FN * b];
            val += (jx ? (T)b / S : (T)(S - b) / S) * sy;
        }
        const int cz = (iz - iz0) % S;  // offset inside the current coarse layer
        if (cz == 0 && iz > iz0) {
            // this plane closes the layer below (top face, z weight 1) and opens the next one
            acc1 += val;
            if (own) {
                const int64_t e = ey + (int64_t)gc.EY * (ex + (int64_t)gc.EX * (e0z + (iz - iz0) / S - 1));
                f[(int64_t)(3 * vtx_of(jy, jx, 0) + oc) * ldk + e] = acc0;
                f[(int64_t)(3 * vtx_of(jy, jx, 1) + oc) * ldk + e] = acc1;
            }
            acc0 = val;
            acc1 = 0;
        } else {
            const T tz = (T)cz / S;
            acc0 += ((T)1 - tz) * val;
            acc1 += tz * val;
        }
    }
}

// Prolongation x_f (+)= P x_c (Solving_InterpolationDeviation, TOP_GPU.m:848-878): the reference
// sums the trilinear interpolants of every non-empty coarse element containing a fine node and
// divides by their count (transferMatCoeffi).  The trilinear interpolant is continuous across
// element faces, so all those interpolants are the same number: each active fine node takes the
// interpolant of the first non-empty coarse element containing it (z, x, y order).
template <typename T, int S, bool ADD>
__global__ __launch_bounds__(256) void k_prolong(const T *__restrict__ xc, T *__restrict__ xf,
                                                 const uint8_t *__restrict__ fMask, const uint8_t *__restrict__ cEle,
                                                 const Grid gc, const Grid gf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= gf.nodes()) return;
    const int iy = (int)(i % gf.NY);
    const int64_t t = i / gf.NY;
    const int ix = (int)(t % gf.NX), iz = (int)(t / gf.NX);
    T a0 = 0, a1 = 0, a2 = 0;
    T o0 = 0, o1 = 0, o2 = 0;
    if (ADD) {  // issued first: independent of the element search
        o0 = xf[3 * i];
        o1 = xf[3 * i + 1];
        o2 = xf[3 * i + 2];
    }
    if (fMask[i] & NODE_ACTIVE) {
        // coarse elements containing the node: e in [ceil(i/S) - 1, floor(i/S)], clamped to the grid;
        // the (up to 8) solidity flags are loaded together, then the first solid one (z, x, y order)
        const int qy = min(iy / S, gc.EY - 1), qx = min(ix / S, gc.EX - 1), qz = min(iz / S, gc.EZ - 1);
        const int ny0 = max((iy + S - 1) / S - 1, 0), nx0 = max((ix + S - 1) / S - 1, 0);
        const int nz0 = max((iz + S - 1) / S - 1, 0);
        uint8_t fl[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int cy = (k & 1) ? qy : ny0, cx = (k & 2) ? qx : nx0, cz = (k & 4) ? qz : nz0;
            fl[k] = cEle[cy + (int64_t)gc.EY * (cx + (int64_t)gc.EX * cz)];
        }
        int ey = -1, ex = -1, ez = -1;
#pragma unroll
        for (int k = 7; k >= 0; --k) {  // descending, so the last assignment is the first solid candidate
            const int cy = (k & 1) ? qy : ny0, cx = (k & 2) ? qx : nx0, cz = (k & 4) ? qz : nz0;
            if (fl[k]) {
                ey = cy;
                ex = cx;
                ez = cz;
            }
        }
        if (ez >= 0) {
            const int a = iy - S * ey, b = ix - S * ex, c = iz - S * ez;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int64_t Jn = (int64_t)(ey + vtx_y(j)) +
                                   (int64_t)gc.NY * ((ex + vtx_x(j)) + (int64_t)gc.NX * (ez + vtx_z(j)));
                const T w = ri_w<T, S>(a, b, c, j);
                a0 += w * xc[3 * Jn];
                a1 += w * xc[3 * Jn + 1];
                a2 += w * xc[3 * Jn + 2];
            }
        }
    }
    if (ADD) {
        xf[3 * i] = o0 + a0;
        xf[3 * i + 1] = o1 + a1;
        xf[3 * i + 2] = o2 + a2;
    } else {
        xf[3 * i] = a0;
        xf[3 * i + 1] = a1;
        xf[3 * i + 2] = a2;
    }
}

// x = w * (r ./ d)  (pre-smoothing, TOP_GPU.m:784)
template <typename T>
__global__ void k_jacobi0(const T *__restrict__ r, const T *__restrict__ d, T *__restrict__ x, T w, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = w * (r[i] / d[i]);
}

// out = x + (w*r) ./ d  ('Adapted' post-smoothing, TOP_GPU.m:807-809)
template <typename T>
__global__ void k_jacobi_adapted(const T *__restrict__ x, const T *__restrict__ r, const T *__restrict__ d,
                                 T *__restrict__ out, T w, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = x[i] + (w * r[i]) / d[i];
}

// ---- conversions / packing ----
template <typename A, typename B>
__global__ void k_convert(const A *__restrict__ a, B *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (B)a[i];
}

// compressed interleaved (reference order) -> full grid; nodFull0 = 0-based full index per solid node
template <typename A, typename B>
__global__ void k_unpack(const A *__restrict__ c, B *__restrict__ f, const int32_t *__restrict__ nodFull0, int64_t Nn) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Nn) return;
    const int64_t n = nodFull0[k];
    f[3 * n] = (B)c[3 * k];
    f[3 * n + 1] = (B)c[3 * k + 1];
    f[3 * n + 2] = (B)c[3 * k + 2];
}

template <typename A, typename B>
__global__ void k_pack(const A *__restrict__ f, B *__restrict__ c, const int32_t *__restrict__ nodFull0, int64_t Nn) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Nn) return;
    const int64_t n = nodFull0[k];
    c[3 * k] = (B)f[3 * n];
    c[3 * k + 1] = (B)f[3 * n + 1];
    c[3 * k + 2] = (B)f[3 * n + 2];
}

// ---- deterministic reductions (fixed grid, fixed order) ----
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

template <typename T>
__device__ inline double block_sum(double v) {
    __shared__ double sh[RED_THREADS];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// partial[b] = sum over this block's fixed contiguous chunk of a[i]*b[i]
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_dot_partial(const T *__restrict__ a, const T *__restrict__ b,
                                                             int64_t n, double *__restrict__ partial) {
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s += (double)a[i] * (double)b[i];
    s = block_sum<T>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[slot] = sum(partial[0..nb))
// host (optional): a device-mapped pinned host copy of out, written too, so the host reads the sum
// without a copy; seq_ptr (optional): then seq is published there (a system-scope release store, so a
// host that sees seq also sees the sum)
static __global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double *__restrict__ partial, int nb,
                                                              double *__restrict__ out, int slot,
                                                              double *__restrict__ host = nullptr,
                                                              unsigned long long *seq_ptr = nullptr,
                                                              unsigned long long seq = 0) {
    double s = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
    s = block_sum<double>(s);
    if (threadIdx.x == 0) {
        out[slot] = s;
        if (host) host[slot] = s;
        if (seq_ptr) __hip_atomic_store(seq_ptr, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace tg
