// Objective, sensitivities, density filters and the optimality-criteria update on the device.
//
//   unit compliance     TopOpti_ComputeUnitCompliance(_Super)   TOP_GPU.m:580-600, 2622-2645
//   sub-voxel energies  pre_compute_der                         TOP_GPU.m:2646-2671
//   SIMP / dc           TOP_GPU.m:168-174, 189-209
//   OC update           TOP_GPU.m:233-253 (bisection on the device, deterministic reductions)
//   PDE filter          TopOpti_SetupPDEfilter_matrixFree / ConductPDEFiltering / MatTimesVec_matrixFree_B
//                       TOP_GPU.m:880-1031 (nodal Helmholtz solve with Jacobi-PCG)
//   convolution filter  imfilter density filter of super-element mode, TOP_GPU.m:112-121, 220-227, 262-265
#include <cmath>

#include "kernels.h"
#include "model.h"
#include "pcg.h"
#include "solver.h"
#include "topopt.h"

namespace tg {
namespace {

// ce_e = u_e' * K * u_e, K = Ke (MODE 0) or the element's K_mat (MODE 1, entry-major).
// u is a full-grid interleaved vector; output in solid-element order.
template <int MODE>
__global__ void k_unit_compliance(const double *__restrict__ u, const double *__restrict__ K0,
                                  const double *__restrict__ Km, int64_t ldk, const int32_t *__restrict__ eleFull0,
                                  int64_t Ne, const Grid g, double *__restrict__ ce) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Ne) return;
    const int64_t e = eleFull0[k];
    const int ey = (int)(e % g.EY), ex = (int)((e / g.EY) % g.EX), ez = (int)(e / ((int64_t)g.EY * g.EX));
    double ue[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int64_t m = (ey + vtx_y(b)) + (int64_t)g.NY * ((ex + vtx_x(b)) + (int64_t)g.NX * (ez + vtx_z(b)));
        ue[3 * b] = u[3 * m];
        ue[3 * b + 1] = u[3 * m + 1];
        ue[3 * b + 2] = u[3 * m + 2];
    }
    double s = 0;
    for (int j = 0; j < 24; ++j) {
        double kj = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i) {
            if (MODE == 0) kj += ue[i] * K0[i + 24 * j];       // (u_e' * Ke)_j
            else kj += Km[(int64_t)(j + 24 * i) * ldk + e] * ue[i];  // (K_mat_e * u_e)_j
        }
        s += kj * ue[j];
    }
    ce[k] = s;
}

// dc(q, e) = u_e' * I_q * u_e
__global__ void k_subvoxel_energy(const double *__restrict__ u, const double *__restrict__ Imat,
                                  const int32_t *__restrict__ eleFull0, int64_t Ne, const Grid g,
                                  double *__restrict__ dc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Ne * 64) return;
    const int64_t k = t / 64;
    const int q = (int)(t % 64);
    const int64_t e = eleFull0[k];
    const int ey = (int)(e % g.EY), ex = (int)((e / g.EY) % g.EX), ez = (int)(e / ((int64_t)g.EY * g.EX));
    double ue[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int64_t m = (ey + vtx_y(b)) + (int64_t)g.NY * ((ex + vtx_x(b)) + (int64_t)g.NX * (ez + vtx_z(b)));
        ue[3 * b] = u[3 * m];
        ue[3 * b + 1] = u[3 * m + 1];
        ue[3 * b + 2] = u[3 * m + 2];
    }
    const double *I = Imat + (size_t)576 * q;
    double s = 0;
    for (int i = 0; i < 24; ++i) {
        double ki = 0;
        for (int j = 0; j < 24; ++j) ki += I[i + 24 * j] * ue[j];
        s += ki * ue[i];
    }
    dc[q + 64 * k] = s;
}

__global__ void k_simp(const double *__restrict__ x, double *__restrict__ E, int64_t n, double Emin, double E0,
                       double p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) E[i] = Emin + pow(x[i], p) * (E0 - Emin);
}

// dc = -(p (E0 - Emin) x^(p-1)) .* (ce / cSolid)
__global__ void k_sens(const double *__restrict__ x, const double *__restrict__ ce, double *__restrict__ dc,
                       int64_t n, int per, double Emin, double E0, double p, double cSolid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dc[i] = -(p * (E0 - Emin) * pow(x[i], p - 1)) * (ce[i / per] / cSolid);
}

__global__ void k_mul(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] * b[i];
}

__global__ void k_div_floor(const double *__restrict__ a, const double *__restrict__ x, double *__restrict__ c,
                            double fl, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] / fmax(fl, x[i]);
}

// c = (b != 0) ? v : a  (xPhys(passiveElements) = 1 after the density filter, TOP_GPU.m:270, 273)
__global__ void k_set_where(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c, double v,
                            int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = b[i] != 0.0 ? v : a[i];
}

// c = a + v b  (the load cases' weighted sum of unit compliances, ceList * objWeightingList_,
// TOP_GPU.m:599, 2644, 2671); c may be a (no restrict: each entry is read before it is written)
__global__ void k_axpy(const double *a, const double *__restrict__ b, double *c, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] + v * b[i];
}

// ---- reductions ------------------------------------------------------------------------------
enum { R_SUM = 0, R_DOT = 1, R_MAXABSDIFF = 2, R_SHARP = 3 };

template <int OP>
__global__ __launch_bounds__(RED_THREADS) void k_red_partial(const double *__restrict__ a, const double *__restrict__ b,
                                                             int64_t n, double *__restrict__ partial) {
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        if (OP == R_SUM) s += a[i];
        else if (OP == R_DOT) s += a[i] * b[i];
        else if (OP == R_SHARP) s += a[i] * (1.0 - a[i]);
        else s = fmax(s, fabs(a[i] - b[i]));
    }
    __shared__ double sh[RED_THREADS];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = OP == R_MAXABSDIFF ? fmax(sh[threadIdx.x], sh[threadIdx.x + w])
                                                                      : sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

template <int OP>
__global__ __launch_bounds__(RED_THREADS) void k_red_final(const double *__restrict__ partial, int nb,
                                                           double *__restrict__ out) {
    __shared__ double sh[RED_THREADS];
    double s = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s = OP == R_MAXABSDIFF ? fmax(s, partial[i]) : s + partial[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = OP == R_MAXABSDIFF ? fmax(sh[threadIdx.x], sh[threadIdx.x + w])
                                                                      : sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// ---- optimality criteria ---------------------------------------------------------------------
// scal: [0] l1, [1] l2, [2] lmid of the last executed step, [3] fval, [4] move, [5] n
__device__ inline bool oc_open(const double *scal) { return (scal[1] - scal[0]) / (scal[0] + scal[1]) > 1e-6; }

__device__ inline double oc_xnew(double x, double dc, double dv, double lmid, double move) {
    return fmax(0.0, fmax(x - move, fmin(1.0, fmin(x + move, x * sqrt(-dc / dv / lmid)))));
}

// passive (may be null): entries with a nonzero indicator take xnew = 1 inside every bisection step, so
// they enter gt (xnew(passiveElements) = 1.0, TOP_GPU.m:239, 247)
__device__ inline double oc_xnew_p(const double *__restrict__ x, const double *__restrict__ dc,
                                   const double *__restrict__ dv, const double *__restrict__ passive, int64_t i,
                                   double lmid, double move) {
    if (passive && passive[i] != 0.0) return 1.0;
    return oc_xnew(x[i], dc[i], dv[i], lmid, move);
}

__global__ __launch_bounds__(RED_THREADS) void k_oc_step(const double *__restrict__ x, const double *__restrict__ dc,
                                                         const double *__restrict__ dv,
                                                         const double *__restrict__ passive, int64_t n,
                                                         const double *__restrict__ scal, double *__restrict__ partial) {
    if (!oc_open(scal)) return;
    const double lmid = 0.5 * (scal[1] + scal[0]), move = scal[4];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s += oc_xnew_p(x, dc, dv, passive, i, lmid, move) - x[i];
    s = block_sum<double>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(RED_THREADS) void k_oc_bisect(const double *__restrict__ partial, int nb,
                                                           double *__restrict__ scal) {
    if (!oc_open(scal)) return;
    double s = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
    s = block_sum<double>(s);
    if (threadIdx.x == 0) {
        const double lmid = 0.5 * (scal[1] + scal[0]);
        const double gt = scal[3] + s / scal[5];
        if (gt > 0) scal[0] = lmid;
        else scal[1] = lmid;
        scal[2] = lmid;
    }
}

__global__ void k_oc_apply(const double *__restrict__ x, const double *__restrict__ dc, const double *__restrict__ dv,
                           const double *__restrict__ passive, double *__restrict__ xnew, int64_t n,
                           const double *__restrict__ scal) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xnew[i] = oc_xnew_p(x, dc, dv, passive, i, scal[2], scal[4]);
}

// ---- PDE filter ------------------------------------------------------------------------------
// element -> node: b_n = sum over solid elements around n of src_e * (1/8)  (TOP_GPU.m:975-996)
__global__ void k_e2n(const double *__restrict__ src, const int32_t *__restrict__ eleFwd, const uint8_t *__restrict__ mask,
                      const Grid g, double *__restrict__ b) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes()) return;
    double s = 0;
    if (mask[n] & NODE_ACTIVE) {
        const int iy = (int)(n % g.NY);
        const int64_t t = n / g.NY;
        const int ix = (int)(t % g.NX), iz = (int)(t / g.NX);
        for (int oz = -1; oz <= 0; ++oz)
            for (int ox = -1; ox <= 0; ++ox)
                for (int oy = -1; oy <= 0; ++oy) {
                    const int ey = iy + oy, ex = ix + ox, ez = iz + oz;
                    if ((unsigned)ey >= (unsigned)g.EY || (unsigned)ex >= (unsigned)g.EX || (unsigned)ez >= (unsigned)g.EZ)
                        continue;
                    const int32_t k = eleFwd[ey + (int64_t)g.EY * (ex + (int64_t)g.EX * ez)];
                    if (k > 0) s += src[k - 1] * (1.0 / 8);
                }
    }
    b[n] = s;
}

// y_n = sum over solid elements around n of (u_e' * iKEF)_a  (MatTimesVec_matrixFree_B, TOP_GPU.m:1008-1031)
template <int MODE>
__global__ void k_filter_apply(const double *__restrict__ u, double *__restrict__ y, const double *__restrict__ K,
                               const uint8_t *__restrict__ eleMask, const uint8_t *__restrict__ mask, const Grid g,
                               const double *__restrict__ inv) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes()) return;
    if (MODE == 1) {  // Jacobi preconditioner diagPrecond .* x
        y[n] = inv[n] * u[n];
        return;
    }
    double s = 0;
    if (mask[n] & NODE_ACTIVE) {
        const int iy = (int)(n % g.NY);
        const int64_t t = n / g.NY;
        const int ix = (int)(t % g.NX), iz = (int)(t / g.NX);
#pragma unroll
        for (int oz = -1; oz <= 0; ++oz)
#pragma unroll
            for (int ox = -1; ox <= 0; ++ox)
#pragma unroll
                for (int oy = -1; oy <= 0; ++oy) {
                    const int ey = iy + oy, ex = ix + ox, ez = iz + oz;
                    if ((unsigned)ey >= (unsigned)g.EY || (unsigned)ex >= (unsigned)g.EX || (unsigned)ez >= (unsigned)g.EZ)
                        continue;
                    if (!eleMask[ey + (int64_t)g.EY * (ex + (int64_t)g.EX * ez)]) continue;
                    const int a = vtx_of(-oy, -ox, -oz);
                    double v = 0;
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const int64_t m = (ey + vtx_y(b)) + (int64_t)g.NY * ((ex + vtx_x(b)) + (int64_t)g.NX * (ez + vtx_z(b)));
                        v += u[m] * K[b + 8 * a];
                    }
                    s += v;
                }
    }
    y[n] = s;
}

__global__ void k_filter_diag(const double *__restrict__ K, const uint8_t *__restrict__ eleMask,
                              const uint8_t *__restrict__ mask, const Grid g, double *__restrict__ inv) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes()) return;
    double s = 0;
    if (mask[n] & NODE_ACTIVE) {
        const int iy = (int)(n % g.NY);
        const int64_t t = n / g.NY;
        const int ix = (int)(t % g.NX), iz = (int)(t / g.NX);
        for (int oz = -1; oz <= 0; ++oz)
            for (int ox = -1; ox <= 0; ++ox)
                for (int oy = -1; oy <= 0; ++oy) {
                    const int ey = iy + oy, ex = ix + ox, ez = iz + oz;
                    if ((unsigned)ey >= (unsigned)g.EY || (unsigned)ex >= (unsigned)g.EX || (unsigned)ez >= (unsigned)g.EZ)
                        continue;
                    if (!eleMask[ey + (int64_t)g.EY * (ex + (int64_t)g.EX * ez)]) continue;
                    const int a = vtx_of(-oy, -ox, -oz);
                    s += K[a + 8 * a];
                }
    }
    inv[n] = (mask[n] & NODE_ACTIVE) ? 1.0 / s : 0.0;
}

// node -> element: tar_e = sum(tar(eNod(e,:))) * (1/8)  (TOP_GPU.m:1004-1005)
__global__ void k_n2e(const double *__restrict__ tar, const int32_t *__restrict__ eleFull0, int64_t Ne, const Grid g,
                      double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Ne) return;
    const int64_t e = eleFull0[k];
    const int ey = (int)(e % g.EY), ex = (int)((e / g.EY) % g.EX), ez = (int)(e / ((int64_t)g.EY * g.EX));
    double s = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b)
        s += tar[(ey + vtx_y(b)) + (int64_t)g.NY * ((ex + vtx_x(b)) + (int64_t)g.NX * (ez + vtx_z(b)))];
    out[k] = s * (1.0 / 8);
}

// ---- convolution filter ----------------------------------------------------------------------
// grid voxel (iy, ix, iz) of the num-times refined box <-> element-major index q + num^3 * element
__device__ inline int64_t cv_new_index(int iy, int ix, int iz, int num, int Y, int X) {
    const int64_t big = iy / num + (int64_t)Y * (ix / num + (int64_t)X * (iz / num));
    const int64_t loc = iy % num + num * (ix % num + num * (iz % num));
    return loc + (int64_t)num * num * num * big;
}

// out[p] = sum_q h(q) * in(p + q - c) on a zero-padded grid (imfilter, 'same', bcF = 0).
// MODE 0: in = grid(src)         ; out = result ./ Hs  (forward: xPhys = imfilter(x)./Hs)
// MODE 1: in = grid(src) ./ Hs   ; out = result       (adjoint: imfilter(dc./Hs))
// MODE 2: in = 1                 ; out = result       (Hs)
template <int MODE>
__global__ void k_conv(const double *__restrict__ src, const double *__restrict__ Hs, const double *__restrict__ h,
                       int hr, int GY, int GX, int GZ, int num, int Y, int X, double *__restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)GY * GX * GZ) return;
    const int iy = (int)(p % GY), ix = (int)((p / GY) % GX), iz = (int)(p / ((int64_t)GY * GX));
    const int w = 2 * hr + 1;
    double s = 0;
    for (int dz = -hr; dz <= hr; ++dz) {
        const int z = iz + dz;
        if ((unsigned)z >= (unsigned)GZ) continue;
        for (int dx = -hr; dx <= hr; ++dx) {
            const int x = ix + dx;
            if ((unsigned)x >= (unsigned)GX) continue;
            for (int dy = -hr; dy <= hr; ++dy) {
                const int y = iy + dy;
                if ((unsigned)y >= (unsigned)GY) continue;
                const double hv = h[(dy + hr) + w * ((dx + hr) + w * (dz + hr))];
                if (hv == 0.0) continue;
                double v;
                if (MODE == 2) v = 1.0;
                else {
                    v = src[cv_new_index(y, x, z, num, Y, X)];
                    if (MODE == 1) v = v / Hs[y + (int64_t)GY * (x + (int64_t)GX * z)];
                }
                s += hv * v;
            }
        }
    }
    if (MODE == 0) out[cv_new_index(iy, ix, iz, num, Y, X)] = s / Hs[p];
    else if (MODE == 1) out[cv_new_index(iy, ix, iz, num, Y, X)] = s;
    else out[p] = s;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
double reduce_dev(int op, const double *a, const double *b, int64_t n, DevBuf<double> &partial, DevBuf<double> &res,
                  double *h, hipStream_t st) {
    const char *id = "tg_reduce:Input";
    switch (op) {
        case R_SUM: k_red_partial<R_SUM><<<RED_BLOCKS, RED_THREADS, 0, st>>>(a, b, n, partial); k_red_final<R_SUM><<<1, RED_THREADS, 0, st>>>(partial, RED_BLOCKS, res); break;
        case R_DOT: k_red_partial<R_DOT><<<RED_BLOCKS, RED_THREADS, 0, st>>>(a, b, n, partial); k_red_final<R_DOT><<<1, RED_THREADS, 0, st>>>(partial, RED_BLOCKS, res); break;
        case R_MAXABSDIFF: k_red_partial<R_MAXABSDIFF><<<RED_BLOCKS, RED_THREADS, 0, st>>>(a, b, n, partial); k_red_final<R_MAXABSDIFF><<<1, RED_THREADS, 0, st>>>(partial, RED_BLOCKS, res); break;
        case R_SHARP: k_red_partial<R_SHARP><<<RED_BLOCKS, RED_THREADS, 0, st>>>(a, b, n, partial); k_red_final<R_SHARP><<<1, RED_THREADS, 0, st>>>(partial, RED_BLOCKS, res); break;
        default: fail(TG_ERR_ARGS, id, "unknown reduction %d", op);
    }
    TG_LAUNCH_CHECK(id);
    TG_HIP(hipMemcpyAsync(h, res.p, sizeof(double), hipMemcpyDeviceToHost, st), id);
    TG_HIP(hipStreamSynchronize(st), id);
    return *h;
}

Filter::Filter(const tg_model &mm, double rMin, hipStream_t s) : m(mm), st(s) {
    const char *id = "tg_filter_create:Input";
    const HostLevel &L = m.lv[0];
    g.EY = L.resY; g.EX = L.resX; g.EZ = L.resZ;
    g.NY = L.resY + 1; g.NX = L.resX + 1; g.NZ = L.resZ + 1;
    Ne = L.Ne;
    double K[64];
    pde_filter_kernel(rMin, 1.0, K);  // eleSize(1) = 1 on the finest level (TOP_GPU.m:2066)
    Kd.upload(K, 64, st);
    std::vector<uint8_t> mask(g.nodes(), 0);
    for (int64_t k = 0; k < L.Nn; ++k) mask[L.nodFull[k] - 1] = NODE_ACTIVE;
    nodeMask.upload(mask, st);
    eleMask.upload(L.solid, st);
    eleFwd.upload(L.eleMapForward, st);
    std::vector<int32_t> ef(L.Ne);
    for (int64_t k = 0; k < L.Ne; ++k) ef[k] = L.eleMapBack[k] - 1;
    eleFull0.upload(ef, st);
    inv.alloc(g.nodes());
    b.alloc(g.nodes());
    y.alloc(g.nodes());
    pw.init(g.nodes(), st);
    k_filter_diag<<<grid_for(g.nodes(), 256), 256, 0, st>>>(Kd, eleMask, nodeMask, g, inv);
    TG_LAUNCH_CHECK(id);
    TG_HIP(hipStreamSynchronize(st), id);
}

void Filter::apply(const double *src, double *tar, double tol, int maxIT, int *its, double *res) {
    const char *id = "tg_filter_apply:Numeric";
    const unsigned gn = grid_for(g.nodes(), 256);
    k_e2n<<<gn, 256, 0, st>>>(src, eleFwd, nodeMask, g, b);
    y.zero(st);
    TG_LAUNCH_CHECK(id);
    pw.run([&](const double *u, double *out, int) {
               k_filter_apply<0><<<gn, 256, 0, st>>>(u, out, Kd, eleMask, nodeMask, g, nullptr);
               return false;
           },
           [&](const double *r, double *z, int) {
               k_filter_apply<1><<<gn, 256, 0, st>>>(r, z, nullptr, nullptr, nullptr, g, inv);
               return false;
           },
           b, y, tol, maxIT, false, its, res, id);
    k_n2e<<<grid_for(Ne, 256), 256, 0, st>>>(y, eleFull0, Ne, g, tar);
    TG_LAUNCH_CHECK(id);
}

ConvFilter::ConvFilter(int nelx, int nely, int nelz, double r, int n, hipStream_t s) : num(n), st(s) {
    const char *id = "tg_convfilter_create:Input";
    X = nelx; Y = nely; Z = nelz;
    GY = num * nely; GX = num * nelx; GZ = num * nelz;
    const int c = (int)std::ceil(r);
    hr = c - 1;
    const int w = 2 * hr + 1;
    std::vector<double> h((size_t)w * w * w);
    // [dy,dx,dz] = meshgrid(-c+1:c-1, ...); h = max(0, r - sqrt(dx^2+dy^2+dz^2)) (TOP_GPU.m:116-118)
    for (int kz = 0; kz < w; ++kz)
        for (int kx = 0; kx < w; ++kx)
            for (int ky = 0; ky < w; ++ky) {
                const double dy = ky - hr, dx = kx - hr, dz = kz - hr;
                h[ky + w * (kx + w * kz)] = std::max(0.0, r - std::sqrt(dx * dx + dy * dy + dz * dz));
            }
    hd.upload(h, st);
    const int64_t ng = (int64_t)GY * GX * GZ;
    Hs.alloc(ng);
    tmp.alloc(ng);
    k_conv<2><<<grid_for(ng, 256), 256, 0, st>>>(nullptr, nullptr, hd, hr, GY, GX, GZ, num, Y, X, Hs);
    TG_LAUNCH_CHECK(id);
    TG_HIP(hipStreamSynchronize(st), id);
}

void ConvFilter::forward(const double *x, double *out) {
    const int64_t ng = (int64_t)GY * GX * GZ;
    k_conv<0><<<grid_for(ng, 256), 256, 0, st>>>(x, Hs, hd, hr, GY, GX, GZ, num, Y, X, out);
    TG_LAUNCH_CHECK("tg_convfilter:forward");
}

void ConvFilter::adjoint(const double *dc, double *out) {
    const int64_t ng = (int64_t)GY * GX * GZ;
    k_conv<1><<<grid_for(ng, 256), 256, 0, st>>>(dc, Hs, hd, hr, GY, GX, GZ, num, Y, X, out);
    TG_LAUNCH_CHECK("tg_convfilter:adjoint");
}

}  // namespace tg

using namespace tg;

extern "C" {

int tg_solver_unit_compliance(tg_solver *s, const double *d_U, double *d_ce) {
    return guard("tg_solver_unit_compliance:Input", [&] {
        const char *id = "tg_solver_unit_compliance:Input";
        if (!s || !s->s || !d_U || !d_ce) fail(TG_ERR_ARGS, id, "null argument");
        Solver &S = *s->s;
        DevLevel &F = S.lv[0];
        S.pw.z.zero(S.st);
        k_unpack<double, double><<<grid_for(F.Nn, 256), 256, 0, S.st>>>(d_U, S.pw.z, F.nodFull0, F.Nn);
        if (S.super) {
            if (!S.assembled) fail(TG_ERR_STATE, "tg_solver_unit_compliance:State", "assemble first (K_mat)");
            k_unit_compliance<1><<<grid_for(F.Ne, 256), 256, 0, S.st>>>(S.pw.z, nullptr, F.Ks64, F.ldk, F.eleFull0, F.Ne, F.g, d_ce);
        } else {
            k_unit_compliance<0><<<grid_for(F.Ne, 256), 256, 0, S.st>>>(S.pw.z, S.K0_64, nullptr, 0, F.eleFull0, F.Ne, F.g, d_ce);
        }
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
    });
}

int tg_compliance_grad(tg_solver *s, tg_filter *f, const double *d_x, double Emin, double E0, double penal,
                       double tol, int32_t max_it, double filter_tol, int32_t warm, double *d_U, double *c_out,
                       double *d_dcdx, int32_t *its_out) {
    return guard("tg_compliance_grad:Input", [&] {
        const char *id = "tg_compliance_grad:Input";
        if (!s || !s->s || !d_x || !d_U || !c_out || max_it < 1) fail(TG_ERR_ARGS, id, "bad arguments");
        Solver &S = *s->s;
        if (S.super) fail(TG_ERR_STATE, "tg_compliance_grad:State", "standard voxel mode only");
        if (f && (!f->f || f->f->Ne != S.lv[0].Ne)) fail(TG_ERR_ARGS, id, "filter does not match the solver's mesh");
        const hipStream_t st = S.st;
        const int64_t Ne = S.lv[0].Ne, ndof = 3 * S.lv[0].Nn;
        if ((int64_t)S.adjF.n != ndof) {
            S.adjF.upload(S.model.F.data(), ndof, st);  // load case 1
            S.adjXp.alloc(Ne);
            S.adjE.alloc(Ne);
            S.adjCe.alloc(Ne);
            S.adjG.alloc(Ne);
            S.adjPart.alloc(RED_BLOCKS);
            S.adjRes.alloc(1);
        }
        int fits = 0;
        double fres = 0;
        // xPhys = filter(x)
        if (f) {
            TG_HIP(hipStreamSynchronize(st), id);
            f->f->apply(d_x, S.adjXp, filter_tol, max_it, &fits, &fres);
            TG_HIP(hipStreamSynchronize(f->f->st), id);
        } else {
            TG_HIP(hipMemcpyAsync(S.adjXp.p, d_x, Ne * sizeof(double), hipMemcpyDeviceToDevice, st), id);
        }
        // E, K(E) U = F
        k_simp<<<grid_for(Ne, 256), 256, 0, st>>>(S.adjXp, S.adjE, Ne, Emin, E0, penal);
        TG_LAUNCH_CHECK(id);
        S.set_modulus(S.adjE);
        S.assemble();
        int its = 0;
        double res = 0;
        S.pcg(S.adjF, d_U, tol, max_it, warm, &its, &res);
        // ce_e = u_e' K0 u_e, c = sum E_e ce_e, g = -p (E0 - Emin) xPhys^(p-1) ce
        DevLevel &F = S.lv[0];
        S.pw.z.zero(st);
        k_unpack<double, double><<<grid_for(F.Nn, 256), 256, 0, st>>>(d_U, S.pw.z, F.nodFull0, F.Nn);
        k_unit_compliance<0><<<grid_for(Ne, 256), 256, 0, st>>>(S.pw.z, S.K0_64, nullptr, 0, F.eleFull0, Ne, F.g, S.adjCe);
        TG_LAUNCH_CHECK(id);
        double *h = nullptr;
        TG_HIP(hipHostMalloc((void **)&h, sizeof(double), hipHostMallocDefault), id);
        try {
            *c_out = reduce_dev(R_DOT, S.adjE, S.adjCe, Ne, S.adjPart, S.adjRes, h, st);
        } catch (...) {
            (void)hipHostFree(h);
            throw;
        }
        (void)hipHostFree(h);
        if (d_dcdx) {
            k_sens<<<grid_for(Ne, 256), 256, 0, st>>>(S.adjXp, S.adjCe, f ? S.adjG.p : d_dcdx, Ne, 1, Emin, E0, penal, 1.0);
            TG_LAUNCH_CHECK(id);
            if (f) {  // the Helmholtz filter is symmetric: its adjoint is the forward filter
                TG_HIP(hipStreamSynchronize(st), id);
                f->f->apply(S.adjG, d_dcdx, filter_tol, max_it, &fits, &fres);
                TG_HIP(hipStreamSynchronize(f->f->st), id);
            }
        }
        TG_HIP(hipStreamSynchronize(st), id);
        if (its_out) *its_out = its;
    });
}

int tg_solver_subvoxel_energy(tg_solver *s, const double *d_U, double *d_dc) {
    return guard("tg_solver_subvoxel_energy:Input", [&] {
        const char *id = "tg_solver_subvoxel_energy:Input";
        if (!s || !s->s || !d_U || !d_dc) fail(TG_ERR_ARGS, id, "null argument");
        Solver &S = *s->s;
        if (!S.super) fail(TG_ERR_STATE, id, "super-element mode only");
        DevLevel &F = S.lv[0];
        S.pw.z.zero(S.st);
        k_unpack<double, double><<<grid_for(F.Nn, 256), 256, 0, S.st>>>(d_U, S.pw.z, F.nodFull0, F.Nn);
        k_subvoxel_energy<<<grid_for(F.Ne * 64, 256), 256, 0, S.st>>>(S.pw.z, S.Imat, F.eleFull0, F.Ne, F.g, d_dc);
        TG_LAUNCH_CHECK(id);
        TG_HIP(hipStreamSynchronize(S.st), id);
    });
}

int tg_simp_modulus(const double *d_x, int64_t n, double Emin, double E0, double penal, double *d_E, void *stream) {
    return guard("tg_simp_modulus:Input", [&] {
        if (!d_x || !d_E || n < 0) fail(TG_ERR_ARGS, "tg_simp_modulus:Input", "bad arguments");
        if (n) k_simp<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(d_x, d_E, n, Emin, E0, penal);
        TG_LAUNCH_CHECK("tg_simp_modulus:HIP");
    });
}

int tg_sensitivity(const double *d_x, const double *d_ce, int64_t n, int32_t per, double Emin, double E0, double penal,
                   double cSolid, double *d_dc, void *stream) {
    return guard("tg_sensitivity:Input", [&] {
        if (!d_x || !d_ce || !d_dc || n < 0 || per < 1) fail(TG_ERR_ARGS, "tg_sensitivity:Input", "bad arguments");
        if (n) k_sens<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(d_x, d_ce, d_dc, n, per, Emin, E0, penal, cSolid);
        TG_LAUNCH_CHECK("tg_sensitivity:HIP");
    });
}

int tg_elementwise(int32_t op, const double *d_a, const double *d_b, double scalar, int64_t n, double *d_out,
                   void *stream) {
    return guard("tg_elementwise:Input", [&] {
        if (!d_a || !d_b || !d_out || n < 0) fail(TG_ERR_ARGS, "tg_elementwise:Input", "bad arguments");
        hipStream_t st = (hipStream_t)stream;
        if (!n) return;
        if (op == TG_EW_MUL) k_mul<<<grid_for(n, 256), 256, 0, st>>>(d_a, d_b, d_out, n);
        else if (op == TG_EW_DIV_FLOOR) k_div_floor<<<grid_for(n, 256), 256, 0, st>>>(d_a, d_b, d_out, scalar, n);
        else if (op == TG_EW_SET_WHERE) k_set_where<<<grid_for(n, 256), 256, 0, st>>>(d_a, d_b, d_out, scalar, n);
        else if (op == TG_EW_AXPY) k_axpy<<<grid_for(n, 256), 256, 0, st>>>(d_a, d_b, d_out, scalar, n);
        else fail(TG_ERR_ARGS, "tg_elementwise:Input", "unknown op %d", op);
        TG_LAUNCH_CHECK("tg_elementwise:HIP");
    });
}

int tg_reduce(int32_t op, const double *d_a, const double *d_b, int64_t n, double *out, void *stream) {
    return guard("tg_reduce:Input", [&] {
        if (!d_a || !out || n < 0 || ((op == TG_RED_DOT || op == TG_RED_MAXABSDIFF) && !d_b))
            fail(TG_ERR_ARGS, "tg_reduce:Input", "bad arguments");
        DevBuf<double> partial(RED_BLOCKS), res(1);
        double *h = nullptr;
        TG_HIP(hipHostMalloc((void **)&h, sizeof(double), hipHostMallocDefault), "tg_reduce:HIP");
        try {
            *out = reduce_dev(op, d_a, d_b, n, partial, res, h, (hipStream_t)stream);
        } catch (...) {
            (void)hipHostFree(h);
            throw;
        }
        (void)hipHostFree(h);
    });
}

int tg_oc_update(const double *d_x, const double *d_dc, const double *d_dv, const double *d_passive, int64_t n,
                 double fval, double move, double *d_xnew, double *change, int32_t *bisections, void *stream) {
    return guard("tg_oc_update:Input", [&] {
        const char *id = "tg_oc_update:Input";
        if (!d_x || !d_dc || !d_dv || !d_xnew || n < 1) fail(TG_ERR_ARGS, id, "bad arguments");
        hipStream_t st = (hipStream_t)stream;
        DevBuf<double> scal(8), partial(RED_BLOCKS), res(1);
        const double init[8] = {0.0, 1e9, 0.0, fval, move, (double)n, 0, 0};
        scal.upload(init, 8, st);
        double *h = nullptr;
        TG_HIP(hipHostMalloc((void **)&h, 8 * sizeof(double), hipHostMallocDefault), id);
        int steps = 0;
        bool open = true;
        try {
            // (l2-l1)/(l1+l2) halves each step from 1: at most ~100 steps; check every 16
            while (open) {
                for (int k = 0; k < 16; ++k) {
                    k_oc_step<<<RED_BLOCKS, RED_THREADS, 0, st>>>(d_x, d_dc, d_dv, d_passive, n, scal, partial);
                    k_oc_bisect<<<1, RED_THREADS, 0, st>>>(partial, RED_BLOCKS, scal);
                }
                TG_LAUNCH_CHECK(id);
                steps += 16;
                TG_HIP(hipMemcpyAsync(h, scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, st), id);
                TG_HIP(hipStreamSynchronize(st), id);
                open = (h[1] - h[0]) / (h[0] + h[1]) > 1e-6;
                if (steps > 400) fail(TG_ERR_NUMERIC, "tg_oc_update:Numeric", "bisection did not converge");
            }
            k_oc_apply<<<grid_for(n, 256), 256, 0, st>>>(d_x, d_dc, d_dv, d_passive, d_xnew, n, scal);
            TG_LAUNCH_CHECK(id);
            if (change) *change = reduce_dev(R_MAXABSDIFF, d_xnew, d_x, n, partial, res, h, st);
            if (bisections) *bisections = steps;
        } catch (...) {
            (void)hipHostFree(h);
            throw;
        }
        (void)hipHostFree(h);
    });
}

int tg_filter_create(const tg_model *m, double rMin, void *stream, tg_filter **out) {
    return guard("tg_filter_create:Input", [&] {
        if (!m || !out) fail(TG_ERR_ARGS, "tg_filter_create:Input", "null argument");
        auto *f = new tg_filter();
        try {
            f->f = new Filter(*m, rMin, (hipStream_t)stream);
        } catch (...) {
            delete f;
            throw;
        }
        *out = f;
    });
}

void tg_filter_destroy(tg_filter *f) {
    if (!f) return;
    if (f->f) (void)hipStreamSynchronize(f->f->st);
    delete f->f;
    delete f;
}

int tg_filter_apply(tg_filter *f, const double *d_src, double *d_tar, double tol, int32_t max_it, int32_t *its) {
    return guard("tg_filter_apply:Input", [&] {
        if (!f || !f->f || !d_src || !d_tar || max_it < 1) fail(TG_ERR_ARGS, "tg_filter_apply:Input", "bad arguments");
        int it = 0;
        double res = 0;
        f->f->apply(d_src, d_tar, tol, max_it, &it, &res);
        TG_HIP(hipStreamSynchronize(f->f->st), "tg_filter_apply:HIP");
        if (its) *its = it;
    });
}

int tg_convfilter_create(int32_t nelx, int32_t nely, int32_t nelz, double rMin_subvoxels, int32_t num, void *stream,
                         tg_convfilter **out) {
    return guard("tg_convfilter_create:Input", [&] {
        if (!out || nelx < 1 || nely < 1 || nelz < 1 || num < 1 || !(rMin_subvoxels > 0))
            fail(TG_ERR_ARGS, "tg_convfilter_create:Input", "bad arguments");
        auto *f = new tg_convfilter();
        try {
            f->f = new ConvFilter(nelx, nely, nelz, rMin_subvoxels, num, (hipStream_t)stream);
        } catch (...) {
            delete f;
            throw;
        }
        *out = f;
    });
}

void tg_convfilter_destroy(tg_convfilter *f) {
    if (!f) return;
    if (f->f) (void)hipStreamSynchronize(f->f->st);
    delete f->f;
    delete f;
}

int tg_convfilter_forward(tg_convfilter *f, const double *d_x, double *d_out) {
    return guard("tg_convfilter_forward:Input", [&] {
        if (!f || !f->f || !d_x || !d_out) fail(TG_ERR_ARGS, "tg_convfilter_forward:Input", "null argument");
        f->f->forward(d_x, d_out);
    });
}

int tg_convfilter_adjoint(tg_convfilter *f, const double *d_dc, double *d_out) {
    return guard("tg_convfilter_adjoint:Input", [&] {
        if (!f || !f->f || !d_dc || !d_out) fail(TG_ERR_ARGS, "tg_convfilter_adjoint:Input", "null argument");
        f->f->adjoint(d_dc, d_out);
    });
}

}  // extern "C"
