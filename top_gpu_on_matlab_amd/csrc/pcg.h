// Preconditioned CG on device vectors (Solving_PCG, TOP_GPU.m:602-667), shared by the elasticity
// solve and the PDE filter.  Dots are deterministic two-stage reductions; the host reads one
// scalar per iteration for the stopping test ||r||/||b|| < tol, exactly as the reference does.
#pragma once

#include <cmath>

#include "kernels.h"
#include "tg_internal.h"

namespace tg {

enum { S_PV = 0, S_RR = 1, S_BB = 2, S_XA = 3, S_XB = 4 };

// A dot left as partial sums for its consumer to finish (one launch fewer per dot).
struct DeferredDot {
    const double *parts = nullptr;
    int n = 0, slot = -1;
};

// The sum of n partials in k_reduce_final's fixed order (the same threads, strides and tree: the same
// value), in every thread of a RED_THREADS-wide workgroup.
__device__ inline double sum_partials(const double *__restrict__ parts, int n) {
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += parts[i];
    s = block_sum<double>(s);
    __syncthreads();  // every thread has read the sum before the shared array is used again
    return s;
}

// r -= lambda v; partial sums of r.r  (TOP_GPU.m:631-632, 637).  The iterate's update y += lambda p
// (:630) is applied later, by the next direction update, which reads p anyway (or by k_pcg_axpy when
// the loop stops): the same operation on the same values, one pass over p and y fewer per iteration.
// R32: also store r rounded to FP32 (the mixed V-cycle's input, TOP_GPU.m:775-777), saving a
// conversion pass.  pvp (optional): p.v as the operator's partial sums, finished here (workgroup 0
// stores it in scal[ipv] for later launches).
template <bool R32>
__global__ __launch_bounds__(RED_THREADS) static void k_pcg_update(double *__restrict__ r, const double *__restrict__ v,
                                                                   int64_t n, double *__restrict__ scal, int ix1,
                                                                   int ipv, double *__restrict__ partial,
                                                                   float *__restrict__ r32, const double *__restrict__ pvp,
                                                                   int npv) {
    double pv;
    if (pvp) {
        pv = sum_partials(pvp, npv);
        if (blockIdx.x == 0 && threadIdx.x == 0) scal[ipv] = pv;
    } else {
        pv = scal[ipv];
    }
    const double lam = scal[ix1] / pv;
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const double rr = r[i] - lam * v[i];
        r[i] = rr;
        if (R32) r32[i] = (float)rr;
        s += rr * rr;
    }
    s = block_sum<double>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// y += (x1 / pv) p, then p = z + (x2 / x1) p  (TOP_GPU.m:630, 652): the previous iteration's update of
// the iterate and the new direction; Z = float: z as the mixed V-cycle left it (widening exact).  One
// thread per entry: a grid-stride form that could also finish the z.r dot measured slower (5 streams).
template <typename Z>
__global__ static void k_pcg_direction(double *__restrict__ p, double *__restrict__ y, const Z *__restrict__ z,
                                       int64_t n, const double *__restrict__ scal, int ix2, int ix1, int ipv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double pi = p[i];
        y[i] = y[i] + scal[ix1] / scal[ipv] * pi;
        p[i] = (double)z[i] + scal[ix2] / scal[ix1] * pi;
    }
}

// y += (x1 / pv) p: the last iteration's update of the iterate, when the loop stops on the residual
__global__ static void k_pcg_axpy(double *__restrict__ y, const double *__restrict__ p, int64_t n,
                                  const double *__restrict__ scal, int ix1, int ipv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + scal[ix1] / scal[ipv] * p[i];
}

// p = (double) z32 (the first search direction, from an FP32 preconditioner output)
__global__ static void k_widen(double *__restrict__ p, const float *__restrict__ z, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (double)z[i];
}

// z64 = (double) z32 and partial sums of z64 . r64 (the mixed V-cycle's exit, fused with z'r)
__global__ __launch_bounds__(RED_THREADS) static void k_convert_dot(const float *__restrict__ z32,
                                                                    const double *__restrict__ r, double *__restrict__ z,
                                                                    int64_t n, double *__restrict__ partial) {
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const double v = (double)z32[i];
        z[i] = v;
        s += v * r[i];
    }
    s = block_sum<double>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ static void k_sub(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c,
                             int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] - b[i];
}

struct PcgWork {
    int64_t n = 0;
    hipStream_t st = nullptr;
    DevBuf<double> r, z, p, v, partial, scal;
    bool defer_ok = false;      // the operator may leave p.v as partial sums ...
    DeferredDot deferred;       // ... here, for the update to finish
    double *h_scal = nullptr;  // pinned host copies of the scalars the host reads (S_BB, S_RR) ...
    double *d_hscal = nullptr;  // ... and their device address, written by the final reductions
    unsigned long long seq = 0;  // the last publish number of S_RR (at h_scal[16], device d_hscal[16])
    float *r32 = nullptr;     // optional FP32 copy of r kept by the update (mixed V-cycle input)
    bool r32_fresh = false;   // r32 holds the current r
    float *z32 = nullptr;     // optional FP32 preconditioner output (mixed V-cycle) ...
    bool z32_fresh = false;   // ... set by ptv when it left z there instead of in z

    void init(int64_t count, hipStream_t s) {
        n = count;
        st = s;
        r.alloc(n);
        z.alloc(n);
        p.alloc(n);
        v.alloc(n);
        partial.alloc(RED_BLOCKS);
        scal.alloc(16);
        scal.zero(st);
        if (!h_scal) {
            TG_HIP(hipHostMalloc((void **)&h_scal, 17 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent),
                   "tg_pcg:alloc");
            *(unsigned long long *)(h_scal + 16) = 0ull;  // publish number 0
            TG_HIP(hipHostGetDevicePointer((void **)&d_hscal, h_scal, 0), "tg_pcg:alloc");
        }
    }
    ~PcgWork() {
        if (h_scal) (void)hipHostFree(h_scal);
    }
    void dot(const double *a, const double *b, int slot) {
        k_dot_partial<double><<<RED_BLOCKS, RED_THREADS, 0, st>>>(a, b, n, partial);
        k_reduce_final<<<1, RED_THREADS, 0, st>>>(partial, RED_BLOCKS, scal, slot, d_hscal);
        TG_LAUNCH_CHECK("tg_pcg:dot");
    }
    // A scalar the stream's last final reduction into slot also left in pinned host memory (dot, or the
    // residual norm of the update), once the stream has finished.
    double read(int slot) {
        TG_HIP(hipStreamSynchronize(st), "tg_pcg");
        return h_scal[slot];
    }
    // r.r of the update: reduced with a publish number, which the host polls in pinned memory (no
    // stream synchronisation in the loop).  The stream is queried now and then, so a device error, or
    // a stream that finished without publishing, ends the wait.
    void reduce_rr() {
        ++seq;
        k_reduce_final<<<1, RED_THREADS, 0, st>>>(partial, RED_BLOCKS, scal, S_RR, d_hscal,
                                                  (unsigned long long *)(d_hscal + 16), seq);
    }
    double wait_rr() {
        unsigned long long *hs = (unsigned long long *)(h_scal + 16);
        for (unsigned spins = 1; __atomic_load_n(hs, __ATOMIC_ACQUIRE) != seq; ++spins) {
            if ((spins & 4095) == 0) {
                const hipError_t e = hipStreamQuery(st);
                if (e != hipErrorNotReady) {
                    TG_HIP(e, "tg_pcg");
                    if (__atomic_load_n(hs, __ATOMIC_ACQUIRE) != seq)
                        fail(TG_ERR_HIP, "tg_pcg", "residual norm not published by a finished stream");
                }
            }
        }
        return h_scal[S_RR];
    }

    // atx(u, out, slot): out = A u;  ptv(r, z, slot): z = M r.  Each returns true when it has also
    // left u.out (resp. z.r) in scal[slot]; otherwise the dot is computed here.  y holds the start
    // vector (zeros unless warm).
    template <class AtX, class PtV>
    void run(AtX atx, PtV ptv, const double *b, double *y, double tol, int maxIT, bool warm, int *its_out,
             double *res_out, const char *id) {
        dot(b, b, S_BB);
        const double normB = std::sqrt(read(S_BB));
        if (warm) {
            atx(y, v.p, -1);
            k_sub<<<grid_for(n, 256), 256, 0, st>>>(b, v, r, n);
            TG_LAUNCH_CHECK(id);
        } else {
            TG_HIP(hipMemcpyAsync(r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, st), id);
        }
        int sx1 = S_XA, sx2 = S_XB;
        r32_fresh = false;
        z32_fresh = false;
        defer_ok = false;
        deferred = DeferredDot{};
        // ptv may leave z in z32 (then with its dot fused); widen it wherever z itself is needed
        auto take = [&](int slot) {  // the dot of slot, if left as partial sums
            DeferredDot d = deferred;
            deferred = DeferredDot{};
            if (d.parts && d.slot != slot) fail(TG_ERR_STATE, id, "deferred dot in an unexpected slot");
            return d;
        };
        auto precond = [&](int slot) {
            const bool dotted = ptv(r.p, z.p, slot);
            if (z32_fresh && !dotted) {
                k_widen<<<grid_for(n, 256), 256, 0, st>>>(z.p, z32, n);
                z32_fresh = false;
            }
            if (!dotted) dot(z, r, slot);
        };
        precond(sx1);
        if (z32_fresh) k_widen<<<grid_for(n, 256), 256, 0, st>>>(p.p, z32, n);
        else TG_HIP(hipMemcpyAsync(p.p, z.p, n * sizeof(double), hipMemcpyDeviceToDevice, st), id);
        int its = 0;
        double res = 0;
        for (its = 1; its <= maxIT; ++its) {
            bool pv_dotted;
            {
                struct Defer {  // deferral allowed for this operator call only (also when it throws)
                    bool &f;
                    explicit Defer(bool &x) : f(x) { f = true; }
                    ~Defer() { f = false; }
                } allow(defer_ok);
                pv_dotted = atx(p.p, v.p, S_PV);
            }
            if (!pv_dotted) dot(p, v, S_PV);
            const DeferredDot dpv = take(S_PV);
            if (r32) {
                k_pcg_update<true><<<RED_BLOCKS, RED_THREADS, 0, st>>>(r, v, n, scal, sx1, S_PV, partial, r32,
                                                                       dpv.parts, dpv.n);
                r32_fresh = true;
            } else {
                k_pcg_update<false><<<RED_BLOCKS, RED_THREADS, 0, st>>>(r, v, n, scal, sx1, S_PV, partial, nullptr,
                                                                        dpv.parts, dpv.n);
            }
            reduce_rr();
            TG_LAUNCH_CHECK(id);
            res = std::sqrt(wait_rr()) / normB;
            if (!std::isfinite(res)) fail(TG_ERR_NUMERIC, id, "non-finite residual at iteration %d", its);
            if (res < tol) {
                k_pcg_axpy<<<grid_for(n, 256), 256, 0, st>>>(y, p, n, scal, sx1, S_PV);
                TG_LAUNCH_CHECK(id);
                break;
            }
            precond(sx2);
            if (z32_fresh) k_pcg_direction<float><<<grid_for(n, 256), 256, 0, st>>>(p, y, z32, n, scal, sx2, sx1, S_PV);
            else k_pcg_direction<double><<<grid_for(n, 256), 256, 0, st>>>(p, y, z, n, scal, sx2, sx1, S_PV);
            TG_LAUNCH_CHECK(id);
            std::swap(sx1, sx2);
        }
        if (its > maxIT) its = maxIT;
        if (its_out) *its_out = its;
        if (res_out) *res_out = res;
    }
};

}  // namespace tg
