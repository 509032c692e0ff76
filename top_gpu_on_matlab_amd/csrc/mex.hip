// The six MEX entry points of the reference as C functions (mexFunction calling convention).
//
// Argument order, in-place contracts, class/shape checks and error identifiers follow the
// reference gateways; the kernels are this library's own deterministic gfx950 kernels:
//   * gather: one lane per element, coalesced column writes;
//   * scatter / scatter_accum3: a stable radix sort of (target row, entry) pairs, then one lane
//     per target row sums its entries in ascending entry order (no floating-point atomics);
//   * Galerkin assembly: one wavefront per coarse element (galerkin.h).
#include <hipcub/hipcub.hpp>

#include "galerkin.h"
#include "tg_internal.h"

namespace tg {
namespace {

struct Arr {
    const tg_array *a = nullptr;
    int64_t d[4] = {1, 1, 1, 1};
    int nd = 0;  // number of dims after dropping trailing singletons beyond 2 (MATLAB ndims)
    int64_t numel = 0;
    int cls = TG_CLASS_NONE;
    void *p = nullptr;
};

Arr arr(const tg_array *a) {
    Arr r;
    r.a = a;
    if (!a) return r;
    r.cls = a->cls;
    r.p = a->data;
    int nd = a->ndim < 0 ? 0 : (a->ndim > 4 ? 4 : a->ndim);
    r.numel = 1;
    for (int i = 0; i < nd; ++i) {
        r.d[i] = a->dims[i];
        r.numel *= a->dims[i];
    }
    if (nd == 0) r.numel = 0;
    int md = nd;
    while (md > 2 && r.d[md - 1] == 1) --md;
    r.nd = md < 2 ? 2 : md;
    return r;
}

bool is_float(int c) { return c == TG_SINGLE || c == TG_DOUBLE; }

hipStream_t S(void *s) { return (hipStream_t)s; }

// ---- gather ----------------------------------------------------------------------------
template <typename T, typename IDX>
__global__ void k_gather(T *__restrict__ uMat, const T *__restrict__ u, const IDX *__restrict__ eNod, int64_t Ne,
                         int64_t Nn) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Ne) return;
    int64_t node[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) node[n] = (int64_t)eNod[e + Ne * n] - 1;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const bool ok = node[n] >= 0 && node[n] < Nn;
            uMat[e + Ne * (3 * n + c)] = ok ? u[node[n] + Nn * c] : (T)0;
        }
}

// ---- deterministic keyed accumulation ---------------------------------------------------
// keys (0-based target rows; out-of-range rows mapped to M) are sorted stably with their entry
// ids; row r then owns the contiguous run of entries whose key is r.
__global__ void k_make_keys_enod(const int32_t *__restrict__ eNod, int64_t Ne, int64_t M, uint32_t *keys,
                                 uint32_t *vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // i = 8e + j (element-major)
    if (i >= Ne * 8) return;
    const int64_t e = i >> 3, j = i & 7;
    const int64_t k = (int64_t)eNod[e + Ne * j] - 1;
    keys[i] = (k >= 0 && k < M) ? (uint32_t)k : (uint32_t)M;
    vals[i] = (uint32_t)i;
}

__global__ void k_make_keys_list(const int32_t *__restrict__ keysIn, int64_t N, int64_t M, uint32_t *keys,
                                 uint32_t *vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t k = (int64_t)keysIn[i] - 1;
    keys[i] = (k >= 0 && k < M) ? (uint32_t)k : (uint32_t)M;
    vals[i] = (uint32_t)i;
}

__device__ inline int64_t lower_bound_u32(const uint32_t *a, int64_t n, uint32_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Y(r, c) += sum over sorted entries with key r of uMat(e, c + 3j), entry = 8e + j
template <typename T>
__global__ void k_seg_scatter_enod(T *__restrict__ Y, int64_t M, const uint32_t *__restrict__ skeys,
                                   const uint32_t *__restrict__ svals, int64_t N, const T *__restrict__ uMat,
                                   int64_t Ne) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    int64_t lo = lower_bound_u32(skeys, N, (uint32_t)r), hi = lower_bound_u32(skeys, N, (uint32_t)r + 1);
    T a0 = 0, a1 = 0, a2 = 0;
    for (int64_t k = lo; k < hi; ++k) {
        const int64_t ent = svals[k], e = ent >> 3, j = ent & 7;
        a0 += uMat[e + Ne * (0 + 3 * j)];
        a1 += uMat[e + Ne * (1 + 3 * j)];
        a2 += uMat[e + Ne * (2 + 3 * j)];
    }
    if (hi > lo) {
        Y[r] += a0;
        Y[r + M] += a1;
        Y[r + 2 * M] += a2;
    }
}

template <typename T>
__global__ void k_seg_scatter_list(T *__restrict__ xF, int64_t M, const uint32_t *__restrict__ skeys,
                                   const uint32_t *__restrict__ svals, int64_t N, const T *__restrict__ Tmp) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    int64_t lo = lower_bound_u32(skeys, N, (uint32_t)r), hi = lower_bound_u32(skeys, N, (uint32_t)r + 1);
    T a0 = 0, a1 = 0, a2 = 0;
    for (int64_t k = lo; k < hi; ++k) {
        const int64_t i = svals[k];
        a0 += Tmp[i];
        a1 += Tmp[i + N];
        a2 += Tmp[i + 2 * N];
    }
    if (hi > lo) {
        xF[r] += a0;
        xF[r + M] += a1;
        xF[r + 2 * M] += a2;
    }
}

// Sort (key, entry) pairs stably; returns sorted arrays in the provided buffers.
void sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *skeys, uint32_t *svals, int64_t n, int64_t M,
                hipStream_t st, const char *id) {
    int bits = 1;
    while ((1ULL << bits) <= (unsigned long long)M) ++bits;
    size_t tmp = 0;
    TG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, skeys, vals, svals, (int)n, 0, bits, st), id);
    DevBuf<char> t(tmp ? tmp : 1);
    TG_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, keys, skeys, vals, svals, (int)n, 0, bits, st), id);
    TG_HIP(hipStreamSynchronize(st), id);  // t is freed on return
}

bool is_device_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

// ---- Galerkin sources (reference layouts: Ks column-major per element, 1-based maps) ----
template <typename T>
struct SrcLevel2 {
    const int32_t *up;
    int64_t Ne;
    int e;
    const T *E, *K0, *uf, *ux;
    const int32_t *mapU;
    bool hasFixed;
    int64_t nFine, nU;
    __device__ int child(int s) const {
        const int32_t f = up[e + Ne * s];
        return (f > 0 && f <= nFine) ? f - 1 : -1;
    }
    __device__ T kch(int f, int i) const {
        const T mod = E[f];
        int32_t uid = hasFixed ? mapU[f] - 1 : -1;
        if (uid >= nU) uid = -1;
        if (uid >= 0) return uf[(size_t)uid * 576 + i] * mod + ux[(size_t)uid * 576 + i];
        return K0[i] * mod;
    }
};

template <typename T>
struct SrcHigher {
    const int32_t *up;
    int64_t Ne;
    int e;
    const T *prev;
    int64_t nFine;
    __device__ int child(int s) const {
        const int32_t f = up[e + Ne * s];
        return (f > 0 && f <= nFine) ? f - 1 : -1;
    }
    __device__ T kch(int f, int i) const { return prev[(size_t)f * 576 + i]; }
};

template <typename T>
struct SrcSuper {
    const int32_t *up;
    int64_t Ne;
    int e;
    const T *Ks0, *E, *uf, *ux;
    const int32_t *mapU;
    bool hasFixed;
    int64_t nFine, nU;
    __device__ int child(int s) const {
        const int32_t f = up[e + Ne * s];
        return (f > 0 && f <= nFine) ? f - 1 : -1;
    }
    __device__ T kch(int f, int i) const {
        int32_t uid1 = hasFixed ? mapU[f] : 0;
        if (uid1 > nU) uid1 = 0;
        if (uid1 > 0) {
            const size_t fb = (size_t)(uid1 - 1) * 576 * 64, mb = (size_t)f * 64;
            T acc = 0;
            for (int q = 0; q < 64; ++q) acc += uf[fb + (size_t)q * 576 + i] * E[mb + q];
            return acc + ux[(size_t)(uid1 - 1) * 576 + i];
        }
        return Ks0[(size_t)f * 576 + i];
    }
};

template <typename T, int MODE>
__global__ __launch_bounds__(GK_THREADS) void k_galerkin_mex(T *__restrict__ Ks, const int32_t *__restrict__ up,
                                                             int64_t Ne, int nCh, const T *__restrict__ P,
                                                             const T *a0, const T *a1, const T *a2, const T *a3,
                                                             const int32_t *mapU, int hasFixed, int64_t nFine,
                                                             int64_t nU) {
    __shared__ T sB[576], sK[576], sT[576];
    const int e = blockIdx.x;
    T acc[9];
    if (MODE == 0) {
        SrcLevel2<T> src{up, Ne, e, a0, a1, a2, a3, mapU, hasFixed != 0, nFine, nU};
        galerkin_block(acc, src, nCh, P, sB, sK, sT);
    } else if (MODE == 1) {
        SrcHigher<T> src{up, Ne, e, a0, nFine};
        galerkin_block(acc, src, nCh, P, sB, sK, sT);
    } else {
        SrcSuper<T> src{up, Ne, e, a0, a1, a2, a3, mapU, hasFixed != 0, nFine, nU};
        galerkin_block(acc, src, nCh, P, sB, sK, sT);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Ks[(size_t)e * 576 + galerkin_entry(i)] = acc[i];
}

}  // namespace
}  // namespace tg

using namespace tg;

extern "C" {

int Gathering_inplace(int nlhs, tg_array *, int nrhs, const tg_array *prhs, void *stream) {
    return guard("Gathering_inplace:Args", [&] {
        const char *A = "Gathering_inplace:Args", *Ty = "Gathering_inplace:Type", *Sh = "Gathering_inplace:Shape";
        if (nrhs != 3 || !prhs) fail(TG_ERR_ARGS, A, "Usage: Gathering_inplace(gUMat, gU, gE)");
        if (nlhs != 0) fail(TG_ERR_ARGS, A, "This function returns no outputs. gUMat is modified in-place.");
        Arr out = arr(&prhs[0]), u = arr(&prhs[1]), e = arr(&prhs[2]);
        if (!is_float(u.cls)) fail(TG_ERR_TYPE, Ty, "gU must be single or double gpuArray.");
        if (out.cls != u.cls) fail(TG_ERR_TYPE, Ty, "gUMat precision must match gU.");
        if (!(e.cls == TG_INT32 || e.cls == TG_UINT32)) fail(TG_ERR_TYPE, Ty, "gE must be int32 or uint32 gpuArray.");
        if (u.nd != 2 || u.d[1] != 3) fail(TG_ERR_SHAPE, Sh, "gU must be (numNodes x 3).");
        if (e.nd != 2 || e.d[1] != 8) fail(TG_ERR_SHAPE, Sh, "gE must be (nElem x 8).");
        const int64_t Ne = e.d[0], Nn = u.d[0];
        if (out.nd != 2 || out.d[0] != Ne || out.d[1] != 24) fail(TG_ERR_SHAPE, Sh, "gUMat must be (nElem x 24).");
        if (Ne == 0) return;
        const unsigned g = grid_for(Ne, 256);
        if (u.cls == TG_DOUBLE) {
            if (e.cls == TG_INT32) k_gather<double, int32_t><<<g, 256, 0, S(stream)>>>((double *)out.p, (const double *)u.p, (const int32_t *)e.p, Ne, Nn);
            else k_gather<double, uint32_t><<<g, 256, 0, S(stream)>>>((double *)out.p, (const double *)u.p, (const uint32_t *)e.p, Ne, Nn);
        } else {
            if (e.cls == TG_INT32) k_gather<float, int32_t><<<g, 256, 0, S(stream)>>>((float *)out.p, (const float *)u.p, (const int32_t *)e.p, Ne, Nn);
            else k_gather<float, uint32_t><<<g, 256, 0, S(stream)>>>((float *)out.p, (const float *)u.p, (const uint32_t *)e.p, Ne, Nn);
        }
        TG_LAUNCH_CHECK("Gathering_inplace:CUDA");
    });
}

int Scattering_inplace(int, tg_array *, int nrhs, const tg_array *prhs, void *stream) {
    return guard("Scattering_inplace:arity", [&] {
        const char *Sh = "Scattering_inplace:shape", *Ty = "Scattering_inplace:type", *Cu = "Scattering_inplace:cuda";
        if (nrhs != 3 || !prhs) fail(TG_ERR_ARGS, "Scattering_inplace:arity", "Expected 3 inputs: (Y, eNodMat, uMat). No outputs.");
        Arr Y = arr(&prhs[0]), en = arr(&prhs[1]), um = arr(&prhs[2]);
        if (Y.nd != 2 || Y.d[1] != 3) fail(TG_ERR_SHAPE, Sh, "Y must be [numNodes x 3].");
        if (!is_float(Y.cls)) fail(TG_ERR_TYPE, Ty, "Y must be single or double gpuArray.");
        if (en.cls != TG_INT32) fail(TG_ERR_TYPE, Ty, "eNodMat must be gpuArray(int32).");
        if (en.nd != 2 || en.d[1] != 8) fail(TG_ERR_SHAPE, Sh, "eNodMat must be [Ne x 8] int32 (1-based).");
        const int64_t Ne = en.d[0], Nn = Y.d[0];
        if (um.nd != 2 || um.d[0] != Ne || um.d[1] != 24) fail(TG_ERR_SHAPE, Sh, "uMat must be [Ne x 24] double.");
        if (um.cls != Y.cls) fail(TG_ERR_TYPE, Ty, "uMat must have the class of Y.");
        hipStream_t st = S(stream);
        if (Ne > 0 && Nn > 0) {
            const int64_t N = Ne * 8;
            DevBuf<uint32_t> k(N), v(N), sk(N), sv(N);
            k_make_keys_enod<<<grid_for(N, 256), 256, 0, st>>>((const int32_t *)en.p, Ne, Nn, k, v);
            TG_LAUNCH_CHECK(Cu);
            sort_pairs(k, v, sk, sv, N, Nn, st, Cu);
            if (Y.cls == TG_DOUBLE)
                k_seg_scatter_enod<double><<<grid_for(Nn, 256), 256, 0, st>>>((double *)Y.p, Nn, sk, sv, N, (const double *)um.p, Ne);
            else
                k_seg_scatter_enod<float><<<grid_for(Nn, 256), 256, 0, st>>>((float *)Y.p, Nn, sk, sv, N, (const float *)um.p, Ne);
            TG_LAUNCH_CHECK(Cu);
        }
        TG_HIP(hipStreamSynchronize(st), Cu);
    });
}

int scatter_accum3_inplace(int nlhs, tg_array *, int nrhs, const tg_array *prhs, void *stream) {
    return guard("scatter_accum3:args", [&] {
        const char *Sh = "scatter_accum3:shape", *Ty = "scatter_accum3:type", *Cu = "scatter_accum3:cuda";
        if (nrhs != 3 || !prhs)
            fail(TG_ERR_ARGS, "scatter_accum3:args", "Usage: scatter_accum3_inplace(transferMatGPU[int32], TmpGPU[Nx3], xFinerGPU[Mx3]).");
        if (nlhs != 0) fail(TG_ERR_ARGS, "scatter_accum3:nlhs", "No output. xFiner is updated in-place.");
        Arr keys = arr(&prhs[0]), tmp = arr(&prhs[1]), xf = arr(&prhs[2]);
        for (int i = 0; i < 3; ++i) {
            Arr a = arr(&prhs[i]);
            if (a.numel > 0 && !is_device_ptr(a.p)) fail(TG_ERR_ARGS, "scatter_accum3:gpu", "All inputs must be gpuArray.");
        }
        if (tmp.nd != 2 || tmp.d[1] != 3) fail(TG_ERR_SHAPE, Sh, "`Tmp` must be [N x 3] gpuArray.");
        const int64_t N = tmp.d[0];
        if (xf.nd != 2 || xf.d[1] != 3) fail(TG_ERR_SHAPE, Sh, "`xFiner` must be [M x 3] gpuArray.");
        const int64_t M = xf.d[0];
        if (keys.cls != TG_INT32) fail(TG_ERR_TYPE, Ty, "`transferMat` must be int32 gpuArray.");
        if (tmp.cls != xf.cls) fail(TG_ERR_TYPE, Ty, "`Tmp` and `xFiner` must have the same class (single/double).");
        if (!is_float(tmp.cls)) fail(TG_ERR_TYPE, Ty, "Only single or double supported for `Tmp`/`xFiner`.");
        if (keys.numel != N) fail(TG_ERR_SHAPE, Sh, "`transferMat` must hold exactly N = size(Tmp,1) keys.");
        if (N == 0 || M == 0) return;
        hipStream_t st = S(stream);
        DevBuf<uint32_t> k(N), v(N), sk(N), sv(N);
        k_make_keys_list<<<grid_for(N, 256), 256, 0, st>>>((const int32_t *)keys.p, N, M, k, v);
        TG_LAUNCH_CHECK(Cu);
        sort_pairs(k, v, sk, sv, N, M, st, Cu);
        if (xf.cls == TG_DOUBLE)
            k_seg_scatter_list<double><<<grid_for(M, 256), 256, 0, st>>>((double *)xf.p, M, sk, sv, N, (const double *)tmp.p);
        else
            k_seg_scatter_list<float><<<grid_for(M, 256), 256, 0, st>>>((float *)xf.p, M, sk, sv, N, (const float *)tmp.p);
        TG_LAUNCH_CHECK(Cu);
        TG_HIP(hipStreamSynchronize(st), Cu);  // scratch buffers are released on return
    });
}

int assembleKs_level2_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream) {
    return guard("assembleKs_level2_inplace:Input", [&] {
        const char *In = "assembleKs_level2_inplace:Input", *Ty = "assembleKs_level2_inplace:Type",
                   *Sh = "assembleKs_level2_inplace:Shape", *Cu = "assembleKs_level2_inplace:CUDA";
        if (nrhs != 8 || !prhs) fail(TG_ERR_ARGS, In, "Expected 8 inputs.");
        Arr Ks = arr(&prhs[0]), up = arr(&prhs[1]), em = arr(&prhs[2]), k0 = arr(&prhs[3]), pch = arr(&prhs[4]),
            mu = arr(&prhs[5]), uf = arr(&prhs[6]), ux = arr(&prhs[7]);
        const int cls = Ks.cls;
        if (!is_float(cls)) fail(TG_ERR_TYPE, Ty, "d_Ks must be single/double gpuArray.");
        if (up.cls != TG_INT32 || mu.cls != TG_INT32) fail(TG_ERR_TYPE, Ty, "d_upMap and d_mapUnique must be int32 gpuArray.");
        if (Ks.d[0] != 24 || Ks.d[1] != 24 || Ks.d[3] != 1) fail(TG_ERR_SHAPE, Sh, "d_Ks must be [24 x 24 x nElem].");
        const int64_t Ne = Ks.d[2];
        if (up.nd != 2 || up.d[0] != Ne) fail(TG_ERR_SHAPE, Sh, "d_upMap must be [nElem x nCh].");
        const int nCh = (int)up.d[1];
        if (pch.d[0] != 24 || pch.d[1] != 24 || pch.d[2] != nCh || pch.d[3] != 1) fail(TG_ERR_SHAPE, Sh, "d_Pch must be [24 x 24 x nCh].");
        if (k0.nd != 2 || k0.d[0] * k0.d[1] != 576) fail(TG_ERR_SHAPE, Sh, "d_K_0 must contain 576 elements.");
        int64_t nU = 0;
        if (uf.numel > 0) {
            if (uf.nd != 2 || uf.d[0] != 576) fail(TG_ERR_SHAPE, Sh, "d_uniqueFree must be [576 x nUnique] or empty.");
            nU = uf.d[1];
        }
        if (ux.numel > 0 && (ux.nd != 2 || ux.d[0] != 576 || ux.d[1] != nU))
            fail(TG_ERR_SHAPE, Sh, "d_uniqueFixed must be [576 x nUnique] matching d_uniqueFree.");
        if (em.cls != cls || k0.cls != cls || pch.cls != cls) fail(TG_ERR_TYPE, Ty, "d_eleModulus/d_K_0/d_Pch must match d_Ks class.");
        if (uf.numel > 0 && uf.cls != cls) fail(TG_ERR_TYPE, Ty, "d_uniqueFree class must match d_Ks.");
        if (ux.numel > 0 && ux.cls != cls) fail(TG_ERR_TYPE, Ty, "d_uniqueFixed class must match d_Ks.");
        const int hasFixed = (nU > 0 && mu.p && uf.p && ux.numel > 0 && ux.p) ? 1 : 0;
        const int64_t nFine = hasFixed ? std::min(em.numel, mu.numel) : em.numel;
        hipStream_t st = S(stream);
        if (Ne > 0) {
            if (cls == TG_DOUBLE)
                k_galerkin_mex<double, 0><<<(unsigned)Ne, GK_THREADS, 0, st>>>((double *)Ks.p, (const int32_t *)up.p, Ne, nCh, (const double *)pch.p,
                    (const double *)em.p, (const double *)k0.p, (const double *)uf.p, (const double *)ux.p, (const int32_t *)mu.p, hasFixed, nFine, nU);
            else
                k_galerkin_mex<float, 0><<<(unsigned)Ne, GK_THREADS, 0, st>>>((float *)Ks.p, (const int32_t *)up.p, Ne, nCh, (const float *)pch.p,
                    (const float *)em.p, (const float *)k0.p, (const float *)uf.p, (const float *)ux.p, (const int32_t *)mu.p, hasFixed, nFine, nU);
            TG_LAUNCH_CHECK(Cu);
        }
        TG_HIP(hipStreamSynchronize(st), Cu);
        if (nlhs > 0 && plhs) plhs[0] = prhs[0];
    });
}

int assembleKs_higherLevel_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream) {
    return guard("assembleKs_higherLevel_inplace:Input", [&] {
        const char *In = "assembleKs_higherLevel_inplace:Input", *Ty = "assembleKs_higherLevel_inplace:Type",
                   *Sh = "assembleKs_higherLevel_inplace:Shape", *Cu = "assembleKs_higherLevel_inplace:CUDA";
        if (nrhs != 4 || !prhs) fail(TG_ERR_ARGS, In, "Expected 4 inputs: (d_Ks, d_upMap, d_KsPrev, d_Psub).");
        Arr Ks = arr(&prhs[0]), up = arr(&prhs[1]), prev = arr(&prhs[2]), ps = arr(&prhs[3]);
        const int cls = Ks.cls;
        if (!is_float(cls)) fail(TG_ERR_TYPE, Ty, "d_Ks must be single/double gpuArray.");
        if (up.cls != TG_INT32) fail(TG_ERR_TYPE, Ty, "d_upMap must be int32 gpuArray.");
        if (prev.cls != cls || ps.cls != cls) fail(TG_ERR_TYPE, Ty, "d_KsPrev and d_Psub must match d_Ks class.");
        if (Ks.d[0] != 24 || Ks.d[1] != 24 || Ks.d[3] != 1) fail(TG_ERR_SHAPE, Sh, "d_Ks must be [24 x 24 x nElem].");
        const int64_t Ne = Ks.d[2];
        if (up.nd != 2 || up.d[0] != Ne) fail(TG_ERR_SHAPE, Sh, "d_upMap must be [nElem x nSub].");
        const int nSub = (int)up.d[1];
        if (prev.d[0] != 24 || prev.d[1] != 24 || prev.d[3] != 1) fail(TG_ERR_SHAPE, Sh, "d_KsPrev must be [24 x 24 x nFinElem].");
        if (ps.d[0] != 24 || ps.d[1] != 24 || ps.d[2] != nSub || ps.d[3] != 1) fail(TG_ERR_SHAPE, Sh, "d_Psub must be [24 x 24 x nSub].");
        const int64_t nFine = prev.d[2];
        hipStream_t st = S(stream);
        if (Ne > 0) {
            if (cls == TG_DOUBLE)
                k_galerkin_mex<double, 1><<<(unsigned)Ne, GK_THREADS, 0, st>>>((double *)Ks.p, (const int32_t *)up.p, Ne, nSub, (const double *)ps.p,
                    (const double *)prev.p, nullptr, nullptr, nullptr, nullptr, 0, nFine, 0);
            else
                k_galerkin_mex<float, 1><<<(unsigned)Ne, GK_THREADS, 0, st>>>((float *)Ks.p, (const int32_t *)up.p, Ne, nSub, (const float *)ps.p,
                    (const float *)prev.p, nullptr, nullptr, nullptr, nullptr, 0, nFine, 0);
            TG_LAUNCH_CHECK(Cu);
        }
        TG_HIP(hipStreamSynchronize(st), Cu);
        if (nlhs > 0 && plhs) plhs[0] = prhs[0];
    });
}

int assembleKs_level2_superEle_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream) {
    return guard("assembleKs_level2_superEle_inplace:Input", [&] {
        const char *In = "assembleKs_level2_superEle_inplace:Input", *Ty = "assembleKs_level2_superEle_inplace:Type",
                   *Sh = "assembleKs_level2_superEle_inplace:Shape", *Cu = "assembleKs_level2_superEle_inplace:CUDA";
        if (nrhs != 8 || !prhs)
            fail(TG_ERR_ARGS, In, "Expected 8 inputs: d_Ks, d_upMap, d_Ks0, d_Psub, d_mapUnique, d_eleModulus, d_uniqueFree, d_uniqueFixed.");
        Arr Ks = arr(&prhs[0]), up = arr(&prhs[1]), ks0 = arr(&prhs[2]), ps = arr(&prhs[3]), mu = arr(&prhs[4]),
            em = arr(&prhs[5]), uf = arr(&prhs[6]), ux = arr(&prhs[7]);
        const int cls = Ks.cls;
        if (!is_float(cls)) fail(TG_ERR_TYPE, Ty, "d_Ks must be single/double gpuArray.");
        if (up.cls != TG_INT32 || mu.cls != TG_INT32) fail(TG_ERR_TYPE, Ty, "d_upMap and d_mapUnique must be int32 gpuArray.");
        if (ks0.cls != cls || ps.cls != cls || em.cls != cls) fail(TG_ERR_TYPE, Ty, "d_Ks0, d_Psub, d_eleModulus must match d_Ks class.");
        if (uf.numel > 0 && uf.cls != cls) fail(TG_ERR_TYPE, Ty, "d_uniqueFree class must match d_Ks.");
        if (ux.numel > 0 && ux.cls != cls) fail(TG_ERR_TYPE, Ty, "d_uniqueFixed class must match d_Ks.");
        if (Ks.d[0] != 24 || Ks.d[1] != 24 || Ks.d[3] != 1) fail(TG_ERR_SHAPE, Sh, "d_Ks must be [24 x 24 x nElem].");
        const int64_t Ne = Ks.d[2];
        if (up.nd != 2 || up.d[0] != Ne) fail(TG_ERR_SHAPE, Sh, "d_upMap must be [nElem x nSub].");
        const int nSub = (int)up.d[1];
        if (ks0.numel % 576 != 0) fail(TG_ERR_SHAPE, Sh, "d_Ks0 total elements must be a multiple of 576.");
        if (ps.d[0] != 24 || ps.d[1] != 24 || ps.d[2] != nSub || ps.d[3] != 1) fail(TG_ERR_SHAPE, Sh, "d_Psub must be [24 x 24 x nSub].");
        if (em.numel % 64 != 0) fail(TG_ERR_SHAPE, Sh, "d_eleModulus total elements must be a multiple of 64.");
        int64_t nU = 0;
        if (uf.numel > 0) {
            if (uf.nd == 3) {
                if (uf.d[0] != 576 || uf.d[1] != 64) fail(TG_ERR_SHAPE, Sh, "d_uniqueFree must be [576 x 64 x nUnique].");
                nU = uf.d[2];
            } else if (uf.nd == 2) {
                if (uf.d[0] != 576 * 64) fail(TG_ERR_SHAPE, Sh, "d_uniqueFree must be [576*64 x nUnique] or [576 x 64 x nUnique].");
                nU = uf.d[1];
            } else {
                fail(TG_ERR_SHAPE, Sh, "d_uniqueFree must be 2-D or 3-D.");
            }
        }
        if (ux.numel > 0 && (ux.nd != 2 || ux.d[0] != 576 || ux.d[1] != nU))
            fail(TG_ERR_SHAPE, Sh, "d_uniqueFixed must be [576 x nUnique] matching d_uniqueFree.");
        const int hasFixed = (nU > 0 && mu.p && uf.p && ux.numel > 0 && ux.p) ? 1 : 0;
        int64_t nFine = std::min(ks0.numel / 576, em.numel / 64);
        if (hasFixed) nFine = std::min(nFine, mu.numel);
        hipStream_t st = S(stream);
        if (Ne > 0) {
            if (cls == TG_DOUBLE)
                k_galerkin_mex<double, 2><<<(unsigned)Ne, GK_THREADS, 0, st>>>((double *)Ks.p, (const int32_t *)up.p, Ne, nSub, (const double *)ps.p,
                    (const double *)ks0.p, (const double *)em.p, (const double *)uf.p, (const double *)ux.p, (const int32_t *)mu.p, hasFixed, nFine, nU);
            else
                k_galerkin_mex<float, 2><<<(unsigned)Ne, GK_THREADS, 0, st>>>((float *)Ks.p, (const int32_t *)up.p, Ne, nSub, (const float *)ps.p,
                    (const float *)ks0.p, (const float *)em.p, (const float *)uf.p, (const float *)ux.p, (const int32_t *)mu.p, hasFixed, nFine, nU);
            TG_LAUNCH_CHECK(Cu);
        }
        TG_HIP(hipStreamSynchronize(st), Cu);
        if (nlhs > 0 && plhs) plhs[0] = prhs[0];
    });
}

}  // extern "C"
