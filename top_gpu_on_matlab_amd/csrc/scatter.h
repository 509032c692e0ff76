// Helpers shared by the scatter entry points (mex.hip, noalloc.hip): MEX-style array views, the
// (row, entry) keys of a connectivity, its stable radix sort and the node-owned segmented sum.
#pragma once

#include <hipcub/hipcub.hpp>

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

// Bits of the radix sort for keys 0..M (M marks out-of-range rows).
int key_bits(int64_t M) {
    int bits = 1;
    while ((1ULL << bits) <= (unsigned long long)M) ++bits;
    return bits;
}

// The item count of hipcub's radix sort is an int: larger inputs are rejected, not truncated.
void check_items(int64_t n, const char *id) {
    if (n > (int64_t)INT32_MAX)
        fail(TG_ERR_SHAPE, id, "%lld (row, entry) pairs exceed the 2^31-1 the sort handles", (long long)n);
}

// Stable radix sort of (key, entry) pairs in double buffers; temp/tmp: caller-provided storage
// (tmp == 0 on entry: query only).  Results are in keys.Current() / vals.Current().
void sort_pairs_db(hipcub::DoubleBuffer<uint32_t> &keys, hipcub::DoubleBuffer<uint32_t> &vals, int64_t n, int64_t M,
                   void *temp, size_t &tmp, hipStream_t st, const char *id) {
    check_items(n, id);
    TG_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tmp, keys, vals, (int)n, 0, key_bits(M), st), id);
}

// Argument checks of Scattering_inplace (Scattering_inplace.cu:95-139), shared by both variants.
struct ScatterArgs {
    Arr Y, en, um;
    int64_t Ne = 0, Nn = 0;
};

ScatterArgs scatter_args(int nrhs, const tg_array *prhs) {
    const char *Sh = "Scattering_inplace:shape", *Ty = "Scattering_inplace:type";
    if (nrhs != 3 || !prhs) fail(TG_ERR_ARGS, "Scattering_inplace:arity", "Expected 3 inputs: (Y, eNodMat, uMat). No outputs.");
    ScatterArgs a;
    a.Y = arr(&prhs[0]);
    a.en = arr(&prhs[1]);
    a.um = arr(&prhs[2]);
    if (a.Y.nd != 2 || a.Y.d[1] != 3) fail(TG_ERR_SHAPE, Sh, "Y must be [numNodes x 3].");
    if (!is_float(a.Y.cls)) fail(TG_ERR_TYPE, Ty, "Y must be single or double gpuArray.");
    if (a.en.cls != TG_INT32) fail(TG_ERR_TYPE, Ty, "eNodMat must be gpuArray(int32).");
    if (a.en.nd != 2 || a.en.d[1] != 8) fail(TG_ERR_SHAPE, Sh, "eNodMat must be [Ne x 8] int32 (1-based).");
    a.Ne = a.en.d[0];
    a.Nn = a.Y.d[0];
    if (a.um.nd != 2 || a.um.d[0] != a.Ne || a.um.d[1] != 24) fail(TG_ERR_SHAPE, Sh, "uMat must be [Ne x 24] double.");
    if (a.um.cls != a.Y.cls) fail(TG_ERR_TYPE, Ty, "uMat must have the class of Y.");
    check_items(8 * a.Ne, Sh);
    return a;
}

}  // namespace
}  // namespace tg
