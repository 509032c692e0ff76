// Scattering_inplace_noalloc: Scattering_inplace for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Same arguments, checks, error identifiers and results as Scattering_inplace (Scattering_inplace.cu:
// 81-146, kernel :45-78): Y(eNodMat(e,j), c) += uMat(e, c+3j), summed per node in ascending entry order.
// The incidence is sorted on every call into `scratch`, a static device array of
// NOALLOC_SCRATCH_BYTES; nothing is cached and nothing is allocated (no hipMalloc, no hipFree, no
// allocating library routine: the sort works in caller-provided storage).
//
// Scratch layout for N = 8 Ne entries: keys[2][N], entry ids[2][N] (uint32, double-buffered, each
// segment 256-byte aligned), then the sort's temporary storage.  Every byte of it is written before it is
// read in the same call, so nothing needs clearing.  All calls on one GPU share that GPU's array:
// calls that overlap on one GPU (two streams, or two host threads) race and must be serialised by the
// caller.  The entry point synchronises the stream before returning, so consecutive calls on one
// stream are ordered.
#include <mutex>

#include "scatter.h"

namespace tg {
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Per-device address of `scratch`.  In host code the array's name is a host-side stand-in, not a device
// address, so every use below goes through the pointer hipGetSymbolAddress returns for the current
// device (each GPU holds its own copy at its own address).
struct ScratchAddr {
    std::mutex mu;
    void *ptr[64] = {};
};

ScratchAddr &scratch_addr() {
    static ScratchAddr s;
    return s;
}

void *scratch_ptr(const char *id) {
    int dev = 0;
    TG_HIP(hipGetDevice(&dev), id);
    if (dev < 0 || dev >= 64) fail(TG_ERR_HIP, id, "device ordinal %d out of range", dev);
    ScratchAddr &S = scratch_addr();
    std::lock_guard<std::mutex> lock(S.mu);
    if (!S.ptr[dev]) {
        void *p = nullptr;
        TG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), id);
        if (!p) fail(TG_ERR_HIP, id, "no device address for the scratch array");
        S.ptr[dev] = p;
    }
    return S.ptr[dev];
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace
}  // namespace tg

using namespace tg;

extern "C" {

int Scattering_inplace_noalloc(int, tg_array *, int nrhs, const tg_array *prhs, void *stream) {
    return guard("Scattering_inplace:arity", [&] {
        const char *Cu = "Scattering_inplace:cuda", *Sh = "Scattering_inplace:shape";
        const ScatterArgs a = scatter_args(nrhs, prhs);
        const int64_t N = 8 * a.Ne;
        const size_t kv = align256((size_t)N * sizeof(uint32_t));
        // the size limit is checked before any device call (the sort's storage is checked below)
        if (4 * kv > NOALLOC_SCRATCH_BYTES)
            fail(TG_ERR_SHAPE, Sh,
                 "Ne = %lld exceeds the allocation-free scratch (%zu bytes hold 16 x 8 Ne bytes of keys and ids "
                 "plus the sort's storage)",
                 (long long)a.Ne, NOALLOC_SCRATCH_BYTES);
        hipStream_t st = (hipStream_t)stream;
        if (a.Ne > 0 && a.Nn > 0) {
            unsigned char *base = (unsigned char *)scratch_ptr(Cu);
            uint32_t *k0 = (uint32_t *)base, *k1 = (uint32_t *)(base + kv), *v0 = (uint32_t *)(base + 2 * kv),
                     *v1 = (uint32_t *)(base + 3 * kv);
            hipcub::DoubleBuffer<uint32_t> kb(k0, k1), vb(v0, v1);
            size_t tmp = 0;
            sort_pairs_db(kb, vb, N, a.Nn, nullptr, tmp, st, Cu);  // size query only
            if (4 * kv + tmp > NOALLOC_SCRATCH_BYTES)
                fail(TG_ERR_SHAPE, Sh,
                     "Ne = %lld exceeds the allocation-free scratch (%zu bytes; keys and ids take %zu, the sort %zu)",
                     (long long)a.Ne, NOALLOC_SCRATCH_BYTES, 4 * kv, tmp);
            k_make_keys_enod<<<grid_for(N, 256), 256, 0, st>>>((const int32_t *)a.en.p, a.Ne, a.Nn, k0, v0);
            TG_LAUNCH_CHECK(Cu);
            sort_pairs_db(kb, vb, N, a.Nn, base + 4 * kv, tmp, st, Cu);
            if (a.Y.cls == TG_DOUBLE)
                k_seg_scatter_enod<double><<<grid_for(a.Nn, 256), 256, 0, st>>>(
                    (double *)a.Y.p, a.Nn, kb.Current(), vb.Current(), N, (const double *)a.um.p, a.Ne);
            else
                k_seg_scatter_enod<float><<<grid_for(a.Nn, 256), 256, 0, st>>>(
                    (float *)a.Y.p, a.Nn, kb.Current(), vb.Current(), N, (const float *)a.um.p, a.Ne);
            TG_LAUNCH_CHECK(Cu);
        }
        TG_HIP(hipStreamSynchronize(st), Cu);
    });
}

}  // extern "C"
