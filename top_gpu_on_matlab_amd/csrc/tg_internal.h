// Internal helpers shared by the library's translation units: thread-local error state,
// HIP error checking, RAII device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/topgpu.h"

namespace tg {

// Raised inside the library; converted to a status code + thread-local message at the C ABI.
struct Error : std::runtime_error {
    int code;
    std::string id;
    Error(int c, std::string i, const std::string &msg) : std::runtime_error(msg), code(c), id(std::move(i)) {}
};

void set_error(int code, const char *id, const char *msg);
void clear_error();

[[noreturn]] inline void fail(int code, const char *id, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, id, buf);
}

#define TG_HIP(call, id)                                                                        \
    do {                                                                                        \
        hipError_t e__ = (call);                                                                \
        if (e__ != hipSuccess)                                                                  \
            ::tg::fail(TG_ERR_HIP, id, "HIP error %d: %s (%s:%d)", (int)e__, hipGetErrorString(e__), \
                       __FILE__, __LINE__);                                                     \
    } while (0)

#define TG_LAUNCH_CHECK(id) TG_HIP(hipGetLastError(), id)

// Run `body` and convert exceptions to status codes (C ABI boundary).
template <class F>
int guard(const char *default_id, F &&body) {
    clear_error();
    try {
        body();
        return TG_OK;
    } catch (const Error &e) {
        set_error(e.code, e.id.c_str(), e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error(TG_ERR_HIP, default_id, "host allocation failed");
        return TG_ERR_HIP;
    } catch (const std::exception &e) {
        set_error(TG_ERR_ARGS, default_id, e.what());
        return TG_ERR_ARGS;
    }
}

// Owning device allocation.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) TG_HIP(hipMalloc((void **)&p, count * sizeof(T)), "topgpu:alloc");
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // The copy has completed when this returns, so the source may be a temporary (uploads are set-up
    // work, outside the solver loops).
    void upload(const T *h, size_t count, hipStream_t s) {
        if (count > n) alloc(count);
        if (count) {
            TG_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s), "topgpu:upload");
            TG_HIP(hipStreamSynchronize(s), "topgpu:upload");
        }
    }
    void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
    void zero(hipStream_t s) {
        if (n) TG_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s), "topgpu:memset");
    }
    operator T *() const { return p; }
};

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace tg
