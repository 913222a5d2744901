// device.hpp -- HIP error check and an owning device buffer (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.hpp"

#define GFX_HIP(x)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            throw ::gfx::Error(GFX_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace gfx {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error(GFX_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                           " bytes failed: " + hipGetErrorString(e));
        }
        n = count;
    }
    void upload(const std::vector<T>& h) {
        alloc(h.size());
        if (!h.empty()) GFX_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    void zero(hipStream_t st) {
        if (p) GFX_HIP(hipMemsetAsync(p, 0, n * sizeof(T), st));
    }
};

}  // namespace gfx
