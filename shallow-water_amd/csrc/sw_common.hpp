// Shared host/device helpers of the HIP library: error handling, device buffers, launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "shallow_water.h"

namespace sw {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SW_HIP(call)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw ::sw::Error(SW_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + " at " +   \
                                              __FILE__ + ":" + std::to_string(__LINE__));                \
    } while (0)

#define SW_CHECK(cond, code, msg)                          \
    do {                                                   \
        if (!(cond)) throw ::sw::Error((code), (msg));     \
    } while (0)

// Owning device allocation.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        if (count) SW_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        release();
        p = o.p;
        n = o.n;
        o.p = nullptr;
        o.n = 0;
        return *this;
    }
};

template <class F>
__global__ void __launch_bounds__(256) k_each(size_t n, F f) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f(i);
}

template <class F>
inline void each(hipStream_t st, size_t n, F f) {
    if (!n) return;
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_each<F>, dim3(blocks), dim3(256), 0, st, n, f);
    SW_HIP(hipGetLastError());
}

// True when p is device (or managed) memory the GPU can address directly.
inline bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace sw
