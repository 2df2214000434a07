// Shared host/device helpers of the HIP library: error handling, device buffers, launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
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
    // Work queued on the library's (non-blocking) stream may still read or write the old block:
    // wait for the device before returning it to the allocator.
    void release() {
        if (p) {
            (void)hipDeviceSynchronize();
            (void)hipFree(p);
        }
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

// Host <-> device copies through page-locked blocks owned by the library (hipHostMalloc): the runtime
// never has to lock the caller's pageable memory for a DMA.  Each copy leases a block from a pool for
// its duration, so copies issued by different host threads (on different streams) run concurrently.
// Each chunk completes before the block is reused.
struct Pinned {
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::vector<void*>& pool() {
        static std::vector<void*> v;
        return v;
    }
    static size_t chunk() { return (size_t)8 << 20; }
    static void* acquire() {
        {
            std::lock_guard<std::mutex> lk(mu());
            if (!pool().empty()) {
                void* p = pool().back();
                pool().pop_back();
                return p;
            }
        }
        void* p = nullptr;
        SW_HIP(hipHostMalloc(&p, chunk(), hipHostMallocDefault));
        return p;
    }
    static void release(void* p) {
        std::lock_guard<std::mutex> lk(mu());
        pool().push_back(p);
    }
};
struct PinLease {
    void* p;
    PinLease() : p(Pinned::acquire()) {}
    ~PinLease() { Pinned::release(p); }
    PinLease(const PinLease&) = delete;
    PinLease& operator=(const PinLease&) = delete;
};

inline void copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    PinLease pin;
    for (size_t off = 0; off < bytes; off += Pinned::chunk()) {
        size_t b = std::min(Pinned::chunk(), bytes - off);
        std::memcpy(pin.p, (const char*)src + off, b);
        SW_HIP(hipMemcpyAsync((char*)dst + off, pin.p, b, hipMemcpyHostToDevice, st));
        SW_HIP(hipStreamSynchronize(st));
    }
}

inline void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
    PinLease pin;
    for (size_t off = 0; off < bytes; off += Pinned::chunk()) {
        size_t b = std::min(Pinned::chunk(), bytes - off);
        SW_HIP(hipMemcpyAsync(pin.p, (const char*)src + off, b, hipMemcpyDeviceToHost, st));
        SW_HIP(hipStreamSynchronize(st));
        std::memcpy((char*)dst + off, pin.p, b);
    }
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
