// device.hpp - small RAII helpers over the HIP runtime used by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <set>
#include <utility>

namespace scrs {

// Device and pinned-host allocations made by the library so far (the
// observable of the zero-allocation steady state).
inline std::atomic<uint64_t> &alloc_counter() {
    static std::atomic<uint64_t> n{0};
    return n;
}

// Device allocation that only grows.
struct DevBuf {
    void *ptr = nullptr;
    size_t size = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        size = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= size) return hipSuccess;
        release();
        alloc_counter().fetch_add(1, std::memory_order_relaxed);
        hipError_t err = hipMalloc(&ptr, n);
        if (err != hipSuccess) {
            ptr = nullptr;
            return err;
        }
        size = n;
        return hipSuccess;
    }
    template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// Pinned host allocation that only grows.
struct HostBuf {
    void *ptr = nullptr;
    size_t size = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf() { release(); }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        size = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= size) return hipSuccess;
        release();
        alloc_counter().fetch_add(1, std::memory_order_relaxed);
        hipError_t err = hipHostMalloc(&ptr, n, hipHostMallocDefault);
        if (err != hipSuccess) {
            ptr = nullptr;
            return err;
        }
        size = n;
        return hipSuccess;
    }
    template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// Opt a kernel into `bytes` of dynamic LDS (needed above 64 KiB) on the
// current device, once per (kernel, device).
inline hipError_t lds_opt_in(const void *kernel, int bytes) {
    static std::mutex mu;
    static std::set<std::pair<const void *, int>> done;
    int dev = -1;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    std::lock_guard<std::mutex> lock(mu);
    if (done.count({kernel, dev})) return hipSuccess;
    err = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (err == hipSuccess) done.insert({kernel, dev});
    return err;
}

// Makes `device` current for the scope and restores the caller's device.
struct DeviceGuard {
    int prev = -1;
    bool changed = false;
    explicit DeviceGuard(int device) {
        if (hipGetDevice(&prev) == hipSuccess && prev != device) {
            changed = hipSetDevice(device) == hipSuccess;
        }
    }
    ~DeviceGuard() {
        if (changed) (void)hipSetDevice(prev);
    }
};

}  // namespace scrs
