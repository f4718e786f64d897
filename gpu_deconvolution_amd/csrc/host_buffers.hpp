// RAII wrappers for device / pinned host memory, streams and events used by
// the host engines.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "hip_check.hpp"

namespace fmvd {

// ------------------------------------------------------------ RAII buffers
struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    ~DeviceBuffer() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n)
    {
        if (n <= bytes && p) return;
        release();
        if (n == 0) return;
        FMVD_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct PinnedBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() { release(); }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n)
    {
        if (n <= bytes && p) return;
        release();
        if (n == 0) return;
        FMVD_HIP(hipHostMalloc(&p, n, hipHostMallocDefault));
        bytes = n;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { FMVD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Stream()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

struct Event {
    hipEvent_t e = nullptr;
    Event() { FMVD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
    ~Event()
    {
        if (e) (void)hipEventDestroy(e);
    }
};

template <typename T> inline void upload(DeviceBuffer& b, const std::vector<T>& v)
{
    b.alloc(sizeof(T) * v.size());
    if (!v.empty()) FMVD_HIP(hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
}

}  // namespace fmvd
