// Device-memory and stream helpers (fmvd_device.h).
#include <hip/hip_runtime.h>

#include "../../include/fmvd_device.h"
#include "hip_check.hpp"

extern "C" {

void* fmvd_device_malloc(size_t bytes)
{
    FMVD_API_GUARD(nullptr, {
        void* p = nullptr;
        FMVD_HIP(hipMalloc(&p, bytes ? bytes : 1));
        return p;
    })
}

void fmvd_device_free(void* p)
{
    if (p) (void)hipFree(p);
}

int fmvd_memcpy(void* dst, const void* src, size_t bytes, int kind)
{
    FMVD_API_GUARD(-1, {
        hipMemcpyKind k;
        switch (kind) {
        case 1: k = hipMemcpyHostToDevice; break;
        case 2: k = hipMemcpyDeviceToHost; break;
        case 3: k = hipMemcpyDeviceToDevice; break;
        default: fmvd::throw_msg("fmvd_memcpy: kind must be 1, 2 or 3");
        }
        if (bytes) FMVD_HIP(hipMemcpy(dst, src, bytes, k));
        return 0;
    })
}

int fmvd_memset_device(void* p, int value, size_t bytes)
{
    FMVD_API_GUARD(-1, {
        if (bytes) FMVD_HIP(hipMemset(p, value, bytes));
        return 0;
    })
}

int fmvd_device_synchronize(void)
{
    FMVD_API_GUARD(-1, {
        FMVD_HIP(hipDeviceSynchronize());
        return 0;
    })
}

void* fmvd_stream_create(void)
{
    FMVD_API_GUARD(nullptr, {
        hipStream_t s = nullptr;
        FMVD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        return (void*)s;
    })
}

int fmvd_stream_destroy(void* stream)
{
    FMVD_API_GUARD(-1, {
        if (stream) FMVD_HIP(hipStreamDestroy((hipStream_t)stream));
        return 0;
    })
}

int fmvd_stream_synchronize(void* stream)
{
    FMVD_API_GUARD(-1, {
        FMVD_HIP(hipStreamSynchronize((hipStream_t)stream));
        return 0;
    })
}

int fmvd_device_memory(size_t* free_bytes, size_t* total_bytes)
{
    FMVD_API_GUARD(-1, {
        size_t f = 0, t = 0;
        FMVD_HIP(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
        return 0;
    })
}

}  // extern "C"
