/*
 * Device-memory and stream helpers of libNativeSPIMReconstructionCuda.so.
 * No counterpart in the reference (its device memory never left the
 * library).  They exist for callers of the GPU-resident entry points
 * (fmvd_deconvolve_volume_device_cuda_*, fmvd_transform_device_*) that have
 * no HIP runtime of their own: buffers and streams made here belong to the
 * same runtime instance as the kernels that use them.  All return 0 on
 * success (NULL pointers on failure) and report errors through the handler
 * of setErrorHandler().
 */
#ifndef FMVD_DEVICE_H
#define FMVD_DEVICE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

void *fmvd_device_malloc(size_t bytes);
void fmvd_device_free(void *p);
/* kind: 1 host->device, 2 device->host, 3 device->device (synchronous) */
int fmvd_memcpy(void *dst, const void *src, size_t bytes, int kind);
int fmvd_memset_device(void *p, int value, size_t bytes);
int fmvd_device_synchronize(void);
void *fmvd_stream_create(void);
int fmvd_stream_destroy(void *stream);
int fmvd_stream_synchronize(void *stream);
/* Free and total device memory in bytes of the current device. */
int fmvd_device_memory(size_t *free_bytes, size_t *total_bytes);

#ifdef __cplusplus
}
#endif

#endif /* FMVD_DEVICE_H */
