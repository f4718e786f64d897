/*
 * Error handler and device helpers of libNativeSPIMReconstructionCuda.so
 * (MI355X / gfx950 build).  Drop-in for fmvd_cuda_utils.h:4-25 of
 * bene51/gpu_deconvolution, with C linkage.
 */
#ifndef FMVD_CUDA_UTILS_H
#define FMVD_CUDA_UTILS_H

#ifdef __cplusplus
extern "C" {
#endif

/* fmvd_cuda_utils.h:4-5 / fmvd_cuda_utils.cpp:22-27.  Installs the
 * process-global error callback; handler(param, message) is called once per
 * failed API call.  NULL restores the default (print, exit(EXIT_FAILURE)). */
void setErrorHandler(void (*handler)(void *, const char *), void *param);

/* fmvd_cuda_utils.h:7-8 */
int iDivUp(int a, int b);

/* fmvd_cuda_utils.h:17-25 / fmvd_cuda_utils.cpp:64-85 */
int getNumCUDADevices(void);
/* copies the device name into name (256 bytes, NUL-terminated) */
void getCudaDeviceName(int dev, char *name);
void setCudaDevice(int dev);

/* Extension: message of the most recent error reported on this thread
 * (empty if none).  Returns the message length. */
int fmvd_last_error(char *buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* FMVD_CUDA_UTILS_H */
