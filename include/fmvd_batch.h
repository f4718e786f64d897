/*
 * Batched plane-wise multi-view RL over many independent volumes of uneven
 * size (libNativeSPIMReconstructionCuda.so, MI355X / gfx950 build).
 *
 * No counterpart in the reference, which reconstructs one volume per plan
 * (fmvd_deconvolve.impl.h:39-55).  Each volume is prepared and iterated
 * exactly as a single-volume plan does it (same PSF preparation, kernel
 * spectra, weight normalisation, E0 = 10 * nViews and update order), so a
 * batched volume's result equals that plan's fmvd_deconvolve_volume_device_*
 * result for the same FFT plan variants.  Volumes are grouped by padded FFT
 * extent; each class is cut into cache-resident chunks (whole volumes or
 * plane ranges of one, so that an iteration's working set fits a 200 MiB
 * budget, FMVD_CHUNK_BYTES), which run all their iterations in turn.  Per
 * iteration of a chunk, one launch per (view index, pass) covers every plane
 * of the chunk, and the iteration is captured once as a hipGraph and
 * replayed.
 *
 * Errors are reported through the handler installed with setErrorHandler()
 * (fmvd_cuda_utils.h); functions return NULL or a negative value.
 * Not reentrant per batch object.  Supported extents: padded FFT heights of
 * 256, 512, 1024 or 1536 and widths of 256, 512, 1024 or 1536.
 */
#ifndef FMVD_BATCH_H
#define FMVD_BATCH_H

#include "fmvd_deconvolve.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One reconstruction.  The device pointers stay owned by the caller and must
 * remain valid while the batch object exists. */
struct fmvd_batch_volume {
    int dataH, dataW, nplanes, nViews;  /* nViews <= 8 */
    int kernelH, kernelW;
    fmvd_psf_type psf_type;
    float const *const *h_weights;  /* nViews host arrays of dataH*dataW */
    float **h_kernels;              /* nViews host arrays of kernelH*kernelW, normalised in place */
    const void *const *d_views;     /* nViews device volumes, nplanes*dataH*dataW samples, plane-major */
    void *d_out;                    /* device samples (same shape) or NULL */
    float *d_est;                   /* device float estimate (same shape) or NULL */
};

struct fmvd_batch_cuda_16; /* opaque */
struct fmvd_batch_cuda_8;

struct fmvd_batch_cuda_16 *fmvd_batch_create_cuda_16(int nvolumes, const struct fmvd_batch_volume *volumes);
struct fmvd_batch_cuda_8 *fmvd_batch_create_cuda_8(int nvolumes, const struct fmvd_batch_volume *volumes);

/* Times every compiled FFT-plan variant of the column, ratio and update
 * passes on each extent class and keeps the fastest; chosen (optional)
 * receives 3 ints per class (column, ratio, update). */
int fmvd_batch_autotune_cuda_16(struct fmvd_batch_cuda_16 *batch, int *chosen);
int fmvd_batch_autotune_cuda_8(struct fmvd_batch_cuda_8 *batch, int *chosen);

/* Runs `iterations` RL iterations from E0 = 10 * nViews on every volume and
 * writes d_out / d_est, enqueued on `stream` (hipStream_t; NULL = default
 * stream).  use_graph != 0 replays one captured hipGraph per chunk and
 * iteration (the graphs are captured on first use and kept until the
 * variants change). */
int fmvd_batch_run_cuda_16(struct fmvd_batch_cuda_16 *batch, int iterations, int use_graph, void *stream);
int fmvd_batch_run_cuda_8(struct fmvd_batch_cuda_8 *batch, int iterations, int use_graph, void *stream);

/* nclasses; extents[2*c] = fftH, extents[2*c+1] = fftW; planes[c] = planes of class c
 * (any output may be NULL; arrays sized for the class count). */
int fmvd_batch_info_cuda_16(const struct fmvd_batch_cuda_16 *batch, int *nclasses, int *extents, int *planes);
int fmvd_batch_info_cuda_8(const struct fmvd_batch_cuda_8 *batch, int *nclasses, int *extents, int *planes);

void fmvd_batch_destroy_cuda_16(struct fmvd_batch_cuda_16 *batch);
void fmvd_batch_destroy_cuda_8(struct fmvd_batch_cuda_8 *batch);

#ifdef __cplusplus
}
#endif

#endif /* FMVD_BATCH_H */
