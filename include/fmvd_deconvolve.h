/*
 * Plane-wise multi-view Richardson-Lucy deconvolution: C ABI of
 * libNativeSPIMReconstructionCuda.so (MI355X / gfx950 build).
 *
 * Drop-in for the reference interface declared in
 *   fmvd_deconvolve.h:9-15        (fmvd_psf_type, fmvd_malloc, fmvd_free)
 *   fmvd_deconvolve.impl.h:6-66   (datasource/datasink, plan, init/process/
 *                                  destroy, files front end)
 * of bene51/gpu_deconvolution.  The names keep the reference's "_cuda"
 * suffix so existing callers (the Fiji plugin through JNI, the CLI) link
 * unchanged; the functions get C linkage here (the reference compiled them
 * with C++ linkage).  The plan struct is opaque: callers only ever held a
 * pointer to it.
 *
 * Errors: every function reports failures once through the handler installed
 * with setErrorHandler() (fmvd_cuda_utils.h; default prints and exits, as the
 * reference does), releases what the failed call allocated and returns
 * (NULL from the initialisers, a negative value from the int-returning
 * extensions).
 */
#ifndef FMVD_DECONVOLVE_H
#define FMVD_DECONVOLVE_H

#include <stddef.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fmvd_deconvolve.h:9 */
enum fmvd_psf_type { independent, efficient_bayesian, optimization_1, optimization_2 };
typedef enum fmvd_psf_type fmvd_psf_type;

/* fmvd_deconvolve.h:11-15: page-locked host memory. */
void *fmvd_malloc(size_t size);
void fmvd_free(void *p);

/* fmvd_deconvolve.impl.h:6-8.  A datasource fills buffer[v] + offset with the
 * next plane of every view (dataW*dataH samples, x fastest) and returns 0 when
 * there is no more data.  A datasink consumes one result plane synchronously
 * (the buffer is reused after it returns).  Planes are requested and returned
 * in z order on the calling thread. */
typedef int (*datasource_t_16)(unsigned short **buffer, int offset);
typedef void (*datasink_t_16)(unsigned short *buffer);
typedef int (*datasource_t_8)(unsigned char **buffer, int offset);
typedef void (*datasink_t_8)(unsigned char *buffer);

struct fmvd_plan_cuda_16; /* opaque (fmvd_deconvolve.impl.h:10-37) */
struct fmvd_plan_cuda_8;

/* fmvd_deconvolve.impl.h:39-48 / fmvd_deconvolve.impl.cpp:130-203.
 * Normalises h_Kernel[v] in place (as the reference does), copies the
 * weights and retains the callbacks.  nstreams >= 1 is the minimum number of
 * planes kept in flight; the plan processes planes in batches of at least
 * that many (fmvd_set_batch_planes_cuda_*), sized from a 2 GiB device budget
 * by default. */
struct fmvd_plan_cuda_16 *fmvd_initialize_cuda_16(
    int dataH, int dataW, float const *const *h_Weights, float **h_Kernel, int kernelH,
    int kernelW, fmvd_psf_type iteration_type, int nViews, int nstreams,
    datasource_t_16 get_next_plane, datasink_t_16 return_next_plane);
struct fmvd_plan_cuda_8 *fmvd_initialize_cuda_8(
    int dataH, int dataW, float const *const *h_Weights, float **h_Kernel, int kernelH,
    int kernelW, fmvd_psf_type iteration_type, int nViews, int nstreams,
    datasource_t_8 get_next_plane, datasink_t_8 return_next_plane);

/* fmvd_deconvolve.impl.h:51-52 / fmvd_deconvolve.impl.cpp:205-300: pull planes
 * until the datasource returns 0, deconvolve each with `iterations` RL
 * iterations, push every result plane to the datasink in z order. */
void fmvd_deconvolve_planes_cuda_16(const struct fmvd_plan_cuda_16 *plan, int iterations);
void fmvd_deconvolve_planes_cuda_8(const struct fmvd_plan_cuda_8 *plan, int iterations);

/* fmvd_deconvolve.impl.h:54-55 / fmvd_deconvolve.impl.cpp:302-336 */
void fmvd_destroy_cuda_16(struct fmvd_plan_cuda_16 *plan);
void fmvd_destroy_cuda_8(struct fmvd_plan_cuda_8 *plan);

/* fmvd_deconvolve.impl.h:57-66 / fmvd_deconvolve_hdd.impl.cpp:36-76: raw
 * planes read from dataFiles[v] (dataD planes of dataW*dataH samples each),
 * results written to resultFile in z order. */
void fmvd_deconvolve_files_cuda_16(FILE **dataFiles, FILE *resultFile, int dataW, int dataH,
                                   int dataD, float **h_Weights, float **h_Kernel, int kernelH,
                                   int kernelW, fmvd_psf_type iteration_type, int nViews,
                                   int iterations);
void fmvd_deconvolve_files_cuda_8(FILE **dataFiles, FILE *resultFile, int dataW, int dataH,
                                  int dataD, float **h_Weights, float **h_Kernel, int kernelH,
                                  int kernelW, fmvd_psf_type iteration_type, int nViews,
                                  int iterations);

/* ------------------------------------------------------------------------
 * Extensions with no counterpart in the reference.
 * ---------------------------------------------------------------------- */

/* Receives the FP32 estimate of each plane (dataW*dataH floats, before the
 * truncation to samples) just before the datasink sees that plane.  For
 * testing and for callers that want the float result. */
typedef void (*fmvd_estimate_sink_t)(const float *estimate, int plane, void *user);
int fmvd_set_estimate_sink_cuda_16(struct fmvd_plan_cuda_16 *plan, fmvd_estimate_sink_t sink,
                                   void *user);
int fmvd_set_estimate_sink_cuda_8(struct fmvd_plan_cuda_8 *plan, fmvd_estimate_sink_t sink,
                                  void *user);

/* Maximum number of planes processed together (0 = automatic, from a device
 * memory budget).  Returns the value in effect. */
int fmvd_set_batch_planes_cuda_16(struct fmvd_plan_cuda_16 *plan, int planes);
int fmvd_set_batch_planes_cuda_8(struct fmvd_plan_cuda_8 *plan, int planes);

/* Planes whose whole iteration chain runs before the next planes start
 * (0 = automatic: the power of two of planes whose per-iteration working
 * set fits a 200 MiB cache budget, 8 at FFT 1536 x 1536 with 4 views).
 * Results do not depend on it.  No counterpart in the reference.  Returns
 * the value in effect. */
int fmvd_set_chunk_planes_cuda_16(struct fmvd_plan_cuda_16 *plan, int planes);
int fmvd_set_chunk_planes_cuda_8(struct fmvd_plan_cuda_8 *plan, int planes);

/* Bound on the planes pulled from the datasource ahead of the datasink
 * (0 = the batched pipeline above, the default).  With planes = L > 0,
 * fmvd_deconvolve_planes_cuda_* makes the reference's callback sequence
 * (fmvd_deconvolve.impl.cpp:216-245,288-294): L source calls, then one
 * sink call followed by one source call per plane until a source call
 * returns 0, then the sinks of the planes still held; with L = nstreams this
 * is exactly the reference's order.  Planes then run in batches of L with no
 * overlap of host callbacks and device work.  The JNI interactive session
 * uses L = nStreams.  No counterpart in the reference.  Returns the value in
 * effect. */
int fmvd_set_lookahead_cuda_16(struct fmvd_plan_cuda_16 *plan, int planes);
int fmvd_set_lookahead_cuda_8(struct fmvd_plan_cuda_8 *plan, int planes);

/* Replaces the plan's view weights with nViews device weight images
 * (d_weights: host array of nViews device pointers to dataH*dataW floats),
 * normalised over views on the device with the float operations of
 * prepareWeights (fmvd_deconvolve.impl.cpp:86-124): bit-identical to
 * passing the same images as h_Weights at init.  Enqueued on `stream`; runs
 * enqueued later on the same stream use them.  fmvd_initialize_cuda_* also
 * accepts h_Weights = NULL (every view weighs 1) for plans whose weights come
 * from here.  No counterpart in the reference.  Returns 0 on success. */
int fmvd_set_weights_device_cuda_16(struct fmvd_plan_cuda_16 *plan, const float *const *d_weights, void *stream);
int fmvd_set_weights_device_cuda_8(struct fmvd_plan_cuda_8 *plan, const float *const *d_weights, void *stream);

/* GPU-resident transform -> weights -> RL chain (no host round trip): view
 * v's acquired volume d_acquired[v] (w x h x d samples, device) is
 * resampled with h_inverse[v] (12 floats, as transform_cuda_*) into
 * d_registered[v] (nPlanes planes of dataW x dataH, device, caller-owned);
 * each view's border mask (fmvd_transform.cu:19-67, plane z = nPlanes/2,
 * `border`, `zspacing`) becomes its weight image, normalised as
 * fmvd_set_weights_device_cuda_*; then RL as
 * fmvd_deconvolve_volume_device_cuda_* into d_out / d_estimate.  The
 * reference chains these through files (transform_cuda_* writes the volume and
 * the mask, the deconvolution reads them back).  Enqueued on `stream`;
 * returns 0 on success. */
int fmvd_reconstruct_device_cuda_16(struct fmvd_plan_cuda_16 *plan, const unsigned short *const *d_acquired,
                                    int w, int h, int d, const float *const *h_inverse, int border,
                                    float zspacing, unsigned short *const *d_registered, int nPlanes,
                                    int iterations, unsigned short *d_out, float *d_estimate, void *stream);
int fmvd_reconstruct_device_cuda_8(struct fmvd_plan_cuda_8 *plan, const unsigned char *const *d_acquired,
                                   int w, int h, int d, const float *const *h_inverse, int border,
                                   float zspacing, unsigned char *const *d_registered, int nPlanes,
                                   int iterations, unsigned char *d_out, float *d_estimate, void *stream);

/* GPU-resident volume deconvolution: d_views[v] (host array of nViews device
 * pointers) each point to nPlanes planes of dataW*dataH samples in device
 * memory.  Results go to d_out (nPlanes planes of samples) and/or d_estimate
 * (nPlanes float planes); either may be NULL.  d_init (nPlanes float planes,
 * may be NULL) replaces the constant start estimate 10*nViews.  All work is
 * enqueued on `stream` (a hipStream_t, NULL = the null stream) and the call
 * returns without synchronising.  Returns 0 on success. */
int fmvd_deconvolve_volume_device_cuda_16(const struct fmvd_plan_cuda_16 *plan,
                                          const unsigned short *const *d_views, int nPlanes,
                                          int iterations, const float *d_init,
                                          unsigned short *d_out, float *d_estimate, void *stream);
int fmvd_deconvolve_volume_device_cuda_8(const struct fmvd_plan_cuda_8 *plan,
                                         const unsigned char *const *d_views, int nPlanes,
                                         int iterations, const float *d_init, unsigned char *d_out,
                                         float *d_estimate, void *stream);

/* Allocation-free variant of fmvd_deconvolve_volume_device_cuda_* (same
 * arguments, same results bit for bit).  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  It calls no hipMalloc/hipFree or other
 * allocating routine: the work buffers (spectrum and estimate of each
 * plane) live in a fixed 128 MiB device array that every GPU the library
 * runs on holds from module load until the process ends, and the planes
 * run in batches of as many as it holds (2 at FFT 2560 x 2560, 6 at
 * 1536 x 1536).  Plans whose single plane needs more (FFT extents above
 * about 4096 x 4096) are rejected with an error.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU race; the array
 * needs no clearing (each plane's buffers are written before they are
 * read).  The plan must have been created beforehand. */
int fmvd_deconvolve_volume_device_noalloc_cuda_16(const struct fmvd_plan_cuda_16 *plan,
                                                  const unsigned short *const *d_views, int nPlanes,
                                                  int iterations, const float *d_init, unsigned short *d_out,
                                                  float *d_estimate, void *stream);
int fmvd_deconvolve_volume_device_noalloc_cuda_8(const struct fmvd_plan_cuda_8 *plan,
                                                 const unsigned char *const *d_views, int nPlanes,
                                                 int iterations, const float *d_init, unsigned char *d_out,
                                                 float *d_estimate, void *stream);

/* Autotuner: times every compiled FFT-plan variant of each device kernel
 * kind (0 column pass, 1 ratio row pass, 2 update row pass, 3 warm-start
 * ingest, 4 egress) on this plan's buffers over `planes` planes (0 = 32)
 * and keeps the fastest.  chosen[5] (may be NULL) receives the picks,
 * times_ms[15] (may be NULL) the mean launch time of kind k, variant v at
 * [3k + v] (negative where that variant does not exist for these extents).
 * Deterministic results: every variant computes the same transform.  Not
 * safe to call while the plan is running. */
int fmvd_autotune_cuda_16(struct fmvd_plan_cuda_16 *plan, int planes, int *chosen, float *times_ms);
int fmvd_autotune_cuda_8(struct fmvd_plan_cuda_8 *plan, int planes, int *chosen, float *times_ms);
/* Autotuner cache file (no counterpart in the reference).  With a path set,
 * fmvd_autotune_cuda_* first looks the plan's configuration up in that text
 * file (device, FFT extents, sample type, kernel-hat form, planes per launch)
 * and applies the recorded picks without timing; otherwise it times every
 * variant and appends a line with the picks and times.  NULL or "" turns the
 * cache off; until this is called, the environment variable
 * FMVD_AUTOTUNE_CACHE sets the path.  fmvd_autotune_cache_selftest checks the
 * line format (a round trip, and rejection of malformed lines) without a
 * device and returns 0 when it holds. */
void fmvd_set_autotune_cache(const char *path);
int fmvd_autotune_cache_selftest(void);

/* Sets the variants directly (variants[5] as above; the egress kernel has no
 * FFT plan, so variants[4] must be 0).  Returns -1 if a value is out of range
 * for this plan's extents. */
int fmvd_set_variants_cuda_16(struct fmvd_plan_cuda_16 *plan, const int *variants);
int fmvd_set_variants_cuda_8(struct fmvd_plan_cuda_8 *plan, const int *variants);

/* Plan geometry: fills dims[0..6] = dataH, dataW, fftH, fftW, nViews, batch
 * planes, planes per cache-resident chunk. */
int fmvd_plan_info_cuda_16(const struct fmvd_plan_cuda_16 *plan, int *dims);
int fmvd_plan_info_cuda_8(const struct fmvd_plan_cuda_8 *plan, int *dims);

/* FFT extent rule of the reference (fmvd_deconvolve.cpp:28-51) and whether
 * the device kernels support the resulting extents. */
int fmvd_snap_transform_size(int n);
int fmvd_fft_extent_supported(int fftH, int fftW);

/* Per-kernel GPU timing with HIP events recorded around each launch on the
 * launch stream (off by default).  kind: 0 column pass, 1 ratio row pass,
 * 2 update row pass, 3 ingest, 4 egress, 5 transform.  fmvd_kernel_timing_read
 * synchronises on the recorded events and returns the number of launches and
 * their summed duration in milliseconds. */
void fmvd_kernel_timing_enable(int on);
void fmvd_kernel_timing_reset(void);
int fmvd_kernel_timing_read(int kind, double *total_ms, long long *launches);

/* Extension (no reference counterpart): RL on a GPU-resident volume, as
 * fmvd_deconvolve_volume_device_cuda_*, that records after ingest and after
 * every iteration k the relative L2 distance ||E_k - ref|| / ||ref|| of the
 * estimate over the data region to d_ref (float, nPlanes*dataH*dataW,
 * device), summed in FP64 on the device.  errs receives iterations + 1
 * values (errs[0] for the start estimate).  Synchronous. */
int fmvd_convergence_cuda_16(const struct fmvd_plan_cuda_16 *plan, const unsigned short *const *d_views,
                             int nPlanes, int iterations, const float *d_init, const float *d_ref, double *errs,
                             void *stream);
int fmvd_convergence_cuda_8(const struct fmvd_plan_cuda_8 *plan, const unsigned char *const *d_views, int nPlanes,
                            int iterations, const float *d_init, const float *d_ref, double *errs, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FMVD_DECONVOLVE_H */
