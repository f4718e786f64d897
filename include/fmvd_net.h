/*
 * Learned warm start for the multi-view RL (libNativeSPIMReconstructionCuda.so,
 * MI355X / gfx950 build).  No counterpart in the reference.
 *
 * A 6-layer 3-D residual conv net (32 channels, 3x3x3, zero padding) maps
 * the registered views of a volume to an estimate that seeds RL
 * (fmvd_deconvolve_volume_device_cuda_* takes it as d_init).  Input: two
 * channels, the mean of the even-index views and the mean of the odd-index
 * views, divided by a per-sample scale; output: y = conv6(h5) + b6 + (x0 +
 * x1)/2, multiplied back by the scale.  Weights and activations are
 * bfloat16, accumulation FP32 (the four 32->32 layers on MFMA), master
 * weights and Adam state FP32.  Flat parameter layout, layer by layer:
 * W1 [27][32][2], b1 [32], W2..W5 [27][32][32], b2..b5 [32], W6 [27][1][32],
 * b6 [1]; tap = (dz*3 + dy)*3 + dx reads voxel (z+dz-1, y+dy-1, x+dx-1).
 *
 * Errors go through the handler of setErrorHandler(); functions return
 * NULL / a negative value.  A network object is not reentrant.
 */
#ifndef FMVD_NET_H
#define FMVD_NET_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct fmvd_net; /* opaque */

size_t fmvd_net_param_count(void);
struct fmvd_net *fmvd_net_create(unsigned long long seed); /* deterministic init */
void fmvd_net_destroy(struct fmvd_net *net);
int fmvd_net_get_params(const struct fmvd_net *net, float *host);  /* fmvd_net_param_count() floats */
int fmvd_net_set_params(struct fmvd_net *net, const float *host);  /* also resets the Adam state */
int fmvd_net_get_grads(const struct fmvd_net *net, float *host);   /* gradients of the last backward */

/* d_x: bfloat16 input [N][D][H][W][2]; d_y: float [N][D][H][W] (scaled units). */
int fmvd_net_forward(struct fmvd_net *net, const void *d_x, int N, int D, int H, int W, float *d_y, void *stream);
/* Forward + gradients of mean((y - t)^2) over the batch (W <= 256);
 * *loss receives the loss (summed in FP64 on the GPU). */
int fmvd_net_backward(struct fmvd_net *net, const void *d_x, const float *d_t, int N, int D, int H, int W,
                      double *loss, void *stream);
/* One Adam step (lr; beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) with
 * the gradients of the last backward. */
int fmvd_net_adam(struct fmvd_net *net, float lr, void *stream);
/* Network input from views: d_views holds N*nviews device pointers (sample
 * major) to D*H*W uint16 volumes; d_inv_scales N floats (device). */
int fmvd_net_prep_input(const unsigned short *const *d_views, int nviews, int N, int D, int H, int W,
                        const float *d_inv_scales, void *d_x, void *stream);
/* Warm start of one volume: d_out (float D*H*W, device) = max(scale * net(views / scale), floor_value),
 * computed tile-wise (interior tiles of at most 128 x 224 x 224 voxels with a
 * 16-voxel halo inside the volume; identical to whole-volume inference).
 * d_views: nviews device pointers (device array). */
int fmvd_net_infer(struct fmvd_net *net, const unsigned short *const *d_views, int nviews, int D, int H, int W,
                   float scale, float floor_value, float *d_out, void *stream);
int fmvd_net_infer_tiled(struct fmvd_net *net, const unsigned short *const *d_views, int nviews, int D, int H, int W,
                         float scale, float floor_value, int tile_z, int tile_y, int tile_x, int halo, float *d_out,
                         void *stream);
/* Unit test of the MFMA operand layout: D (32x32 float) = A (32x16 bf16) * B (16x32 bf16), row-major. */
int fmvd_net_mfma_layout_test(const unsigned short *d_a, const unsigned short *d_b, float *d_d);

#ifdef __cplusplus
}
#endif

#endif /* FMVD_NET_H */
