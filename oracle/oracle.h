/*
 * CPU oracle for the multi-view SPIM reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker (or as the timed CPU baseline).
 *
 * It restates the reference algorithm (bene51/gpu_deconvolution) in plain C:
 * every function cites the reference file:line it follows.  The RL core and
 * the complex FFT are compiled twice from oracle_real.inc: once in double
 * (suffix _f64, the oracle proper) and once in float (suffix _f32, the FP32
 * twin used only as a CPU baseline).
 *
 * Conventions: planes are row-major [y][x] (x fastest); volumes are [z][y][x];
 * complex arrays are interleaved (re, im).  Sample data is passed as uint16
 * for both bit depths (bits = 8 or 16 selects the clamp value 2^bits-1).
 *
 * Parity pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_snap_transform_size  closed-form table (incl. 512*k, k=3,5,7)
 *   orc_fft_c2c_*            brute-force O(N^2) DFT and numpy.fft (library)
 *   orc_convolve_single_plane brute-force over the explicitly built padded torus,
 *                            delta-kernel identity, numpy FFT convolution
 *   orc_kernel_hat           flip index map, elementwise pow, normalisation sum,
 *                            numpy-built compound kernels (types 1 and 2)
 *   orc_normalized_weights   closed form, float sequential sum (golden fixture)
 *   orc_rl_torus_*           KATs: uniform-data scalar recurrence, fixed point,
 *                            flux conservation, zero-weight pixels, transpose
 *                            symmetry, numpy FFT-convolution re-implementation
 *   orc_rl_planes_*          golden fixtures written by numpy (tests/golden)
 *   orc_transform            identity/+1 quirk, integer shifts, half-voxel means,
 *                            brute force trilinear definition
 *   orc_mask                 closed form at known points
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fmvd_deconvolve.cpp:22-51 */
int orc_ialign_up(int a, int b);
int orc_snap_transform_size(int n);

/* Pad map shared by data, weights and PSF-prep padding
 * (fmvd_deconvolve_cuda.cu:96-112, fmvd_deconvolve_cuda.impl.cu:39-59):
 * i < n -> i ; n <= i < n+half -> n-1 ; otherwise 0. */
int orc_pad_index(int i, int n, int half);

/* Unnormalised complex FFT, forward sign -1, inverse sign +1. */
int orc_fft_c2c_f64(double *data, int n, int inverse);
int orc_fft_c2c_f32(float *data, int n, int inverse);
/* 2D complex FFT over an h x w interleaved complex array (rows then columns). */
int orc_fft2_c2c_f64(double *data, int h, int w, int inverse);

/* fmvd_deconvolve.cpp:53-68: sum in double, divide each float by (float)sum. */
void orc_normalize_f32(float *k, int len);
/* fmvd_deconvolve.cpp:70-79 */
void orc_flip_kernel_f64(const double *in, double *out, int kw, int kh);
/* fmvd_deconvolve.cpp:90-137: clamp-padded circular convolution on the
 * snap-sized torus, cropped back to dw x dh; computed directly in FP64. */
void orc_convolve_single_plane(double *data, int dw, int dh,
                               const double *kernel, int kw, int kh);
/* fmvd_deconvolve.cpp:139-204 (computeKernel2), evaluated in FP64 from the
 * RAW (un-normalised) kernels, then normalised.  kernels: V x kh x kw. */
int orc_kernel_hat(const float *kernels, int kw, int kh, int nviews,
                   int psf_type, double *out);

/* fmvd_deconvolve.impl.cpp:86-124 + fmvd_deconvolve_cuda.cu:160-177 with
 * reference float semantics: sum = 0f; sum += w_v in view order; w_v /= sum
 * where sum > 0.  Evaluated on the unpadded grid (padding commutes). */
void orc_normalized_weights(const float *w, int dw, int dh, int nviews,
                            float *out);

/* RL iterations on the padded torus (fmvd_deconvolve.impl.cpp:262-273):
 * for it, for v: Q = E (*) kp_v ; R = Q>1 ? Dp/Q : Q ; C = R (*) kph_v ;
 * E = E + Wp*(E*C - E).  All arrays h x w; per-view arrays stacked V x h x w.
 * kp / kph are the padded (wrapped, centred at (0,0)) kernels. */
int orc_rl_torus_f64(int h, int w, int nviews, const double *Dp,
                     const double *Wp, const double *kp, const double *kph,
                     int iterations, double *E);
int orc_rl_torus_f32(int h, int w, int nviews, const float *Dp,
                     const float *Wp, const float *kp, const float *kph,
                     int iterations, float *E);

/* Full per-plane reconstruction (fmvd_deconvolve.impl.cpp:130-300 +
 * fmvd_deconvolve_cuda.impl.cu:14-157).
 *   data     : nplanes x V x dh x dw samples
 *   weights  : V x dh x dw raw weights
 *   kernels  : V x kh x kw raw PSFs, normalised IN PLACE (as the reference)
 *   init     : optional nplanes x dh x dw initial estimate; NULL -> E0 = 10*V
 *              (the padded torus is then filled through the data pad map)
 *   est_out  : optional nplanes x dh x dw float estimate before truncation
 *   out      : optional nplanes x dh x dw samples (truncated, clamped)
 * Planes are independent; nthreads > 1 parallelises over planes. */
int orc_rl_planes_f64(int dw, int dh, int nplanes, int nviews, int bits,
                      const uint16_t *data, const float *weights,
                      float *kernels, int kw, int kh, int psf_type,
                      int iterations, const double *init, double *est_out,
                      uint16_t *out, int nthreads);
int orc_rl_planes_f32(int dw, int dh, int nplanes, int nviews, int bits,
                      const uint16_t *data, const float *weights,
                      float *kernels, int kw, int kh, int psf_type,
                      int iterations, const double *init, double *est_out,
                      uint16_t *out, int nthreads);

/* Build the padded per-view operands exactly as the reference does, for use
 * with orc_rl_torus_*: Dp (data pad map with swapped half-widths), Wp
 * (normalised weights through the weight pad map), kp/kph padded kernels. */
int orc_build_operands_f64(int dw, int dh, int nviews, const uint16_t *plane,
                           const float *weights, float *kernels, int kw,
                           int kh, int psf_type, int ffth, int fftw,
                           double *Dp, double *Wp, double *kp, double *kph);

/* fmvd_transform.impl.cu:24-56 with FP64 coordinates and exact trilinear
 * weights (the reference's texture unit uses 9-bit fixed-point weights):
 * out[z][y][x] for z in [z0, z0+nz). src is d x h x w. m = 12 floats,
 * row-major 3x4 target->source. */
void orc_transform(const uint16_t *src, int w, int h, int d, int bits,
                   const float *m, int tw, int th, int z0, int nz,
                   uint16_t *out);
/* fmvd_transform.cu:19-67, evaluated in FP64 at plane z = td/2. */
void orc_mask(int w, int h, int d, int tw, int th, int td, const float *m,
              int border, float zspacing, float *out);

#ifdef __cplusplus
}
#endif
#endif
