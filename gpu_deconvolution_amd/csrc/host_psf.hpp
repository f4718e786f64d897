// Host-side plan preparation: FFT extents, PSF preparation, weight
// normalisation and kernel spectra.  Runs once per plan; all arithmetic that
// the reference did in FP32 on the GPU (the cuFFT convolutions of its PSF
// preparation and the kernel spectra) is done here in FP64 and rounded once.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace fmvd {

// fmvd_deconvolve.cpp:22-51
int snap_transform_size(int n);

// fmvd_deconvolve.cpp:53-68: sum in double, each value divided by (float)sum.
void normalize_kernel_f32(float* k, int len);

// fmvd_deconvolve.cpp:139-204 (computeKernel2) evaluated in FP64 from the raw
// (un-normalised) kernels and normalised; returns nviews * kh * kw values.
std::vector<double> compute_kernel_hat(const float* const* raw, int kw, int kh, int nviews,
                                       int psf_type);

// fmvd_deconvolve.impl.cpp:86-124 with fmvd_deconvolve_cuda.cu:71-177 float
// semantics, on the unpadded grid: per pixel sum_v w_v (float, view order),
// then w_v / sum where sum > 0.  Padding commutes with this normalisation, so
// the row kernels index the result through the weight pad map.
std::vector<float> normalized_weights(const float* const* w, int dw, int dh, int nviews);

// Spectrum of a kh x kw kernel padded onto the H x W torus centred at (0,0)
// with wraparound (fmvd_deconvolve_cuda.cu:13-38), multiplied by `scale`,
// in the device layout of rl_launch.hpp: K[ky][kx] at kx H + ky for
// kx <= W/2 (one contiguous column per kx).
std::vector<float2> kernel_spectrum_transposed(const double* k, int kw, int kh, int H, int W,
                                               double scale);

// tw[m] = exp(-2 pi i m / n) for m < count, computed in double.
std::vector<float2> twiddles(int n, int count);

// Unnormalised in-place complex FFT in FP64 (any length whose prime factors
// are small), sign -1 forward / +1 inverse.
void host_fft(std::vector<double>& re_im, int n, int sign);

}  // namespace fmvd
