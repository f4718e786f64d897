// Launch interface between the RL host engine and the RL device kernels.
//
// Device layout of one plane's "row spectrum" (transposed, packed, with kx
// pairs interleaved):
//   X(kx, y) at xpos(kx, y) = (kx >> 1) * 2 fftH + 2 y + (kx & 1),
//   kx in [0, N) with N = fftW/2, y in [0, fftH), complex64,
//   X(kx, y) = sum_x f[y][x] exp(-2 pi i kx x / fftW)   for kx in [1, N),
//   X(0, y)  = (DC(y), Nyquist(y))  packed as (re, im),
// i.e. one plane occupies N*fftH complex = fftW*fftH floats.  The column pass
// transforms each kx sequence (stride 2 complex along y, which loads at full
// bandwidth); the row passes read and write it transposed in tiles of G rows,
// whose runs are 2G complex long (128 bytes at G = 8, whole cache lines).
//
// Kernel spectra hold one contiguous column per kx in [0, N] (K[ky][kx] at
// kx*fftH + ky, kx = 0 holding K[ky][0] and kx = N the Nyquist column), all
// complex and already multiplied by the pass's normalisation: the column
// pass, their only reader, then loads them with unit stride.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace fmvd {

__host__ __device__ __forceinline__ long xpos(int kx, int y, int H)
{
    return (long)(kx >> 1) * 2 * H + 2 * y + (kx & 1);
}

struct RowGeom {
    int H, W, N;          // fftH, fftW, fftW/2
    int dataH, dataW;
    int dkY, dkX;         // data pad half-widths (kernelW/2, kernelH/2: swapped as the reference)
    int wkY, wkX;         // weight pad half-widths (kernelH/2, kernelW/2)
    const float2* twP[3]; // per-pass twiddle tables of the row FFT plan variants 0..2
    const float2* twW;    // exp(-2 pi i p / W), p < N
};

struct ColGeom {
    int H, N;
    const float2* twP[3]; // per-pass twiddle tables of the column FFT plan variants 0..2
};

// Pass B: X <- R2C(ratio(D, C2R(X))) * scale for nplanes planes, in place.
// data: first plane of this view; dstride: elements between planes.
template <typename SAMPLE>
void launch_row_ratio(const RowGeom& g, float2* X, long xstride, const SAMPLE* data, long dstride,
                      int nplanes, int variant, hipStream_t s, float scale = 1.0f);
// Pass D: C = C2R(XC); E += w*(E*C - E) on the spatial estimate ES (fftH x
// fftW floats per plane, row-major); XE <- R2C(E)/W.
void launch_row_update(const RowGeom& g, float2* XE, const float2* XC, float* ES, long xstride, long estride,
                       const float* weights, int nplanes, int variant, hipStream_t s);
// Egress: crop ES to dataH x dataW; out = sample(E), est = E (either may be null).
template <typename SAMPLE>
void launch_final(const RowGeom& g, const float* ES, long estride, SAMPLE* out, float* est, long ostride,
                  int nplanes, hipStream_t s);
// Warm-start ingest: ES <- pad(E0), XE <- R2C(ES)/W; E0 is dataH x dataW per plane.
void launch_row_init(const RowGeom& g, float2* XE, float* ES, long xstride, long estride, const float* e0,
                     long e0stride, int nplanes, int variant, hipStream_t s);
// Constant start E0 = value everywhere: ES = value, XE[0][y] = (value, 0), rest 0.
void launch_const_init(const RowGeom& g, float2* XE, float* ES, long xstride, long estride, float value,
                       int nplanes, hipStream_t s);
// Passes A / C: Xout[kx] <- IFFT_y(FFT_y(Xin[kx]) * K[kx]) for every kx
// (conj(K[kx]) when conj is set).
void launch_col_conv(const ColGeom& g, const float2* Xin, float2* Xout, long xstride,
                     const float2* K, int nplanes, int variant, hipStream_t s, bool conj = false);

// Per-pass twiddle tables (host side) of a row / column FFT plan variant.
std::vector<float2> row_pass_twiddles(int N, int variant);
std::vector<float2> col_pass_twiddles(int H, int variant);

// acc[0] += sum (ES_crop - ref)^2, acc[1] += sum ref^2 over nplanes planes
// (FP64 accumulation on the device).
void launch_error_sums(const RowGeom& g, const float* ES, long estride, const float* ref, long rstride, int nplanes,
                       double* acc, hipStream_t s);

// Per-pixel view normalisation of n-pixel device weight images, in the host
// rule's float operations: sum = ((0 + w_0) + w_1) + ... accumulated one
// view per launch (first = view 0), then out_v = sum > 0 ? w_v / sum : w_v.
void launch_weight_accumulate(float* sum, const float* w, long n, bool first, hipStream_t s);
void launch_weight_normalize(float* out, const float* w, const float* sum, long n, hipStream_t s);

bool row_size_supported(int N);
bool col_size_supported(int H);
// Number of compiled FFT-plan variants for a row / column length (>= 1 when
// supported).  Variant 0 is the default plan.
int row_variants(int N);
int col_variants(int H);

// Plan variant chosen for each kernel kind.
struct Variants {
    int column = 0, ratio = 0, update = 0, ingest = 0, egress = 0;
};

}  // namespace fmvd
