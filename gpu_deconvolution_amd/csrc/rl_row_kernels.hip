// Row passes of the plane-wise multi-view RL update (gfx950).
//
// Each workgroup owns G consecutive rows of one padded plane.  It loads the
// rows' half-spectra from the transposed layout X[kx][y] (G consecutive
// complex per kx: coalesced G*8-byte runs), performs per row
//   C2R (inverse real FFT of length fftW via a complex FFT of length N=fftW/2)
//   -> fused pointwise stage -> R2C,
// and stores the result transposed again.  The pointwise stages are the
// reference's divide_kernel (fmvd_deconvolve_cuda.impl.cu:159-176) and
// multiply32_kernel (fmvd_deconvolve_cuda.cu:374-392); ingest and egress
// replace padDataClampToBorderAndInitialize (impl.cu:14-64) and unpadData
// (impl.cu:107-128).
//
// Every kernel is templated on its FFT plan; the host picks one of the
// compiled plan variants per kernel kind (see rl_launch.hpp).
#include "fft_device.hpp"
#include "hip_check.hpp"
#include "rl_launch.hpp"
#include "rl_tiles.hpp"
#include "timing.hpp"

namespace fmvd {

// ------------------------------------------------------------------ pass B
template <typename SAMPLE, typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS, PL::WPE) void row_ratio_kernel(
    RowGeom g, float2* __restrict__ X, long xstride, const SAMPLE* __restrict__ data, long dstride, int nplanes,
    float scale)
{
    using C = RowCfg<PL>;
    __shared__ typename C::E lds[C::G * C::L];
    const int nrg = g.H / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    ratio_tile<SAMPLE, PL>(g, X + p * xstride, data + p * dstride, rg * C::G, lds, scale);
}

// ------------------------------------------------------------------ pass D
// Row group / plane of a block when planes are taken in groups of kPlaneGroup
// with the plane fastest inside a group: the blocks that run together on one
// XCD then share the weight rows in its L2, while a row group's neighbours
// (which share the tile's cache lines) stay only kPlaneGroup blocks away.
constexpr int kPlaneGroup = 8;
__device__ __forceinline__ void grouped_tile(unsigned lid, int nrg, int nplanes, int& rg, int& p)
{
    const int pg = (int)(lid / (unsigned)(nrg * kPlaneGroup));
    const int base = pg * kPlaneGroup;
    const int kk = min(kPlaneGroup, nplanes - base);
    const int rem = (int)lid - pg * nrg * kPlaneGroup;
    rg = rem / kk;
    p = base + rem % kk;
}

template <typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS, PL::WPE) void row_update_kernel(
    RowGeom g, float2* XE, const float2* XC, float* __restrict__ ES, long xstride,
    long estride, const float* __restrict__ weights, int nplanes)
{
    using C = RowCfg<PL>;
    __shared__ typename C::E lds[C::G * C::L];
    const int nrg = g.H / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    int rg, p;
    grouped_tile(lid, nrg, nplanes, rg, p);
    update_tile<PL>(g, XE + p * xstride, XC + p * xstride, ES + p * estride, weights, rg * C::G, lds);
}

// ------------------------------------------------------------------ egress
// fmvd_deconvolve_cuda.impl.cu:107-128: crop, min(E, max), truncate; negative
// and NaN saturate to 0.
template <typename SAMPLE>
__global__ void final_kernel(const float* __restrict__ ES, long estride, int W, int dataH, int dataW,
                             SAMPLE* __restrict__ out, float* __restrict__ est, long ostride)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int p = blockIdx.z;
    if (x >= dataW) return;
    const float e = ES[p * estride + (long)y * W + x];
    const long o = p * ostride + (long)y * dataW + x;
    if (est) est[o] = e;
    if (out) {
        constexpr float maxv = (float)((1u << (8 * sizeof(SAMPLE))) - 1u);
        const float tv = e > maxv ? maxv : e;
        out[o] = tv > 0.0f ? (SAMPLE)tv : (SAMPLE)0;
    }
}

// ------------------------------------------------------------------ ingest
template <typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS, PL::WPE) void row_init_kernel(
    RowGeom g, float2* __restrict__ XE, float* __restrict__ ES, long xstride, long estride,
    const float* __restrict__ e0, long e0stride, int nplanes)
{
    using C = RowCfg<PL>;
    __shared__ typename C::E lds[C::G * C::L];
    const int nrg = g.H / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    init_tile<PL>(g, XE + p * xstride, ES + p * estride, e0 + p * e0stride, rg * C::G, lds);
}

__global__ void const_init_kernel(float2* __restrict__ XE, float* __restrict__ ES, long xstride, long estride,
                                  long per_plane, int H, float value)
{
    // per_plane = N*H complex of XE = H*W/2 float2 of ES
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= per_plane) return;
    // kx = 0 (DC + i Nyquist) sits at the even positions below 2H
    XE[p * xstride + i] = (i < 2 * H && !(i & 1)) ? make_float2(value, 0.0f) : make_float2(0.0f, 0.0f);
    reinterpret_cast<float2*>(ES + p * estride)[i] = make_float2(value, value);
}

// ------------------------------------------------------- convergence sums
__global__ void error_sums_kernel(const float* __restrict__ ES, long estride, int W, int dataH, int dataW,
                                  const float* __restrict__ ref, long rstride, double* __restrict__ acc)
{
    __shared__ double se[256], sr[256];
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, p = blockIdx.z;
    double e = 0.0, r = 0.0;
    if (x < dataW) {
        const double rv = ref[p * rstride + (long)y * dataW + x];
        const double d = (double)ES[p * estride + (long)y * W + x] - rv;
        e = d * d;
        r = rv * rv;
    }
    se[threadIdx.x] = e;
    sr[threadIdx.x] = r;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            se[threadIdx.x] += se[threadIdx.x + k];
            sr[threadIdx.x] += sr[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(acc, se[0]);
        atomicAdd(acc + 1, sr[0]);
    }
}

void launch_error_sums(const RowGeom& g, const float* ES, long estride, const float* ref, long rstride, int nplanes,
                       double* acc, hipStream_t s)
{
    if (nplanes <= 0) return;
    dim3 grid((unsigned)((g.dataW + 255) / 256), (unsigned)g.dataH, (unsigned)nplanes);
    error_sums_kernel<<<grid, 256, 0, s>>>(ES, estride, g.W, g.dataH, g.dataW, ref, rstride, acc);
    FMVD_LAUNCH_CHECK();
}

// ------------------------------------------------------- device weights
// fmvd_deconvolve.impl.cpp:86-124 / normalizeWeights_kernel
// (fmvd_deconvolve_cuda.cu:160-199) on unpadded weight images: the sum over
// views is accumulated in view order, then each view is divided by it where
// it is positive.  The same float operations as the host rule, so weights
// normalised here equal the host's bit for bit.
__global__ void weight_accumulate_kernel(float* __restrict__ sum, const float* __restrict__ w, long n, int first)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sum[i] = (first ? 0.0f : sum[i]) + w[i];
}

__global__ void weight_normalize_kernel(float* __restrict__ out, const float* __restrict__ w,
                                        const float* __restrict__ sum, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = sum[i], v = w[i];
    out[i] = s > 0.0f ? v / s : v;
}

void launch_weight_accumulate(float* sum, const float* w, long n, bool first, hipStream_t s)
{
    if (n <= 0) return;
    weight_accumulate_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(sum, w, n, first ? 1 : 0);
    FMVD_LAUNCH_CHECK();
}

void launch_weight_normalize(float* out, const float* w, const float* sum, long n, hipStream_t s)
{
    if (n <= 0) return;
    weight_normalize_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(out, w, sum, n);
    FMVD_LAUNCH_CHECK();
}

// ------------------------------------------------------------- dispatchers
#define FMVD_ROW_SIZES(X)                                                                              \
    X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(768) X(1024) X(1280) X(1536) X(1792) X(2048) X(2304) \
    X(2560) X(2816) X(3072) X(3328) X(3584) X(3840) X(4096)

bool row_size_supported(int N)
{
    switch (N) {
#define FMVD_CASE(n) case n:
        FMVD_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
        return true;
    default:
        return false;
    }
}

int row_variants(int N)
{
    int n = 0;
    switch (N) {
#define FMVD_CASE(nn)                              \
    case nn:                                       \
        n = 1 + (PlanV<nn, 1>::distinct ? 1 : 0) + \
            (PlanV<nn, 2>::distinct ? 1 : 0);      \
        break;
        FMVD_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        break;
    }
    return n;
}

template <typename PL> static inline dim3 row_grid(int rows, int nplanes)
{
    return dim3((unsigned)((rows + RowCfg<PL>::G - 1) / RowCfg<PL>::G * nplanes));
}

// Calls f.template operator()<PlanV<N, v>>() for the requested variant
// (falling back to variant 0 when that variant is not distinct).
template <int N, typename F> static void with_variant(int v, F&& f)
{
    if (v == 1 && PlanV<N, 1>::distinct) {
        if constexpr (PlanV<N, 1>::distinct) f.template operator()<PlanV<N, 1>>();
    } else if (v == 2 && PlanV<N, 2>::distinct) {
        if constexpr (PlanV<N, 2>::distinct) f.template operator()<PlanV<N, 2>>();
    } else {
        f.template operator()<PlanV<N, 0>>();
    }
}

template <typename F> static void dispatch_row(int N, int v, F&& f)
{
    switch (N) {
#define FMVD_CASE(n)             \
    case n:                      \
        with_variant<n>(v, f);   \
        break;
        FMVD_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("unsupported row FFT length");
    }
}

std::vector<float2> row_pass_twiddles(int N, int variant)
{
    std::vector<float2> t;
    dispatch_row(N, variant, [&]<typename PL>() { t = pass_twiddle_table<PL>(); });
    return t;
}

template <typename SAMPLE>
void launch_row_ratio(const RowGeom& g, float2* X, long xstride, const SAMPLE* data, long dstride, int nplanes,
                      int variant, hipStream_t s, float scale)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        launch(row_ratio_kernel<SAMPLE, PL>, row_grid<PL>(g.H, nplanes), dim3(RowCfg<PL>::THREADS), 0, s, g, X,
               xstride, data, dstride, nplanes, scale);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_row_update(const RowGeom& g, float2* XE, const float2* XC, float* ES, long xstride, long estride,
                       const float* weights, int nplanes, int variant, hipStream_t s)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        launch(row_update_kernel<PL>, row_grid<PL>(g.H, nplanes), dim3(RowCfg<PL>::THREADS), 0, s, g, XE, XC, ES,
               xstride, estride, weights, nplanes);
    });
    FMVD_LAUNCH_CHECK();
}

template <typename SAMPLE>
void launch_final(const RowGeom& g, const float* ES, long estride, SAMPLE* out, float* est, long ostride,
                  int nplanes, hipStream_t s)
{
    if (nplanes <= 0) return;
    dim3 grid((unsigned)((g.dataW + 255) / 256), (unsigned)g.dataH, (unsigned)nplanes);
    launch(final_kernel<SAMPLE>, grid, dim3(256), 0, s, ES, estride, g.W, g.dataH, g.dataW, out, est, ostride);
    FMVD_LAUNCH_CHECK();
}

void launch_row_init(const RowGeom& g, float2* XE, float* ES, long xstride, long estride, const float* e0,
                     long e0stride, int nplanes, int variant, hipStream_t s)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        launch(row_init_kernel<PL>, row_grid<PL>(g.H, nplanes), dim3(RowCfg<PL>::THREADS), 0, s, g, XE, ES,
               xstride, estride, e0, e0stride, nplanes);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_const_init(const RowGeom& g, float2* XE, float* ES, long xstride, long estride, float value,
                       int nplanes, hipStream_t s)
{
    const long per_plane = (long)g.N * g.H;
    dim3 grid((unsigned)((per_plane + 255) / 256), (unsigned)nplanes);
    launch(const_init_kernel, grid, dim3(256), 0, s, XE, ES, xstride, estride, per_plane, g.H, value);
    FMVD_LAUNCH_CHECK();
}

template void launch_row_ratio<uint16_t>(const RowGeom&, float2*, long, const uint16_t*, long, int, int,
                                         hipStream_t, float);
template void launch_row_ratio<uint8_t>(const RowGeom&, float2*, long, const uint8_t*, long, int, int,
                                        hipStream_t, float);
template void launch_final<uint16_t>(const RowGeom&, const float*, long, uint16_t*, float*, long, int,
                                     hipStream_t);
template void launch_final<uint8_t>(const RowGeom&, const float*, long, uint8_t*, float*, long, int, hipStream_t);

}  // namespace fmvd
