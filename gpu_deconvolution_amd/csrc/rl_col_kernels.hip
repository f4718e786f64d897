// Column pass of the plane-wise multi-view RL update (gfx950).
//
// Replaces, per view-iteration, the column halves of cuFFT R2C/C2R plus
// modulateAndNormalize (fmvd_deconvolve_cuda.cu:327-372, called at
// fmvd_deconvolve.impl.cpp:265,270): every kx-row of the transposed row
// spectrum (contiguous, length fftH) is transformed along y, multiplied by the
// kernel spectrum (normalisation folded in on the host) and transformed back,
// all in registers/LDS of one workgroup, so each element is read once and
// written once.
//
// kx = 0 holds two real columns packed as DC + i*Nyquist.  Its spectrum is
// split with A[k] = (C[k] + conj C[-k])/2, B[k] = (C[k] - conj C[-k])/(2i),
// multiplied by the DC and Nyquist kernel columns separately and recombined
// as A*K0 + i*B*KN, whose inverse is again DC + i*Nyquist of the result.
//
// Pass C multiplies by conj(K) (ksign = -1) when the kernel-hat spectrum is
// the conjugate of the kernel's.
#include "fft_device.hpp"
#include "hip_check.hpp"
#include "rl_launch.hpp"
#include "rl_tiles.hpp"
#include "timing.hpp"

namespace fmvd {

template <typename PL>
__global__ __launch_bounds__(ColCfg<PL>::THREADS, PL::WPE) void col_conv_kernel(ColGeom g, const float2* Xin,
                                                                       float2* Xout, long xstride,
                                                                       const float2* __restrict__ K, int nplanes,
                                                                       float ksign)
{
    using C = ColCfg<PL>;
    __shared__ typename C::E lds[C::SQ * C::L];
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    int p, grp;
    if constexpr (C::SQ == 1) {
        // one column per workgroup: the two columns of a kx pair, which share
        // every 16-byte word of the spectrum, run back to back so that their
        // half-word writes meet in L2; then the planes, so that consecutive
        // pairs of blocks reuse the same kernel columns (N is even)
        const unsigned q = lid >> 1;
        p = (int)(q % (unsigned)nplanes);
        grp = (int)((q / (unsigned)nplanes) * 2u + (lid & 1u));
    } else {
        // plane-fastest: consecutive logical blocks reuse the same kernel rows
        p = (int)(lid % (unsigned)nplanes);
        grp = (int)(lid / (unsigned)nplanes);
    }
    col_tile<PL>(g, Xin + p * xstride, Xout + p * xstride, K, grp, lds, ksign);
}

#define FMVD_COL_SIZES(X)                                                                               \
    X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(1536) X(2048) X(2560) X(3072) X(3584) X(4096) X(4608) \
    X(5120) X(5632) X(6144) X(6656) X(7168) X(7680) X(8192)

bool col_size_supported(int H)
{
    switch (H) {
#define FMVD_CASE(n) case n:
        FMVD_COL_SIZES(FMVD_CASE)
#undef FMVD_CASE
        return true;
    default:
        return false;
    }
}

int col_variants(int H)
{
    int n = 0;
    switch (H) {
#define FMVD_CASE(nn)                                                                        \
    case nn:                                                                                 \
        n = 1 + (PlanV<nn, 1>::distinct ? 1 : 0) + (PlanV<nn, 2>::distinct ? 1 : 0);         \
        break;
        FMVD_COL_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        break;
    }
    return n;
}

template <typename PL>
static void launch_col_plan(const ColGeom& g, const float2* Xin, float2* Xout, long xstride, const float2* K,
                            int nplanes, hipStream_t s, bool conj)
{
    using C = ColCfg<PL>;
    if (g.N % C::SQ && C::T > 64) throw_msg("row count not divisible by column group");
    if (C::SQ == 1 && g.N % 2) throw_msg("single-column plans need an even column count");
    dim3 grid((unsigned)((g.N + C::SQ - 1) / C::SQ * nplanes));
    launch(col_conv_kernel<PL>, grid, dim3(C::THREADS), 0, s, g, Xin, Xout, xstride, K, nplanes, conj ? -1.0f : 1.0f);
}

template <int H>
static void col_variant(int v, const ColGeom& g, const float2* Xin, float2* Xout, long xstride, const float2* K,
                        int nplanes, hipStream_t s, bool conj)
{
    if (v == 1 && PlanV<H, 1>::distinct) {
        if constexpr (PlanV<H, 1>::distinct) launch_col_plan<PlanV<H, 1>>(g, Xin, Xout, xstride, K, nplanes, s, conj);
    } else if (v == 2 && PlanV<H, 2>::distinct) {
        if constexpr (PlanV<H, 2>::distinct) launch_col_plan<PlanV<H, 2>>(g, Xin, Xout, xstride, K, nplanes, s, conj);
    } else {
        launch_col_plan<PlanV<H, 0>>(g, Xin, Xout, xstride, K, nplanes, s, conj);
    }
}

std::vector<float2> col_pass_twiddles(int H, int variant)
{
    std::vector<float2> t;
    auto get = [&]<typename PL>() { t = pass_twiddle_table<PL>(); };
    switch (H) {
#define FMVD_CASE(n)                                                                   \
    case n:                                                                            \
        if (variant == 1 && PlanV<n, 1>::distinct) get.template operator()<PlanV<n, 1>>(); \
        else if (variant == 2 && PlanV<n, 2>::distinct) get.template operator()<PlanV<n, 2>>(); \
        else get.template operator()<PlanV<n, 0>>();                                   \
        break;
        FMVD_COL_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("unsupported column FFT length");
    }
    return t;
}

void launch_col_conv(const ColGeom& g, const float2* Xin, float2* Xout, long xstride, const float2* K, int nplanes,
                     int variant, hipStream_t s, bool conj)
{
    switch (g.H) {
#define FMVD_CASE(n)                                                          \
    case n:                                                                   \
        col_variant<n>(variant, g, Xin, Xout, xstride, K, nplanes, s, conj);  \
        break;
        FMVD_COL_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("unsupported column FFT length");
    }
    FMVD_LAUNCH_CHECK();
}

}  // namespace fmvd
