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
#include "fft_device.hpp"
#include "hip_check.hpp"
#include "rl_launch.hpp"

namespace fmvd {

__device__ __forceinline__ unsigned xcd_remap_col(unsigned bid, unsigned nblocks)
{
    const unsigned q = nblocks >> 3, rem = nblocks & 7u;
    const unsigned x = bid & 7u, k = bid >> 3;
    return x * q + (x < rem ? x : rem) + k;
}

constexpr int pow2_floor(int x)
{
    int p = 1;
    while (p * 2 <= x) p *= 2;
    return p;
}

template <typename PL_> struct ColCfg {
    using PL = PL_;
    static_assert(plan_valid<PL>(), "invalid FFT plan");
    static constexpr int H = PL::N;
    static constexpr int T = PL::T;
    static constexpr int P = PL::P;
    static constexpr int L = lds_len<H>() + 1;
    // sequences per workgroup (a power of two, so it divides the row count):
    // aim for 256 threads, keep LDS <= ~64 KB
    static constexpr int SQ0 = (256 / T) < 1 ? 1 : (256 / T);
    static constexpr int SQ1 = SQ0 > 16 ? 16 : SQ0;
    static constexpr int SQ2 = 65536 / (L * 8) < 1 ? 1 : 65536 / (L * 8);
    static constexpr int SQ = pow2_floor(SQ1 < SQ2 ? SQ1 : SQ2);
    static constexpr int THREADS = SQ * T;
    static constexpr int RF = first_radix<PL, -1>();
    static constexpr int RL = last_radix<PL, -1>();
    static_assert(RL == first_radix<PL, +1>(), "forward exit layout must match inverse entry");
    static_assert(RF == last_radix<PL, +1>(), "inverse exit layout must match forward entry");
};

template <typename PL>
__global__ __launch_bounds__(ColCfg<PL>::THREADS) void col_conv_kernel(ColGeom g, const float2* __restrict__ Xin,
                                                                       float2* __restrict__ Xout, long xstride,
                                                                       const float2* __restrict__ K, int nplanes)
{
    using C = ColCfg<PL>;
    constexpr int H = C::H;
    __shared__ float2 lds[C::SQ * C::L];
    const unsigned lid = xcd_remap_col(blockIdx.x, gridDim.x);
    // plane-fastest: consecutive logical blocks reuse the same kernel rows
    const int p = lid % nplanes, grp = lid / nplanes;
    const int seq = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const int kx = grp * C::SQ + seq;
    const float2* xin = Xin + p * xstride + (long)kx * H;
    float2* row = lds + seq * C::L;

    float2 v[C::P];
#pragma unroll
    for (int i = 0; i < C::P; ++i) v[i] = xin[slot_pos<PL, C::RF>(t, i)];
    fft_regs<PL, -1>(v, row, t, g.twP[PL::variant]);

    if (grp == 0) {  // workgroup-uniform: this group holds the packed DC/Nyquist sequence
        if (seq == 0) {
#pragma unroll
            for (int i = 0; i < C::P; ++i) row[lds_pad(slot_pos<PL, C::RL>(t, i))] = v[i];
        }
        seq_sync<PL>();
        if (seq == 0) {
            const float2* K0 = K;
            const float2* KN = K + (long)g.N * H;
#pragma unroll
            for (int i = 0; i < C::P; ++i) {
                const int k = slot_pos<PL, C::RL>(t, i);
                const float2 ck = row[lds_pad(k)];
                const float2 cm = cconj(row[lds_pad(k == 0 ? 0 : H - k)]);
                const float2 a = cscale(cadd(ck, cm), 0.5f);
                const float2 d = csub(ck, cm);
                const float2 b = make_float2(0.5f * d.y, -0.5f * d.x);
                const float2 ak = cmul(a, K0[k]);
                const float2 bk = cmul(b, KN[k]);
                v[i] = make_float2(ak.x - bk.y, ak.y + bk.x);
            }
        } else {
            const float2* Kr = K + (long)kx * H;
#pragma unroll
            for (int i = 0; i < C::P; ++i) v[i] = cmul(v[i], Kr[slot_pos<PL, C::RL>(t, i)]);
        }
        seq_sync<PL>();
    } else {
        const float2* Kr = K + (long)kx * H;
#pragma unroll
        for (int i = 0; i < C::P; ++i) v[i] = cmul(v[i], Kr[slot_pos<PL, C::RL>(t, i)]);
    }

    fft_regs<PL, +1>(v, row, t, g.twP[PL::variant]);
    float2* xo = Xout + p * xstride + (long)kx * H;
#pragma unroll
    for (int i = 0; i < C::P; ++i) xo[slot_pos<PL, C::RF>(t, i)] = v[i];
}

#define FMVD_COL_SIZES(X) \
    X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(1536) X(2048) X(2560) X(3072) X(3584) X(4096)

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
                            int nplanes, hipStream_t s)
{
    using C = ColCfg<PL>;
    if (g.N % C::SQ) throw_msg("row count not divisible by column group");
    dim3 grid((unsigned)(g.N / C::SQ * nplanes));
    col_conv_kernel<PL><<<grid, C::THREADS, 0, s>>>(g, Xin, Xout, xstride, K, nplanes);
}

template <int H>
static void col_variant(int v, const ColGeom& g, const float2* Xin, float2* Xout, long xstride, const float2* K,
                        int nplanes, hipStream_t s)
{
    if (v == 1 && PlanV<H, 1>::distinct) {
        if constexpr (PlanV<H, 1>::distinct) launch_col_plan<PlanV<H, 1>>(g, Xin, Xout, xstride, K, nplanes, s);
    } else if (v == 2 && PlanV<H, 2>::distinct) {
        if constexpr (PlanV<H, 2>::distinct) launch_col_plan<PlanV<H, 2>>(g, Xin, Xout, xstride, K, nplanes, s);
    } else {
        launch_col_plan<PlanV<H, 0>>(g, Xin, Xout, xstride, K, nplanes, s);
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
                     int variant, hipStream_t s)
{
    switch (g.H) {
#define FMVD_CASE(n)                                                          \
    case n:                                                                   \
        col_variant<n>(variant, g, Xin, Xout, xstride, K, nplanes, s);        \
        break;
        FMVD_COL_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("unsupported column FFT length");
    }
    FMVD_LAUNCH_CHECK();
}

}  // namespace fmvd
