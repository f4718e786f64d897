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
#include "fft_device.hpp"
#include "hip_check.hpp"
#include "rl_launch.hpp"

namespace fmvd {

__device__ __forceinline__ int pad_index(int i, int n, int half)
{
    return i < n ? i : (i < n + half ? n - 1 : 0);
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of logical block ids so neighbouring tiles share its L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nblocks)
{
    const unsigned q = nblocks >> 3, rem = nblocks & 7u;
    const unsigned x = bid & 7u, k = bid >> 3;
    return x * q + (x < rem ? x : rem) + k;
}

template <int N> struct RowCfg {
    using PL = Plan<N>;
    static constexpr int T = PL::T;
    static constexpr int P = PL::P;
    static constexpr int G = (512 / T) > 16 ? 16 : ((512 / T) < 4 ? 4 : (512 / T));
    static constexpr int L = lds_len<N>() + 1;  // odd-ish pitch spreads the transposed tile over banks
    static constexpr int THREADS = G * T;
    static constexpr int RFI = first_radix<PL, +1>();  // layout on entry of the inverse
    static constexpr int RLI = last_radix<PL, +1>();   // layout on exit of the inverse (= forward entry)
    static constexpr int RLF = last_radix<PL, -1>();   // layout on exit of the forward
    static_assert(RLI == first_radix<PL, -1>(), "inverse exit layout must match forward entry");
};

// Transposed tile load: lds[g][kx] <- X[kx][y0+g]
template <int N>
__device__ __forceinline__ void load_tile(float2* lds, const float2* __restrict__ X, int H, int y0, int tid)
{
    using C = RowCfg<N>;
#pragma unroll 4
    for (int k = 0; k < C::P; ++k) {
        const int e = tid + k * C::THREADS;
        const int g = e % C::G, kx = e / C::G;
        lds[g * C::L + lds_pad(kx)] = X[(long)kx * H + y0 + g];
    }
}

// C2R pre-processing: from the packed half-spectrum of one row (in LDS),
// form the length-N complex input of the inverse FFT at this thread's
// first-pass positions:  Z'[p] = A + iB,  A = X[p] + conj(X[N-p]),
// B = (X[p] - conj(X[N-p])) exp(+2 pi i p / W);  Z'[0] = (DC+Ny, DC-Ny).
template <int N>
__device__ __forceinline__ void c2r_pre(float2* v, const float2* row, int t, const float2* __restrict__ twW)
{
    using C = RowCfg<N>;
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int p = slot_pos<typename C::PL, C::RFI>(t, i);
        const float2 xk = row[lds_pad(p)];
        if (p == 0) {
            v[i] = make_float2(xk.x + xk.y, xk.x - xk.y);
        } else {
            const float2 xm = cconj(row[lds_pad(N - p)]);
            const float2 a = cadd(xk, xm);
            const float2 b = cmulc(csub(xk, xm), twW[p]);
            v[i] = make_float2(a.x - b.y, a.y + b.x);
        }
    }
}

// R2C post-processing and transposed store: v holds Z (forward FFT of the
// packed row) at the forward's exit layout.  X[k] = (Z[k]+conj Z[N-k])/2 +
// exp(-2 pi i k/W) (Z[k]-conj Z[N-k])/(2i);  X[0] = (ReZ0+ImZ0, ReZ0-ImZ0).
template <int N>
__device__ __forceinline__ void r2c_post_store(const float2* v, float2* lds, float2* __restrict__ X,
                                               int H, int y0, int tid, int seq, int t,
                                               const float2* __restrict__ twW, float scale)
{
    using C = RowCfg<N>;
    float2* row = lds + seq * C::L;
#pragma unroll
    for (int i = 0; i < C::P; ++i) row[lds_pad(slot_pos<typename C::PL, C::RLF>(t, i))] = v[i];
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < C::P; ++k) {
        const int e = tid + k * C::THREADS;
        const int g = e % C::G, kx = e / C::G;
        const float2* r = lds + g * C::L;
        const float2 zk = r[lds_pad(kx)];
        float2 xo;
        if (kx == 0) {
            xo = make_float2(zk.x + zk.y, zk.x - zk.y);
        } else {
            const float2 zm = cconj(r[lds_pad(N - kx)]);
            const float2 fe = cscale(cadd(zk, zm), 0.5f);
            const float2 d = csub(zk, zm);
            const float2 fo = make_float2(0.5f * d.y, -0.5f * d.x);
            xo = cadd(fe, cmul(twW[kx], fo));
        }
        X[(long)kx * H + y0 + g] = cscale(xo, scale);
    }
}

// ------------------------------------------------------------------ pass B
template <typename SAMPLE, int N>
__global__ __launch_bounds__(RowCfg<N>::THREADS) void row_ratio_kernel(
    RowGeom g, float2* __restrict__ X, long xstride, const SAMPLE* __restrict__ data, long dstride,
    int nplanes)
{
    using C = RowCfg<N>;
    using PL = typename C::PL;
    __shared__ float2 lds[C::G * C::L];
    const int nrg = g.H / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    const int y0 = rg * C::G;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    float2* Xp = X + p * xstride;
    float2* row = lds + seq * C::L;

    load_tile<N>(lds, Xp, g.H, y0, tid);
    __syncthreads();
    float2 v[C::P];
    c2r_pre<N>(v, row, t, g.twW);
    __syncthreads();
    fft_regs<PL, +1>(v, row, t, g.twN);

    // R = Q > 1 ? D / Q : Q  on the padded row (data pad map, swapped half-widths)
    const SAMPLE* drow = data + p * dstride + (long)pad_index(y, g.dataH, g.dkY) * g.dataW;
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int n = slot_pos<PL, C::RLI>(t, i);
        const int x0 = 2 * n, x1 = 2 * n + 1;
        const float q0 = v[i].x, q1 = v[i].y;
        const float d0 = (float)drow[pad_index(x0, g.dataW, g.dkX)];
        const float d1 = (float)drow[pad_index(x1, g.dataW, g.dkX)];
        v[i].x = q0 > 1.0f ? d0 / q0 : q0;
        v[i].y = q1 > 1.0f ? d1 / q1 : q1;
    }
    fft_regs<PL, -1>(v, row, t, g.twN);
    r2c_post_store<N>(v, lds, Xp, g.H, y0, tid, seq, t, g.twW, 1.0f);
}

// ------------------------------------------------------------------ pass D
template <int N>
__global__ __launch_bounds__(RowCfg<N>::THREADS) void row_update_kernel(
    RowGeom g, float2* __restrict__ XE, const float2* __restrict__ XC, long xstride,
    const float* __restrict__ weights, int nplanes)
{
    using C = RowCfg<N>;
    using PL = typename C::PL;
    __shared__ float2 lds[C::G * C::L];
    const int nrg = g.H / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    const int y0 = rg * C::G;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    float2* XEp = XE + p * xstride;
    const float2* XCp = XC + p * xstride;
    float2* row = lds + seq * C::L;

    // current estimate E = C2R(XE)  (XE carries the 1/W normalisation)
    load_tile<N>(lds, XEp, g.H, y0, tid);
    __syncthreads();
    float2 e[C::P];
    c2r_pre<N>(e, row, t, g.twW);
    __syncthreads();
    fft_regs<PL, +1>(e, row, t, g.twN);

    // correction C = C2R(XC)
    load_tile<N>(lds, XCp, g.H, y0, tid);
    __syncthreads();
    float2 v[C::P];
    c2r_pre<N>(v, row, t, g.twW);
    __syncthreads();
    fft_regs<PL, +1>(v, row, t, g.twN);

    // E <- E + w (E*C - E)   (fmvd_deconvolve_cuda.cu:387-391 operation order)
    const float* wrow = weights + (long)pad_index(y, g.dataH, g.wkY) * g.dataW;
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int n = slot_pos<PL, C::RLI>(t, i);
        const float w0 = wrow[pad_index(2 * n, g.dataW, g.wkX)];
        const float w1 = wrow[pad_index(2 * n + 1, g.dataW, g.wkX)];
        float t0 = e[i].x * v[i].x, t1 = e[i].y * v[i].y;
        float c0 = t0 - e[i].x, c1 = t1 - e[i].y;
        c0 *= w0;
        c1 *= w1;
        v[i].x = e[i].x + c0;
        v[i].y = e[i].y + c1;
    }
    fft_regs<PL, -1>(v, row, t, g.twN);
    r2c_post_store<N>(v, lds, XEp, g.H, y0, tid, seq, t, g.twW, 1.0f / (float)g.W);
}

// ------------------------------------------------------------------ egress
template <typename SAMPLE, int N>
__global__ __launch_bounds__(RowCfg<N>::THREADS) void row_final_kernel(
    RowGeom g, const float2* __restrict__ XE, long xstride, SAMPLE* __restrict__ out,
    float* __restrict__ est, long ostride, int nplanes)
{
    using C = RowCfg<N>;
    using PL = typename C::PL;
    __shared__ float2 lds[C::G * C::L];
    const int nrg = (g.dataH + C::G - 1) / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    const int y0 = rg * C::G;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    const float2* XEp = XE + p * xstride;
    float2* row = lds + seq * C::L;

    load_tile<N>(lds, XEp, g.H, y0, tid);
    __syncthreads();
    float2 v[C::P];
    c2r_pre<N>(v, row, t, g.twW);
    __syncthreads();
    fft_regs<PL, +1>(v, row, t, g.twN);
    if (y >= g.dataH) return;
    constexpr float maxv = (float)((1u << (8 * sizeof(SAMPLE))) - 1u);
    const long obase = p * ostride + (long)y * g.dataW;
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int n = slot_pos<PL, C::RLI>(t, i);
        const float ev[2] = {v[i].x, v[i].y};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int x = 2 * n + c;
            if (x < g.dataW) {
                if (est) est[obase + x] = ev[c];
                if (out) {
                    // fmvd_deconvolve_cuda.impl.cu:121-126: min(E, max) then
                    // truncation; negative and NaN saturate to 0.
                    const float tv = ev[c] > maxv ? maxv : ev[c];
                    out[obase + x] = tv > 0.0f ? (SAMPLE)tv : (SAMPLE)0;
                }
            }
        }
    }
}

// ------------------------------------------------------------------ ingest
template <int N>
__global__ __launch_bounds__(RowCfg<N>::THREADS) void row_init_kernel(
    RowGeom g, float2* __restrict__ XE, long xstride, const float* __restrict__ e0, long estride,
    int nplanes)
{
    using C = RowCfg<N>;
    using PL = typename C::PL;
    __shared__ float2 lds[C::G * C::L];
    const int nrg = g.H / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    const int y0 = rg * C::G;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    float2* row = lds + seq * C::L;
    const float* erow = e0 + p * estride + (long)pad_index(y, g.dataH, g.dkY) * g.dataW;
    float2 v[C::P];
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int n = slot_pos<PL, C::RLI>(t, i);
        v[i].x = erow[pad_index(2 * n, g.dataW, g.dkX)];
        v[i].y = erow[pad_index(2 * n + 1, g.dataW, g.dkX)];
    }
    fft_regs<PL, -1>(v, row, t, g.twN);
    r2c_post_store<N>(v, lds, XE + p * xstride, g.H, y0, tid, seq, t, g.twW, 1.0f / (float)g.W);
}

__global__ void const_init_kernel(float2* __restrict__ XE, long xstride, long per_plane, int H,
                                  float value)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= per_plane) return;
    XE[p * xstride + i] = (i < H) ? make_float2(value, 0.0f) : make_float2(0.0f, 0.0f);
}

// ------------------------------------------------------------- dispatchers
#define FMVD_ROW_SIZES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(768) X(1024) X(1280) X(1536) X(1792) X(2048)

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

template <int N> static inline dim3 row_grid(int rows, int nplanes)
{
    return dim3((unsigned)((rows + RowCfg<N>::G - 1) / RowCfg<N>::G * nplanes));
}

template <typename SAMPLE>
void launch_row_ratio(const RowGeom& g, float2* X, long xstride, const SAMPLE* data, long dstride,
                      int nplanes, hipStream_t s)
{
    switch (g.N) {
#define FMVD_CASE(n)                                                                          \
    case n:                                                                                   \
        row_ratio_kernel<SAMPLE, n><<<row_grid<n>(g.H, nplanes), RowCfg<n>::THREADS, 0, s>>>( \
            g, X, xstride, data, dstride, nplanes);                                           \
        break;
        FMVD_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("unsupported row FFT length");
    }
    FMVD_LAUNCH_CHECK();
}

void launch_row_update(const RowGeom& g, float2* XE, const float2* XC, long xstride,
                       const float* weights, int nplanes, hipStream_t s)
{
    switch (g.N) {
#define FMVD_CASE(n)                                                                     \
    case n:                                                                              \
        row_update_kernel<n><<<row_grid<n>(g.H, nplanes), RowCfg<n>::THREADS, 0, s>>>(   \
            g, XE, XC, xstride, weights, nplanes);                                       \
        break;
        FMVD_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("unsupported row FFT length");
    }
    FMVD_LAUNCH_CHECK();
}

template <typename SAMPLE>
void launch_row_final(const RowGeom& g, const float2* XE, long xstride, SAMPLE* out, float* est,
                      long ostride, int nplanes, hipStream_t s)
{
    switch (g.N) {
#define FMVD_CASE(n)                                                                              \
    case n:                                                                                       \
        row_final_kernel<SAMPLE, n><<<row_grid<n>(g.dataH, nplanes), RowCfg<n>::THREADS, 0, s>>>( \
            g, XE, xstride, out, est, ostride, nplanes);                                          \
        break;
        FMVD_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("unsupported row FFT length");
    }
    FMVD_LAUNCH_CHECK();
}

void launch_row_init(const RowGeom& g, float2* XE, long xstride, const float* e0, long estride,
                     int nplanes, hipStream_t s)
{
    switch (g.N) {
#define FMVD_CASE(n)                                                                 \
    case n:                                                                          \
        row_init_kernel<n><<<row_grid<n>(g.H, nplanes), RowCfg<n>::THREADS, 0, s>>>( \
            g, XE, xstride, e0, estride, nplanes);                                   \
        break;
        FMVD_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("unsupported row FFT length");
    }
    FMVD_LAUNCH_CHECK();
}

void launch_const_init(const RowGeom& g, float2* XE, long xstride, float value, int nplanes,
                       hipStream_t s)
{
    const long per_plane = (long)g.N * g.H;
    dim3 grid((unsigned)((per_plane + 255) / 256), (unsigned)nplanes);
    const_init_kernel<<<grid, 256, 0, s>>>(XE, xstride, per_plane, g.H, value);
    FMVD_LAUNCH_CHECK();
}

template void launch_row_ratio<uint16_t>(const RowGeom&, float2*, long, const uint16_t*, long, int, hipStream_t);
template void launch_row_ratio<uint8_t>(const RowGeom&, float2*, long, const uint8_t*, long, int, hipStream_t);
template void launch_row_final<uint16_t>(const RowGeom&, const float2*, long, uint16_t*, float*, long, int, hipStream_t);
template void launch_row_final<uint8_t>(const RowGeom&, const float2*, long, uint8_t*, float*, long, int, hipStream_t);

}  // namespace fmvd
