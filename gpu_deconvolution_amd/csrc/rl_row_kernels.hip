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

// LDS row pitch >= len with 2*L = 64/G (mod 64) (4-byte banks): the
// transposed tile access (G consecutive rows per 8-byte column element) then
// touches distinct bank pairs across a half-wavefront.
constexpr int row_pitch(int len, int G)
{
    const int want = (32 / G) % 32;
    int L = len;
    while (L % 32 != want) ++L;
    return L;
}

template <typename PL_> struct RowCfg {
    using PL = PL_;
    static_assert(plan_valid<PL>(), "invalid FFT plan");
    static constexpr int N = PL::N;
    static constexpr int T = PL::T;
    static constexpr int P = PL::P;
    // rows per workgroup: 256 threads where possible, at least 8 rows so the
    // transposed loads/stores move 64-byte runs
    static constexpr int G = (256 / T) > 16 ? 16 : ((256 / T) < 8 ? 8 : (256 / T));
    static constexpr int L = row_pitch(lds_len<N>(), G);
    static constexpr int THREADS = G * T;
    static constexpr int RFI = first_radix<PL, +1>();  // layout on entry of the inverse
    static constexpr int RLI = last_radix<PL, +1>();   // layout on exit of the inverse (= forward entry)
    static constexpr int RLF = last_radix<PL, -1>();   // layout on exit of the forward
    static_assert(RLI == first_radix<PL, -1>(), "inverse exit layout must match forward entry");
};

// Transposed tile load: lds[g][kx] <- X[kx][y0+g]
template <typename C>
__device__ __forceinline__ void load_tile(float2* lds, const float2* __restrict__ X, int H, int y0, int tid)
{
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
template <typename C>
__device__ __forceinline__ void c2r_pre(float2* v, const float2* row, int t, const float2* __restrict__ twW)
{
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int p = slot_pos<typename C::PL, C::RFI>(t, i);
        const float2 xk = row[lds_pad(p)];
        if (p == 0) {
            v[i] = make_float2(xk.x + xk.y, xk.x - xk.y);
        } else {
            const float2 xm = cconj(row[lds_pad(C::N - p)]);
            const float2 a = cadd(xk, xm);
            const float2 b = cmulc(csub(xk, xm), twW[p]);
            v[i] = make_float2(a.x - b.y, a.y + b.x);
        }
    }
}

// R2C post-processing and transposed store: v holds Z (forward FFT of the
// packed row) at the forward's exit layout.  X[k] = (Z[k]+conj Z[N-k])/2 +
// exp(-2 pi i k/W) (Z[k]-conj Z[N-k])/(2i);  X[0] = (ReZ0+ImZ0, ReZ0-ImZ0).
template <typename C>
__device__ __forceinline__ void r2c_post_store(const float2* v, float2* lds, float2* __restrict__ X, int H,
                                               int y0, int tid, int seq, int t, const float2* __restrict__ twW,
                                               float scale)
{
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
            const float2 zm = cconj(r[lds_pad(C::N - kx)]);
            const float2 fe = cscale(cadd(zk, zm), 0.5f);
            const float2 d = csub(zk, zm);
            const float2 fo = make_float2(0.5f * d.y, -0.5f * d.x);
            xo = cadd(fe, cmul(twW[kx], fo));
        }
        X[(long)kx * H + y0 + g] = cscale(xo, scale);
    }
}

// Load a G-row tile and run the inverse real FFT of this thread's row:
// on return v holds (f[2n], f[2n+1]) at the positions slot_pos<PL, RLI>.
template <typename C>
__device__ __forceinline__ void tile_c2r(float2* v, float2* lds, const float2* __restrict__ X, int H, int y0,
                                         int tid, int seq, int t, const RowGeom& g)
{
    float2* row = lds + seq * C::L;
    load_tile<C>(lds, X, H, y0, tid);
    __syncthreads();
    c2r_pre<C>(v, row, t, g.twW);
    seq_sync<typename C::PL>();
    fft_regs<typename C::PL, +1>(v, row, t, g.twN);
}

// ------------------------------------------------------------------ pass B
template <typename SAMPLE, typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS) void row_ratio_kernel(
    RowGeom g, float2* __restrict__ X, long xstride, const SAMPLE* __restrict__ data, long dstride, int nplanes)
{
    using C = RowCfg<PL>;
    __shared__ float2 lds[C::G * C::L];
    const int nrg = g.H / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    const int y0 = rg * C::G;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    float2* Xp = X + p * xstride;
    float2* row = lds + seq * C::L;

    float2 v[C::P];
    tile_c2r<C>(v, lds, Xp, g.H, y0, tid, seq, t, g);

    // R = Q > 1 ? D / Q : Q  on the padded row (data pad map, swapped half-widths)
    const SAMPLE* drow = data + p * dstride + (long)pad_index(y, g.dataH, g.dkY) * g.dataW;
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int n = slot_pos<PL, C::RLI>(t, i);
        const float q0 = v[i].x, q1 = v[i].y;
        const float d0 = (float)drow[pad_index(2 * n, g.dataW, g.dkX)];
        const float d1 = (float)drow[pad_index(2 * n + 1, g.dataW, g.dkX)];
        // hardware reciprocal (1 ulp) instead of the IEEE division sequence
        v[i].x = q0 > 1.0f ? d0 * __builtin_amdgcn_rcpf(q0) : q0;
        v[i].y = q1 > 1.0f ? d1 * __builtin_amdgcn_rcpf(q1) : q1;
    }
    fft_regs<PL, -1>(v, row, t, g.twN);
    r2c_post_store<C>(v, lds, Xp, g.H, y0, tid, seq, t, g.twW, 1.0f);
}

// ------------------------------------------------------------------ pass D
template <typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS) void row_update_kernel(
    RowGeom g, float2* __restrict__ XE, const float2* __restrict__ XC, long xstride,
    const float* __restrict__ weights, int nplanes)
{
    using C = RowCfg<PL>;
    __shared__ float2 lds[C::G * C::L];
    const int nrg = g.H / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    const int y0 = rg * C::G;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    float2* XEp = XE + p * xstride;
    float2* row = lds + seq * C::L;

    // current estimate E = C2R(XE)  (XE carries the 1/W normalisation)
    float2 e[C::P];
    tile_c2r<C>(e, lds, XEp, g.H, y0, tid, seq, t, g);
    // every row must be done with the tile before it is overwritten
    __syncthreads();
    // correction C = C2R(XC)
    float2 v[C::P];
    tile_c2r<C>(v, lds, XC + p * xstride, g.H, y0, tid, seq, t, g);

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
    r2c_post_store<C>(v, lds, XEp, g.H, y0, tid, seq, t, g.twW, 1.0f / (float)g.W);
}

// ------------------------------------------------------------------ egress
template <typename SAMPLE, typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS) void row_final_kernel(
    RowGeom g, const float2* __restrict__ XE, long xstride, SAMPLE* __restrict__ out, float* __restrict__ est,
    long ostride, int nplanes)
{
    using C = RowCfg<PL>;
    __shared__ float2 lds[C::G * C::L];
    const int nrg = (g.dataH + C::G - 1) / C::G;
    const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
    const int rg = lid % nrg, p = lid / nrg;
    const int y0 = rg * C::G;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;

    float2 v[C::P];
    tile_c2r<C>(v, lds, XE + p * xstride, g.H, y0, tid, seq, t, g);
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
template <typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS) void row_init_kernel(
    RowGeom g, float2* __restrict__ XE, long xstride, const float* __restrict__ e0, long estride, int nplanes)
{
    using C = RowCfg<PL>;
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
    r2c_post_store<C>(v, lds, XE + p * xstride, g.H, y0, tid, seq, t, g.twW, 1.0f / (float)g.W);
}

__global__ void const_init_kernel(float2* __restrict__ XE, long xstride, long per_plane, int H, float value)
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

template <typename SAMPLE>
void launch_row_ratio(const RowGeom& g, float2* X, long xstride, const SAMPLE* data, long dstride, int nplanes,
                      int variant, hipStream_t s)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        row_ratio_kernel<SAMPLE, PL><<<row_grid<PL>(g.H, nplanes), RowCfg<PL>::THREADS, 0, s>>>(
            g, X, xstride, data, dstride, nplanes);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_row_update(const RowGeom& g, float2* XE, const float2* XC, long xstride, const float* weights,
                       int nplanes, int variant, hipStream_t s)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        row_update_kernel<PL><<<row_grid<PL>(g.H, nplanes), RowCfg<PL>::THREADS, 0, s>>>(g, XE, XC, xstride,
                                                                                         weights, nplanes);
    });
    FMVD_LAUNCH_CHECK();
}

template <typename SAMPLE>
void launch_row_final(const RowGeom& g, const float2* XE, long xstride, SAMPLE* out, float* est, long ostride,
                      int nplanes, int variant, hipStream_t s)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        row_final_kernel<SAMPLE, PL><<<row_grid<PL>(g.dataH, nplanes), RowCfg<PL>::THREADS, 0, s>>>(
            g, XE, xstride, out, est, ostride, nplanes);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_row_init(const RowGeom& g, float2* XE, long xstride, const float* e0, long estride, int nplanes,
                     int variant, hipStream_t s)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        row_init_kernel<PL><<<row_grid<PL>(g.H, nplanes), RowCfg<PL>::THREADS, 0, s>>>(g, XE, xstride, e0,
                                                                                       estride, nplanes);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_const_init(const RowGeom& g, float2* XE, long xstride, float value, int nplanes, hipStream_t s)
{
    const long per_plane = (long)g.N * g.H;
    dim3 grid((unsigned)((per_plane + 255) / 256), (unsigned)nplanes);
    const_init_kernel<<<grid, 256, 0, s>>>(XE, xstride, per_plane, g.H, value);
    FMVD_LAUNCH_CHECK();
}

template void launch_row_ratio<uint16_t>(const RowGeom&, float2*, long, const uint16_t*, long, int, int,
                                         hipStream_t);
template void launch_row_ratio<uint8_t>(const RowGeom&, float2*, long, const uint8_t*, long, int, int,
                                        hipStream_t);
template void launch_row_final<uint16_t>(const RowGeom&, const float2*, long, uint16_t*, float*, long, int, int,
                                         hipStream_t);
template void launch_row_final<uint8_t>(const RowGeom&, const float2*, long, uint8_t*, float*, long, int, int,
                                        hipStream_t);

}  // namespace fmvd
