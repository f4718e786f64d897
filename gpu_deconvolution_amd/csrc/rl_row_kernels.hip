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

// LDS addressing.  lds_pad(base + c) = lds_pad(base) + c + c/16 whenever c is a
// multiple of 16, and lds_pad(A - t) = A + A/16 + lds_mirror(t) whenever A is;
// the affine forms let every access of a thread share one address register
// and use immediate offsets.
__device__ __forceinline__ int lds_mirror(int t) { return -t + ((-t) >> 4); }
template <int C> constexpr int pad_c() { return C + C / 16; }

// The affine forms apply when the sequence's thread count and every slot
// offset of the given register layout are multiples of 16.
template <typename C, int R> constexpr bool affine_layout()
{
    return C::T % 16 == 0 && (C::N / R) % 16 == 0 && C::N % 16 == 0;
}

// Transposed tile load: lds[g][kx] <- X[kx][y0+g].  Thread tid owns
// g = tid % G and kx = tid / G + k*T (k < P).
template <typename C>
__device__ __forceinline__ void load_tile(float2* lds, const float2* __restrict__ X, int H, int y0, int tid)
{
    const int g = tid % C::G, kx0 = tid / C::G;
    if constexpr (C::T % 16 == 0) {
        const unsigned off0 = (unsigned)(kx0 * H + y0 + g), step = (unsigned)(C::T * H);
        float2 tmp[C::P];
#pragma unroll
        for (int k = 0; k < C::P; ++k) tmp[k] = X[off0 + k * step];
        float2* dst = lds + g * C::L + lds_pad(kx0);
#pragma unroll
        for (int k = 0; k < C::P; ++k) dst[k * pad_c<C::T>()] = tmp[k];
    } else {
#pragma unroll 4
        for (int k = 0; k < C::P; ++k) {
            const int kx = kx0 + k * C::T;
            lds[g * C::L + lds_pad(kx)] = X[(unsigned)(kx * H + y0 + g)];
        }
    }
}

// C2R pre-processing: from the packed half-spectrum of one row (in LDS),
// form the length-N complex input of the inverse FFT at this thread's
// first-pass positions:  Z'[p] = A + iB,  A = X[p] + conj(X[N-p]),
// B = (X[p] - conj(X[N-p])) exp(+2 pi i p / W);  Z'[0] = (DC+Ny, DC-Ny).
__device__ __forceinline__ float2 c2r_combine(float2 xk, float2 xmc, float2 w)
{
    const float2 xm = cconj(xmc);
    const float2 a = cadd(xk, xm);
    const float2 b = cmulc(csub(xk, xm), w);
    return make_float2(a.x - b.y, a.y + b.x);
}

template <typename C>
__device__ __forceinline__ void c2r_pre(float2* v, const float2* row, int t, const float2* __restrict__ twW)
{
    using PL = typename C::PL;
    if constexpr (affine_layout<C, C::RFI>()) {
        const int pt = lds_pad(t), mt = lds_mirror(t);
#pragma unroll
        for (int i = 0; i < C::P; ++i) {
            constexpr int M = C::N / C::RFI;
            const int c = (i / C::RFI) * C::T + (i % C::RFI) * M;  // slot_pos = t + c
            const float2 xk = row[pt + c + c / 16];
            if (i == 0) {
                const int im = t == 0 ? pt : C::N + C::N / 16 + mt;
                const float2 z = c2r_combine(xk, row[im], twW[t]);
                v[i] = t == 0 ? make_float2(xk.x + xk.y, xk.x - xk.y) : z;
            } else {
                v[i] = c2r_combine(xk, row[(C::N - c) + (C::N - c) / 16 + mt], twW[t + c]);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < C::P; ++i) {
            const int p = slot_pos<PL, C::RFI>(t, i);
            const float2 xk = row[lds_pad(p)];
            if (p == 0) v[i] = make_float2(xk.x + xk.y, xk.x - xk.y);
            else v[i] = c2r_combine(xk, row[lds_pad(C::N - p)], twW[p]);
        }
    }
}

// R2C post-processing and transposed store: v holds Z (forward FFT of the
// packed row) at the forward's exit layout.  X[k] = (Z[k]+conj Z[N-k])/2 +
// exp(-2 pi i k/W) (Z[k]-conj Z[N-k])/(2i);  X[0] = (ReZ0+ImZ0, ReZ0-ImZ0).
__device__ __forceinline__ float2 r2c_combine(float2 zk, float2 zmc, float2 w)
{
    const float2 zm = cconj(zmc);
    const float2 fe = cscale(cadd(zk, zm), 0.5f);
    const float2 d = csub(zk, zm);
    const float2 fo = make_float2(0.5f * d.y, -0.5f * d.x);
    return cadd(fe, cmul(w, fo));
}

template <typename C>
__device__ __forceinline__ void r2c_post_store(const float2* v, float2* lds, float2* __restrict__ X, int H,
                                               int y0, int tid, int seq, int t, const float2* __restrict__ twW,
                                               float scale)
{
    using PL = typename C::PL;
    float2* row = lds + seq * C::L;
    if constexpr (affine_layout<C, C::RLF>()) {
        constexpr int M = C::N / C::RLF;
        float2* dst = row + lds_pad(t);
#pragma unroll
        for (int i = 0; i < C::P; ++i) {
            const int c = (i / C::RLF) * C::T + (i % C::RLF) * M;
            dst[c + c / 16] = v[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < C::P; ++i) row[lds_pad(slot_pos<PL, C::RLF>(t, i))] = v[i];
    }
    __syncthreads();
    const int g = tid % C::G, kx0 = tid / C::G;
    const float2* r = lds + g * C::L;
    if constexpr (C::T % 16 == 0) {
        const unsigned off0 = (unsigned)(kx0 * H + y0 + g), step = (unsigned)(C::T * H);
        const int pk = lds_pad(kx0), mk = lds_mirror(kx0);
#pragma unroll
        for (int k = 0; k < C::P; ++k) {
            const int c = k * C::T;
            const float2 zk = r[pk + c + c / 16];
            float2 xo;
            if (k == 0) {
                const int im = kx0 == 0 ? pk : C::N + C::N / 16 + mk;
                const float2 z = r2c_combine(zk, r[im], twW[kx0]);
                xo = kx0 == 0 ? make_float2(zk.x + zk.y, zk.x - zk.y) : z;
            } else {
                xo = r2c_combine(zk, r[(C::N - c) + (C::N - c) / 16 + mk], twW[kx0 + c]);
            }
            X[off0 + k * step] = cscale(xo, scale);
        }
    } else {
#pragma unroll 4
        for (int k = 0; k < C::P; ++k) {
            const int kx = kx0 + k * C::T;
            const float2 zk = r[lds_pad(kx)];
            const float2 xo = kx == 0 ? make_float2(zk.x + zk.y, zk.x - zk.y)
                                      : r2c_combine(zk, r[lds_pad(C::N - kx)], twW[kx]);
            X[(unsigned)(kx * H + y0 + g)] = cscale(xo, scale);
        }
    }
}

// Reads the two values at x = 2n and 2n+1 (n = t + c) of one padded row
// through the pad map (fmvd_deconvolve_cuda.impl.cu:34-50, pad_index).  c is
// the register slot's compile-time offset, so every thread of a sequence
// takes the same branch: all positions inside the data (one paired load when
// `pairs`), all in the region that maps to x = 0, or the per-lane map.
template <typename T, typename PairT>
__device__ __forceinline__ void load_padded_pair(const T* __restrict__ row, int n, int c, int span, int dataW,
                                                 int half, bool pairs, float& a, float& b)
{
    if (pairs && 2 * (c + span) <= dataW) {
        const PairT w = reinterpret_cast<const PairT*>(row)[n];
        if constexpr (sizeof(T) == 4) {
            a = w.x;
            b = w.y;
        } else {
            constexpr unsigned bits = 8 * sizeof(T), mask = (1u << bits) - 1u;
            a = (float)(w & mask);
            b = (float)((unsigned)w >> bits);
        }
    } else if (2 * c >= dataW + half) {
        a = b = (float)row[0];
    } else {
        a = (float)row[pad_index(2 * n, dataW, half)];
        b = (float)row[pad_index(2 * n + 1, dataW, half)];
    }
}
template <typename T> struct PairOf;
template <> struct PairOf<uint16_t> { using type = uint32_t; };
template <> struct PairOf<uint8_t> { using type = uint16_t; };
template <> struct PairOf<float> { using type = float2; };

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
    fft_regs<typename C::PL, +1>(v, row, t, g.twP[C::PL::variant]);
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
    const SAMPLE* dplane = data + p * dstride;
    const SAMPLE* drow = dplane + (unsigned)(pad_index(y, g.dataH, g.dkY) * g.dataW);
    const bool pairs = (g.dataW % 2 == 0) && ((uintptr_t)dplane % (2 * sizeof(SAMPLE)) == 0);
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int c = (i / C::RLI) * C::T + (i % C::RLI) * (C::N / C::RLI);
        const int n = t + c;
        const float q0 = v[i].x, q1 = v[i].y;
        float d0, d1;
        load_padded_pair<SAMPLE, typename PairOf<SAMPLE>::type>(drow, n, c, C::T, g.dataW, g.dkX, pairs, d0, d1);
        // hardware reciprocal (1 ulp) instead of the IEEE division sequence
        v[i].x = q0 > 1.0f ? d0 * __builtin_amdgcn_rcpf(q0) : q0;
        v[i].y = q1 > 1.0f ? d1 * __builtin_amdgcn_rcpf(q1) : q1;
    }
    fft_regs<PL, -1>(v, row, t, g.twP[PL::variant]);
    r2c_post_store<C>(v, lds, Xp, g.H, y0, tid, seq, t, g.twW, 1.0f);
}

// ------------------------------------------------------------------ pass D
// The estimate is carried in the spatial domain (ES, one fftH x fftW float
// plane per z, as the reference keeps it), so this pass needs one C2R and one
// R2C per row.
template <typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS) void row_update_kernel(
    RowGeom g, float2* __restrict__ XE, const float2* __restrict__ XC, float* __restrict__ ES, long xstride,
    long estride, const float* __restrict__ weights, int nplanes)
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

    // correction C = C2R(XC)
    float2 v[C::P];
    tile_c2r<C>(v, lds, XC + p * xstride, g.H, y0, tid, seq, t, g);

    // E <- E + w (E*C - E)   (fmvd_deconvolve_cuda.cu:387-391 operation order)
    float2* erow = reinterpret_cast<float2*>(ES + p * estride) + (unsigned)(y * (g.W / 2));
    const float* wrow = weights + (unsigned)(pad_index(y, g.dataH, g.wkY) * g.dataW);
    const bool pairs = (g.dataW % 2 == 0) && ((uintptr_t)weights % 8 == 0);
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int c = (i / C::RLI) * C::T + (i % C::RLI) * (C::N / C::RLI);
        const int n = t + c;
        const float2 e = erow[n];
        float w0, w1;
        load_padded_pair<float, float2>(wrow, n, c, C::T, g.dataW, g.wkX, pairs, w0, w1);
        float t0 = e.x * v[i].x, t1 = e.y * v[i].y;
        float c0 = t0 - e.x, c1 = t1 - e.y;
        c0 *= w0;
        c1 *= w1;
        v[i].x = e.x + c0;
        v[i].y = e.y + c1;
        erow[n] = v[i];
    }
    fft_regs<PL, -1>(v, row, t, g.twP[PL::variant]);
    r2c_post_store<C>(v, lds, XE + p * xstride, g.H, y0, tid, seq, t, g.twW, 1.0f / (float)g.W);
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
__global__ __launch_bounds__(RowCfg<PL>::THREADS) void row_init_kernel(
    RowGeom g, float2* __restrict__ XE, float* __restrict__ ES, long xstride, long estride,
    const float* __restrict__ e0, long e0stride, int nplanes)
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
    const float* erow = e0 + p * e0stride + (long)pad_index(y, g.dataH, g.dkY) * g.dataW;
    float2* esrow = reinterpret_cast<float2*>(ES + p * estride + (long)y * g.W);
    float2 v[C::P];
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int n = slot_pos<PL, C::RLI>(t, i);
        v[i].x = erow[pad_index(2 * n, g.dataW, g.dkX)];
        v[i].y = erow[pad_index(2 * n + 1, g.dataW, g.dkX)];
        esrow[n] = v[i];
    }
    fft_regs<PL, -1>(v, row, t, g.twP[PL::variant]);
    r2c_post_store<C>(v, lds, XE + p * xstride, g.H, y0, tid, seq, t, g.twW, 1.0f / (float)g.W);
}

__global__ void const_init_kernel(float2* __restrict__ XE, float* __restrict__ ES, long xstride, long estride,
                                  long per_plane, int H, float value)
{
    // per_plane = N*H complex of XE = H*W/2 float2 of ES
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= per_plane) return;
    XE[p * xstride + i] = (i < H) ? make_float2(value, 0.0f) : make_float2(0.0f, 0.0f);
    reinterpret_cast<float2*>(ES + p * estride)[i] = make_float2(value, value);
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

std::vector<float2> row_pass_twiddles(int N, int variant)
{
    std::vector<float2> t;
    dispatch_row(N, variant, [&]<typename PL>() { t = pass_twiddle_table<PL>(); });
    return t;
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

void launch_row_update(const RowGeom& g, float2* XE, const float2* XC, float* ES, long xstride, long estride,
                       const float* weights, int nplanes, int variant, hipStream_t s)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        row_update_kernel<PL><<<row_grid<PL>(g.H, nplanes), RowCfg<PL>::THREADS, 0, s>>>(
            g, XE, XC, ES, xstride, estride, weights, nplanes);
    });
    FMVD_LAUNCH_CHECK();
}

template <typename SAMPLE>
void launch_final(const RowGeom& g, const float* ES, long estride, SAMPLE* out, float* est, long ostride,
                  int nplanes, hipStream_t s)
{
    if (nplanes <= 0) return;
    dim3 grid((unsigned)((g.dataW + 255) / 256), (unsigned)g.dataH, (unsigned)nplanes);
    final_kernel<SAMPLE><<<grid, 256, 0, s>>>(ES, estride, g.W, g.dataH, g.dataW, out, est, ostride);
    FMVD_LAUNCH_CHECK();
}

void launch_row_init(const RowGeom& g, float2* XE, float* ES, long xstride, long estride, const float* e0,
                     long e0stride, int nplanes, int variant, hipStream_t s)
{
    dispatch_row(g.N, variant, [&]<typename PL>() {
        row_init_kernel<PL><<<row_grid<PL>(g.H, nplanes), RowCfg<PL>::THREADS, 0, s>>>(
            g, XE, ES, xstride, estride, e0, e0stride, nplanes);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_const_init(const RowGeom& g, float2* XE, float* ES, long xstride, long estride, float value,
                       int nplanes, hipStream_t s)
{
    const long per_plane = (long)g.N * g.H;
    dim3 grid((unsigned)((per_plane + 255) / 256), (unsigned)nplanes);
    const_init_kernel<<<grid, 256, 0, s>>>(XE, ES, xstride, estride, per_plane, g.H, value);
    FMVD_LAUNCH_CHECK();
}

template void launch_row_ratio<uint16_t>(const RowGeom&, float2*, long, const uint16_t*, long, int, int,
                                         hipStream_t);
template void launch_row_ratio<uint8_t>(const RowGeom&, float2*, long, const uint8_t*, long, int, int,
                                        hipStream_t);
template void launch_final<uint16_t>(const RowGeom&, const float*, long, uint16_t*, float*, long, int,
                                     hipStream_t);
template void launch_final<uint8_t>(const RowGeom&, const float*, long, uint8_t*, float*, long, int, hipStream_t);

}  // namespace fmvd
