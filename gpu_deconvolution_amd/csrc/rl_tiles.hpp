// Tile-level device code of the RL passes, shared by the single-volume
// kernels (rl_row_kernels.hip, rl_col_kernels.hip) and the batched kernels
// over many volumes (rl_batch_kernels.hip).
#pragma once
#include "fft_device.hpp"
#include "rl_launch.hpp"

#include <type_traits>

namespace fmvd {

// An opaque copy of a per-lane value.  Addresses derived from it cannot be
// shared with the same addresses computed earlier (the tile's loads), so the
// compiler recomputes them at the stores (two instructions each) instead of
// holding one register per element for the whole kernel.
__device__ __forceinline__ int fresh(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

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

// LDS row pitch >= len with L = 16/G (mod 16).  The LDS serves 8-byte
// accesses a quarter-wavefront (16 lanes, 32 four-byte banks) per cycle; in
// the transposed tile access those lanes hold G consecutive rows times 16/G
// consecutive columns, which then land on 16 distinct bank pairs.
constexpr int row_pitch(int len, int G)
{
    const int want = G >= 16 ? 1 : (16 / G) % 16;  // one row (G = 1): any multiple of 16
    int L = len;
    while (L % 16 != want) ++L;
    return L;
}
// The same for the float buffer of split exchanges: 4-byte accesses are served
// 32 lanes at a time, which then hold G rows times 32/G columns, so the pitch
// (in floats) is 32/G (mod 32).
constexpr int row_pitch_split(int len, int G)
{
    const int want = G >= 32 ? 1 : (32 / G) % 32;
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
    // rows per workgroup: the plan's choice, else 256 threads where possible
    // with at least 8 rows so the transposed loads/stores move 64-byte runs;
    // halved while the tile would exceed the 160 KiB of LDS a workgroup may
    // hold or 1024 threads (long rows: 4 rows of N >= 2304)
    static constexpr int G0 = PL::G ? PL::G : ((256 / T) > 16 ? 16 : ((256 / T) < 8 ? 8 : (256 / T)));
    // LDS element: float2, or float for split exchanges (the tile is staged
    // and read back one part at a time as well)
    using E = typename PL::E;
    static constexpr int pitch(int g)
    {
        return PL::SPLIT ? row_pitch_split(lds_len<N>(), g) : row_pitch(lds_len<N>(), g);
    }
    static constexpr int fit_rows(int g)
    {
        while (g > 1 && ((long)g * pitch(g) * (long)sizeof(E) > 160L * 1024 || g * T > 1024)) g /= 2;
        return g;
    }
    static constexpr int G = fit_rows(G0);
    static constexpr int L = pitch(G);
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

// Tile element k of thread tid (q = tid + k * THREADS) in the spectrum
// layout of xpos: run j = q / (2G), row g = (q % (2G)) / 2, kx = 2j + (q & 1),
// so that 2G consecutive lanes move one 16G-byte run.  For even T a thread's
// elements are kx = kx0 + k T at global offsets xpos(kx0, y0 + g) + k T H.
template <typename C> __device__ __forceinline__ void tile_elem(int tid, int k, int& g, int& kx)
{
    const int q = tid + k * C::THREADS, e = q % (2 * C::G);
    g = e >> 1;
    kx = 2 * (q / (2 * C::G)) + (e & 1);
}

// Transposed tile load: lds[g][kx] <- X(kx, y0 + g).  fetch_tile issues the
// global loads into registers and stage_tile writes them to LDS, so that a
// caller can put further loads in flight between the two: the wait before
// the LDS writes then covers only the tile.
template <typename C>
__device__ __forceinline__ void fetch_tile(float2* pre, const float2* __restrict__ X, int H, int y0, int tid)
{
    if constexpr (C::T % 2 == 0) {
        int g, kx0;
        tile_elem<C>(tid, 0, g, kx0);
        const unsigned off0 = (unsigned)xpos(kx0, y0 + g, H), step = (unsigned)(C::T * H);
#pragma unroll
        for (int k = 0; k < C::P; ++k) pre[k] = X[off0 + k * step];
    } else {
#pragma unroll
        for (int k = 0; k < C::P; ++k) {
            int g, kx;
            tile_elem<C>(tid, k, g, kx);
            pre[k] = X[(unsigned)xpos(kx, y0 + g, H)];
        }
    }
}
template <typename C, int COMP = 0>
__device__ __forceinline__ void stage_tile(typename Comp<COMP>::E* lds, const float2* pre, int tid)
{
    if constexpr (C::T % 16 == 0) {
        int g, kx0;
        tile_elem<C>(tid, 0, g, kx0);
        typename Comp<COMP>::E* dst = lds + g * C::L + lds_pad(kx0);
#pragma unroll
        for (int k = 0; k < C::P; ++k) {
            float2 x = pre[k];
            dst[k * pad_c<C::T>()] = Comp<COMP>::of(x);
        }
    } else {
#pragma unroll
        for (int k = 0; k < C::P; ++k) {
            int g, kx;
            tile_elem<C>(tid, k, g, kx);
            float2 x = pre[k];
            lds[g * C::L + lds_pad(kx)] = Comp<COMP>::of(x);
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

// exp(-2 pi i (p + c) / W) for a compile-time c, from w0 = exp(-2 pi i p / W)
// (one load per thread) times the constant exp(-2 pi i c / W): the tables'
// loads, not their arithmetic, bound the row passes.
template <int N, int c> __device__ __forceinline__ float2 twiddle_w(float2 w0)
{
    if constexpr (c == 0) return w0;
    else return cmul(w0, ct_tw<-1, (long)c, 2L * N>());
}

template <typename C>
__device__ __forceinline__ void c2r_pre(float2* v, const float2* row, int t, const float2* __restrict__ twW)
{
    using PL = typename C::PL;
    if constexpr (affine_layout<C, C::RFI>()) {
        const int pt = lds_pad(t), mt = lds_mirror(t);
        const float2 w0 = twW[t];
        static_for<0, C::P>([&]<int i>() {
            constexpr int M = C::N / C::RFI;
            constexpr int c = (i / C::RFI) * C::T + (i % C::RFI) * M;  // slot_pos = t + c
            const float2 xk = row[pt + c + c / 16];
            if constexpr (i == 0) {
                const int im = t == 0 ? pt : C::N + C::N / 16 + mt;
                const float2 z = c2r_combine(xk, row[im], w0);
                v[i] = t == 0 ? make_float2(xk.x + xk.y, xk.x - xk.y) : z;
            } else {
                v[i] = c2r_combine(xk, row[(C::N - c) + (C::N - c) / 16 + mt], twiddle_w<C::N, c>(w0));
            }
        });
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

// Split exchanges: stage the tile that v holds (as fetch_tile loaded it) one
// part at a time and form the same inverse-FFT input as stage_tile + c2r_pre,
// with the same operations: after the real parts, a.x = re X[p] + re X[N-p]
// and d.x = re X[p] - re X[N-p]; after the imaginary parts, a.y and d.y; then
// Z'[p] = (a.x - b.y, a.y + b.x), b = d * conj(w).
template <typename C>
__device__ __forceinline__ void stage_c2r_pre_split(float2* v, float* lds, int tid, int seq, int t,
                                                    const float2* __restrict__ twW)
{
    static_assert(affine_layout<C, C::RFI>() && C::T % 16 == 0, "split exchanges need the affine layout");
    constexpr int M = C::N / C::RFI;
    const float* row = lds + seq * C::L;
    const int pt = lds_pad(t), mt = lds_mirror(t);
    float ax[C::P], dx[C::P], x0 = 0.0f;
    stage_tile<C, 1>(lds, v, tid);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int c = (i / C::RFI) * C::T + (i % C::RFI) * M;
        const float xk = row[pt + c + c / 16];
        const float xm = i == 0 ? row[t == 0 ? pt : C::N + C::N / 16 + mt] : row[(C::N - c) + (C::N - c) / 16 + mt];
        if (i == 0) x0 = xk;
        ax[i] = xk + xm;
        dx[i] = xk - xm;
    }
    __syncthreads();
    stage_tile<C, 2>(lds, v, tid);
    const float2 w0 = twW[t];
    __syncthreads();
    static_for<0, C::P>([&]<int i>() {
        constexpr int c = (i / C::RFI) * C::T + (i % C::RFI) * M;
        const float yk = row[pt + c + c / 16];
        const float ym = i == 0 ? row[t == 0 ? pt : C::N + C::N / 16 + mt] : row[(C::N - c) + (C::N - c) / 16 + mt];
        const float2 b = cmulc(make_float2(dx[i], yk + ym), twiddle_w<C::N, c>(w0));
        const float2 z = make_float2(ax[i] - b.y, (yk - ym) + b.x);
        v[i] = (i == 0 && t == 0) ? make_float2(x0 + yk, x0 - yk) : z;
    });
}

// R2C post-processing and transposed store: v holds Z (forward FFT of the
// packed row) at the forward's exit layout.  X[k] = (Z[k]+conj Z[N-k])/2 +
// exp(-2 pi i k/W) (Z[k]-conj Z[N-k])/(2i);  X[0] = (ReZ0+ImZ0, ReZ0-ImZ0).
// h = scale / 2: the output scale rides on the halving (no extra multiplies)
__device__ __forceinline__ float2 r2c_combine(float2 zk, float2 zmc, float2 w, float h)
{
    const float2 zm = cconj(zmc);
    const float2 fe = cscale(cadd(zk, zm), h);
    const float2 d = csub(zk, zm);
    const float2 fo = make_float2(h * d.y, -h * d.x);
    return cadd(fe, cmul(w, fo));
}

// Split exchanges: the same, one part at a time, with the same operations
// (s = Z[k] + conj Z[N-k] and d = Z[k] - conj Z[N-k] by parts).
template <typename C>
__device__ __forceinline__ void r2c_post_store_split(float2* v, float* lds, float2* __restrict__ X, int H, int y0,
                                                     int tid, int seq, int t, const float2* __restrict__ twW,
                                                     float scale)
{
    static_assert(affine_layout<C, C::RLF>() && C::T % 16 == 0, "split exchanges need the affine layout");
    constexpr int M = C::N / C::RLF;
    const float h = 0.5f * scale;
    float* dst = lds + seq * C::L + lds_pad(t);
    int g, kx0;
    tile_elem<C>(fresh(tid), 0, g, kx0);
    const float* r = lds + g * C::L;
    const unsigned off0 = (unsigned)xpos(kx0, y0 + g, H), step = (unsigned)(C::T * H);
    const int pk = lds_pad(kx0), mk = lds_mirror(kx0);
    float sx[C::P], dx[C::P], z0 = 0.0f;
    const float2 w0 = twW[kx0];
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int c = (i / C::RLF) * C::T + (i % C::RLF) * M;
        dst[c + c / 16] = v[i].x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::P; ++k) {
        const int c = k * C::T;
        const float zk = r[pk + c + c / 16];
        const float zm = k == 0 ? r[kx0 == 0 ? pk : C::N + C::N / 16 + mk] : r[(C::N - c) + (C::N - c) / 16 + mk];
        if (k == 0) z0 = zk;
        sx[k] = zk + zm;
        dx[k] = zk - zm;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int c = (i / C::RLF) * C::T + (i % C::RLF) * M;
        dst[c + c / 16] = v[i].y;
    }
    __syncthreads();
    static_for<0, C::P>([&]<int k>() {
        constexpr int c = k * C::T;
        const float zk = r[pk + c + c / 16];
        const float zm = k == 0 ? r[kx0 == 0 ? pk : C::N + C::N / 16 + mk] : r[(C::N - c) + (C::N - c) / 16 + mk];
        const float2 fe = make_float2(sx[k] * h, (zk - zm) * h);
        const float2 fo = make_float2(h * (zk + zm), -h * dx[k]);
        const float2 z = cadd(fe, cmul(twiddle_w<C::N, c>(w0), fo));
        X[off0 + k * step] = (k == 0 && kx0 == 0) ? make_float2((z0 + zk) * scale, (z0 - zk) * scale) : z;
    });
}

template <typename C>
__device__ __forceinline__ void r2c_post_store(float2* v, typename C::E* lds, float2* __restrict__ X, int H,
                                               int y0, int tid, int seq, int t, const float2* __restrict__ twW,
                                               float scale)
{
    if constexpr (C::PL::SPLIT) {
        r2c_post_store_split<C>(v, lds, X, H, y0, tid, seq, t, twW, scale);
        return;
    } else {
    using PL = typename C::PL;
    const float h = 0.5f * scale;
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
    if constexpr (C::T % 16 == 0) {
        int g, kx0;
        tile_elem<C>(fresh(tid), 0, g, kx0);
        const float2* r = lds + g * C::L;
        const unsigned off0 = (unsigned)xpos(kx0, y0 + g, H), step = (unsigned)(C::T * H);
        const int pk = lds_pad(kx0), mk = lds_mirror(kx0);
        const float2 w0 = twW[kx0];
        static_for<0, C::P>([&]<int k>() {
            constexpr int c = k * C::T;
            const float2 zk = r[pk + c + c / 16];
            float2 xo;
            if constexpr (k == 0) {
                const int im = kx0 == 0 ? pk : C::N + C::N / 16 + mk;
                const float2 z = r2c_combine(zk, r[im], w0, h);
                xo = kx0 == 0 ? make_float2((zk.x + zk.y) * scale, (zk.x - zk.y) * scale) : z;
            } else {
                xo = r2c_combine(zk, r[(C::N - c) + (C::N - c) / 16 + mk], twiddle_w<C::N, c>(w0), h);
            }
            X[off0 + k * step] = xo;
        });
    } else {
#pragma unroll 4
        for (int k = 0; k < C::P; ++k) {
            int g, kx;
            tile_elem<C>(tid, k, g, kx);
            const float2* r = lds + g * C::L;
            const float2 zk = r[lds_pad(kx)];
            const float2 xo = kx == 0 ? make_float2((zk.x + zk.y) * scale, (zk.x - zk.y) * scale)
                                      : r2c_combine(zk, r[lds_pad(C::N - kx)], twW[kx], h);
            X[(unsigned)xpos(kx, y0 + g, H)] = xo;
        }
    }
    }
}

// The two values at x = 2n and 2n+1 (n = t + c) of one padded row, read
// through the pad map (fmvd_deconvolve_cuda.impl.cu:34-50, pad_index) and
// returned still packed (two samples in one word, or a float pair) so that
// the load can be issued long before the values are used.  c is the register
// slot's compile-time offset, so every thread of a sequence takes the same
// branch: all positions inside the data (one paired load when `pairs`), all
// in the region that maps to x = 0, or the per-lane map.
template <typename T> struct PairOf;
template <> struct PairOf<uint16_t> { using type = uint32_t; };
template <> struct PairOf<uint8_t> { using type = uint16_t; };
template <> struct PairOf<float> { using type = float2; };

template <typename T, typename PairT = typename PairOf<T>::type>
__device__ __forceinline__ PairT pack_pair(T a, T b)
{
    if constexpr (sizeof(T) == 4) return make_float2(a, b);
    else return (PairT)((unsigned)a | ((unsigned)b << (8 * sizeof(T))));
}

template <typename T, typename PairT = typename PairOf<T>::type>
__device__ __forceinline__ PairT load_padded_pair(const T* __restrict__ row, int n, int c, int span, int dataW,
                                                  int half, bool pairs)
{
    if (pairs && 2 * (c + span) <= dataW) {
        // samples are read once per view-iteration: non-temporal, so they do
        // not displace the spectra and estimates from the memory-side cache
        if constexpr (std::is_integral_v<PairT>) return __builtin_nontemporal_load(reinterpret_cast<const PairT*>(row) + n);
        else return reinterpret_cast<const PairT*>(row)[n];
    }
    if (2 * c >= dataW + half) return pack_pair<T>(row[0], row[0]);
    return pack_pair<T>(row[pad_index(2 * n, dataW, half)], row[pad_index(2 * n + 1, dataW, half)]);
}

template <typename T, typename PairT>
__device__ __forceinline__ void unpack_pair(PairT w, float& a, float& b)
{
    if constexpr (sizeof(T) == 4) {
        a = w.x;
        b = w.y;
    } else {
        constexpr unsigned bits = 8 * sizeof(T), mask = (1u << bits) - 1u;
        a = (float)(w & mask);
        b = (float)((unsigned)w >> bits);
    }
}

// Inverse real FFT of this thread's row from the tile that v holds (as
// fetch_tile loaded it): the tile is staged in LDS, the inverse FFT's input
// formed from it and transformed; on return v holds (f[2n], f[2n+1]) at the
// positions slot_pos<PL, RLI>.
template <typename C>
__device__ __forceinline__ void row_c2r(float2* v, typename C::E* lds, int tid, int seq, int t, const RowGeom& g)
{
    typename C::E* row = lds + seq * C::L;
    if constexpr (C::PL::SPLIT) {
        stage_c2r_pre_split<C>(v, lds, tid, seq, t, g.twW);
    } else {
        stage_tile<C>(lds, v, tid);
        __syncthreads();
        c2r_pre<C>(v, row, t, g.twW);
    }
    seq_sync<typename C::PL>();
    fft_regs<typename C::PL, +1>(v, row, t, g.twP[C::PL::variant]);
}

// ------------------------------------------------------------- row tiles
// Each works on the G rows [y0, y0+G) of one padded plane; lds holds
// G*RowCfg::L complex.  g carries the FFT extent (uniform over the launch)
// and the data extent and pad half-widths of the plane's volume.

// Pass B: X <- R2C(Q > 1 ? D / Q : Q) * scale, Q = C2R(X).  divide_kernel_{16,8},
// fmvd_deconvolve_cuda.impl.cu:159-176; the data is read unpadded through the
// data pad map (swapped half-widths, fmvd_deconvolve.impl.cpp:36-37).  scale
// is 1, or 1/W when pass C's kernel spectrum is pass A's (which carries 1/H only).
template <typename SAMPLE, typename PL>
__device__ __forceinline__ void ratio_tile(const RowGeom& g, float2* __restrict__ Xp,
                                           const SAMPLE* __restrict__ dplane, int y0, typename PL::E* lds,
                                           float scale)
{
    using C = RowCfg<PL>;
    using PairT = typename PairOf<SAMPLE>::type;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    typename C::E* row = lds + seq * C::L;
    float2 v[C::P];
    fetch_tile<C>(v, Xp, g.H, y0, tid);
    // the row's samples (streamed from HBM once per view-iteration) are loaded
    // behind the tile, ahead of the FFT, so their latency hides behind it
    const SAMPLE* drow = dplane + (unsigned)(pad_index(y, g.dataH, g.dkY) * g.dataW);
    const bool pairs = (g.dataW % 2 == 0) && ((uintptr_t)dplane % (2 * sizeof(SAMPLE)) == 0);
    PairT d[C::P];
    auto load_samples = [&]() {
#pragma unroll
        for (int i = 0; i < C::P; ++i) {
            const int c = (i / C::RLI) * C::T + (i % C::RLI) * (C::N / C::RLI);
            d[i] = load_padded_pair<SAMPLE>(drow, t + c, c, C::T, g.dataW, g.dkX, pairs);
        }
    };
    load_samples();
    row_c2r<C>(v, lds, tid, seq, t, g);
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const float q0 = v[i].x, q1 = v[i].y;
        float d0, d1;
        unpack_pair<SAMPLE>(d[i], d0, d1);
        // hardware reciprocal (1 ulp) instead of the IEEE division sequence
        v[i].x = q0 > 1.0f ? d0 * __builtin_amdgcn_rcpf(q0) : q0;
        v[i].y = q1 > 1.0f ? d1 * __builtin_amdgcn_rcpf(q1) : q1;
    }
    fft_regs<PL, -1>(v, row, t, g.twP[PL::variant]);
    r2c_post_store<C>(v, lds, Xp, g.H, y0, tid, seq, t, g.twW, scale);
}

// Pass D: C = C2R(XC); E <- E + w (E C - E) on the spatial estimate (one
// fftH x fftW float plane, kept in the spatial domain as the reference does);
// XE <- R2C(E) / W.  multiply32_kernel, fmvd_deconvolve_cuda.cu:374-392.
// XE may be XC (in place): the tile is in LDS before any of it is written.
template <typename PL>
__device__ __forceinline__ void update_tile(const RowGeom& g, float2* XEp,
                                            const float2* XCp, float* __restrict__ ESp,
                                            const float* __restrict__ weights, int y0, typename PL::E* lds)
{
    using C = RowCfg<PL>;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    typename C::E* row = lds + seq * C::L;
    float2 v[C::P];
    fetch_tile<C>(v, XCp, g.H, y0, tid);
    row_c2r<C>(v, lds, tid, seq, t, g);

    float2* erow = reinterpret_cast<float2*>(ESp) + (unsigned)(y * (g.W / 2));
    const float* wrow = weights + (unsigned)(pad_index(y, g.dataH, g.wkY) * g.dataW);
    const bool pairs = (g.dataW % 2 == 0) && ((uintptr_t)weights % 8 == 0);
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int c = (i / C::RLI) * C::T + (i % C::RLI) * (C::N / C::RLI);
        const float2 e = erow[t + c];
        float w0, w1;
        unpack_pair<float>(load_padded_pair<float>(wrow, t + c, c, C::T, g.dataW, g.wkX, pairs), w0, w1);
        // fmvd_deconvolve_cuda.cu:387-391 operation order
        float t0 = e.x * v[i].x, t1 = e.y * v[i].y;
        float c0 = t0 - e.x, c1 = t1 - e.y;
        c0 *= w0;
        c1 *= w1;
        v[i].x = e.x + c0;
        v[i].y = e.y + c1;
        erow[t + c] = v[i];
    }
    fft_regs<PL, -1>(v, row, t, g.twP[PL::variant]);
    r2c_post_store<C>(v, lds, XEp, g.H, y0, tid, seq, t, g.twW, 1.0f / (float)g.W);
}

// Warm-start ingest: ES <- pad(E0) (data pad map), XE <- R2C(ES) / W.
template <typename PL>
__device__ __forceinline__ void init_tile(const RowGeom& g, float2* __restrict__ XEp, float* __restrict__ ESp,
                                          const float* __restrict__ e0p, int y0, typename PL::E* lds)
{
    using C = RowCfg<PL>;
    const int tid = threadIdx.x, seq = tid / C::T, t = tid % C::T;
    const int y = y0 + seq;
    typename C::E* row = lds + seq * C::L;
    const float* erow = e0p + (unsigned)(pad_index(y, g.dataH, g.dkY) * g.dataW);
    float2* esrow = reinterpret_cast<float2*>(ESp) + (unsigned)(y * (g.W / 2));
    float2 v[C::P];
#pragma unroll
    for (int i = 0; i < C::P; ++i) {
        const int n = slot_pos<PL, C::RLI>(t, i);
        v[i].x = erow[pad_index(2 * n, g.dataW, g.dkX)];
        v[i].y = erow[pad_index(2 * n + 1, g.dataW, g.dkX)];
        esrow[n] = v[i];
    }
    fft_regs<PL, -1>(v, row, t, g.twP[PL::variant]);
    r2c_post_store<C>(v, lds, XEp, g.H, y0, tid, seq, t, g.twW, 1.0f / (float)g.W);
}

// ---------------------------------------------------------- column tile
constexpr int pow2_floor(int x)
{
    int p = 1;
    while (p * 2 <= x) p *= 2;
    return p;
}

template <typename PL_> struct ColCfg {
    using PL = PL_;
    static_assert(plan_valid<PL>(), "invalid FFT plan");
    // column lengths from the size rule (16 ... 8192 as listed in the column
    // kernels) have complex-exchange plans only
    static_assert(!PL::SPLIT, "column plans exchange complex values");
    static constexpr int H = PL::N;
    static constexpr int T = PL::T;
    static constexpr int P = PL::P;
    using E = typename PL::E;  // LDS element (float for split exchanges)
    static constexpr int L = lds_len<H>() + 1;
    // sequences per workgroup (a power of two, so it divides the row count):
    // aim for 256 threads, keep LDS <= ~64 KB
    static constexpr int SQ0 = (256 / T) < 1 ? 1 : (256 / T);
    static constexpr int SQ1 = SQ0 > 16 ? 16 : SQ0;
    static constexpr int SQ2 = 65536 / (L * (int)sizeof(E)) < 1 ? 1 : 65536 / (L * (int)sizeof(E));
    static constexpr int SQ = PL::G ? PL::G : pow2_floor(SQ1 < SQ2 ? SQ1 : SQ2);
    static constexpr int THREADS = SQ * T;
    static constexpr int RF = first_radix<PL, -1>();
    static constexpr int RL = last_radix<PL, -1>();
    static_assert(RL == first_radix<PL, +1>(), "forward exit layout must match inverse entry");
    static_assert(RF == last_radix<PL, +1>(), "inverse exit layout must match forward entry");
};

// v[i] *= Kr[y] at the inverse-entry positions y of thread t, with Kr[y]'s
// imaginary part times ksign (-1: the conjugate); Kr is one kx column of a
// kernel spectrum (contiguous along y).  The sign is a multiply, not a second
// code path, so every kernel that inlines this computes the same bits.
template <typename PL>
__device__ __forceinline__ void col_kmul(float2* v, const float2* __restrict__ Kr, int t, float ksign)
{
    constexpr int RL = last_radix<PL, -1>();
#pragma unroll
    for (int i = 0; i < PL::P; ++i) {
        const float2 k = Kr[slot_pos<PL, RL>(t, i)];
        v[i] = cmul(v[i], make_float2(k.x, ksign * k.y));
    }
}

// Passes A / C on the kx-rows [grp*SQ, grp*SQ+SQ) of one plane:
// Xout[kx] <- IFFT_y(FFT_y(Xin[kx]) * K[kx]); Xout may be Xin (each wave holds
// its whole sequence in registers before it stores).  The cuFFT column halves plus
// modulateAndNormalize of fmvd_deconvolve_cuda.cu:327-372 (normalisation
// folded into K on the host).  kx = 0 holds two real columns packed as
// DC + i*Nyquist: its spectrum is split with A[k] = (C[k] + conj C[-k])/2,
// B[k] = (C[k] - conj C[-k])/(2i), multiplied by the DC and Nyquist kernel
// columns and recombined as A*K0 + i*B*KN.
// ksign = -1: multiply by conj(K) instead.  The spectrum of a real kernel
// flipped about its centre is the conjugate of the kernel's spectrum, so pass
// C can read pass A's K.
template <typename PL>
__device__ __forceinline__ void col_tile(const ColGeom& g, const float2* Xin, float2* Xout,
                                         const float2* __restrict__ K, int grp, typename PL::E* lds, float ksign)
{
    using C = ColCfg<PL>;
    constexpr int H = C::H;
    const int seq = threadIdx.x / C::T, t = threadIdx.x % C::T;
    // wave-uniform (a sequence never spans two wavefronts here): keep it scalar
    // so the loads below address from one SGPR base with 32-bit lane offsets
    const int kx = C::T % 64 == 0 ? __builtin_amdgcn_readfirstlane(grp * C::SQ + seq) : grp * C::SQ + seq;
    // only a group of 16 sequences of at most 16 threads (columns up to 256)
    // can overhang N (= 8 at fftW = 16); such sequences synchronise within
    // their wavefront, so the overhanging ones just leave
    if (kx >= g.N) return;
    const float2* xin = Xin + (unsigned)xpos(kx, 0, H);  // stride 2 along y
    typename C::E* row = lds + seq * C::L;

    float2 v[C::P];
#pragma unroll
    for (int i = 0; i < C::P; ++i) v[i] = xin[2 * slot_pos<PL, C::RF>(t, i)];
    fft_regs<PL, -1>(v, row, t, g.twP[PL::variant]);

    if (grp == 0) {  // workgroup-uniform: this group holds the packed DC/Nyquist sequence
        if (seq == 0) {
#pragma unroll
            for (int i = 0; i < C::P; ++i) row[lds_pad(slot_pos<PL, C::RL>(t, i))] = v[i];
        }
        seq_sync<PL>();
        if (seq == 0) {
            const float2* K0 = K;                         // kx = 0
            const float2* KN = K + (unsigned)(g.N * H);  // kx = N, contiguous
#pragma unroll
            for (int i = 0; i < C::P; ++i) {
                const int k = slot_pos<PL, C::RL>(t, i);
                const float2 ck = row[lds_pad(k)];
                const float2 cm = cconj(row[lds_pad(k == 0 ? 0 : H - k)]);
                const float2 a = cscale(cadd(ck, cm), 0.5f);
                const float2 d = csub(ck, cm);
                const float2 b = make_float2(0.5f * d.y, -0.5f * d.x);
                const float2 k0 = K0[k], kn = KN[k];
                const float2 ak = cmul(a, make_float2(k0.x, ksign * k0.y));
                const float2 bk = cmul(b, make_float2(kn.x, ksign * kn.y));
                v[i] = make_float2(ak.x - bk.y, ak.y + bk.x);
            }
        } else {
            col_kmul<PL>(v, K + (unsigned)(kx * H), t, ksign);
        }
        seq_sync<PL>();
    } else {
        col_kmul<PL>(v, K + (unsigned)(kx * H), t, ksign);
    }

    fft_regs<PL, +1>(v, row, t, g.twP[PL::variant]);
    float2* xo = Xout + (unsigned)xpos(kx, 0, H);
    const int ts = fresh(t);
#pragma unroll
    for (int i = 0; i < C::P; ++i) xo[2 * slot_pos<PL, C::RF>(ts, i)] = v[i];
}

}  // namespace fmvd
