// Batched RL passes over many independent volumes of uneven size (gfx950).
//
// Volumes whose padded planes share one FFT extent form a class; a class's
// planes live back to back in its XE / XT / ES buffers.  One launch runs one
// pass (column, ratio, column, update) of one view index over every plane of
// every volume of the class that has that view: the block id is mapped to
// (volume, plane, tile) through a launch list in device memory, and the tile
// code is the single-volume kernels' (rl_tiles.hpp) with the volume's data
// extent, pad half-widths and operands substituted.
#include "fft_device.hpp"
#include "hip_check.hpp"
#include "rl_batch.hpp"
#include "rl_launch.hpp"
#include "rl_tiles.hpp"

namespace fmvd {

// Launch list layout (device ints): ids[n] then starts[n + 1] (block offsets;
// starts[n] = total blocks).  Every thread of a block finds the same entry.
__device__ __forceinline__ int find_entry(const int* __restrict__ list, int n, int lid)
{
    const int* starts = list + n;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= lid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ RowGeom seg_geom(const RowGeom& g, const BatchSeg& S)
{
    RowGeom r = g;
    r.dataH = S.dataH;
    r.dataW = S.dataW;
    r.dkY = S.dkY;
    r.dkX = S.dkX;
    r.wkY = S.wkY;
    r.wkX = S.wkX;
    return r;
}

template <typename SAMPLE, typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS, PL::WPE) void batch_ratio_kernel(RowGeom g, const BatchSeg* __restrict__ segs,
                                                                          const int* __restrict__ list, int n, int view,
                                                                          float2* __restrict__ XT, long xstride)
{
    using C = RowCfg<PL>;
    __shared__ typename C::E lds[C::G * C::L];
    const int lid = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int e = find_entry(list, n, lid);
    const BatchSeg& S = segs[list[e]];
    const int local = lid - list[n + e], nrg = g.H / C::G;
    const int rg = local % nrg, pl = local / nrg;
    const SAMPLE* d = static_cast<const SAMPLE*>(S.data[view]) + (long)pl * S.dstride;
    ratio_tile<SAMPLE, PL>(seg_geom(g, S), XT + (long)(S.plane0 + pl) * xstride, d, rg * C::G, lds, S.rscale);
}

template <typename PL>
__global__ __launch_bounds__(RowCfg<PL>::THREADS, PL::WPE) void batch_update_kernel(RowGeom g, const BatchSeg* __restrict__ segs,
                                                                           const int* __restrict__ list, int n, int view,
                                                                           float2* XE,
                                                                           const float2* XT,
                                                                           float* __restrict__ ES, long xstride,
                                                                           long estride)
{
    using C = RowCfg<PL>;
    __shared__ typename C::E lds[C::G * C::L];
    const int lid = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int e = find_entry(list, n, lid);
    const BatchSeg& S = segs[list[e]];
    const int local = lid - list[n + e], nrg = g.H / C::G;
    const int rg = local % nrg, pl = local / nrg;
    const long p = S.plane0 + pl;
    update_tile<PL>(seg_geom(g, S), XE + p * xstride, XT + p * xstride, ES + p * estride, S.w[view], rg * C::G,
                    lds);
}

template <typename PL>
__global__ __launch_bounds__(ColCfg<PL>::THREADS, PL::WPE) void batch_col_kernel(ColGeom g, const BatchSeg* __restrict__ segs,
                                                                        const int* __restrict__ list, int n, int view,
                                                                        int hat, const float2* Xin,
                                                                        float2* Xout, long xstride)
{
    using C = ColCfg<PL>;
    __shared__ typename C::E lds[C::SQ * C::L];
    const int lid = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int e = find_entry(list, n, lid);
    const BatchSeg& S = segs[list[e]];
    const int local = lid - list[n + e];
    // plane-fastest inside a volume: consecutive blocks reuse the kernel rows
    const int pl = local % S.nplanes, grp = local / S.nplanes;
    const long p = S.plane0 + pl;
    col_tile<PL>(g, Xin + p * xstride, Xout + p * xstride, hat ? S.Kh[view] : S.K[view], grp, lds,
                 hat && S.khconj ? -1.0f : 1.0f);
}

// E0 = 10 * nViews of the plane's volume (fmvd_deconvolve_cuda.impl.cu:57-60).
__global__ void batch_const_init_kernel(const BatchSeg* __restrict__ segs, const int* __restrict__ list, int n,
                                        float2* __restrict__ XE, float* __restrict__ ES, long xstride, long estride,
                                        long per_plane, int H)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;  // class plane
    if (i >= per_plane) return;
    const int e = find_entry(list, n, p);
    const float value = segs[list[e]].e0;
    XE[p * xstride + i] = (i < 2 * H && !(i & 1)) ? make_float2(value, 0.0f) : make_float2(0.0f, 0.0f);
    reinterpret_cast<float2*>(ES + p * estride)[i] = make_float2(value, value);
}

// fmvd_deconvolve_cuda.impl.cu:107-128 per volume: crop, min(E, max), truncate.
template <typename SAMPLE>
__global__ void batch_final_kernel(const BatchSeg* __restrict__ segs, const int* __restrict__ list, int n,
                                   const float* __restrict__ ES, long estride, int W)
{
    const int p = blockIdx.z;  // class plane
    const int e = find_entry(list, n, p);
    const BatchSeg& S = segs[list[e]];
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= S.dataW || y >= S.dataH) return;
    const float v = ES[p * estride + (long)y * W + x];
    const long o = (long)(p - S.plane0) * S.dstride + (long)y * S.dataW + x;
    if (S.est) S.est[o] = v;
    if (S.out) {
        constexpr float maxv = (float)((1u << (8 * sizeof(SAMPLE))) - 1u);
        const float tv = v > maxv ? maxv : v;
        static_cast<SAMPLE*>(S.out)[o] = tv > 0.0f ? (SAMPLE)tv : (SAMPLE)0;
    }
}

// ------------------------------------------------------------- launchers
template <int N, typename F> static void batch_with_variant(int v, F&& f)
{
    if (v == 1 && PlanV<N, 1>::distinct) {
        if constexpr (PlanV<N, 1>::distinct) f.template operator()<PlanV<N, 1>>();
    } else if (v == 2 && PlanV<N, 2>::distinct) {
        if constexpr (PlanV<N, 2>::distinct) f.template operator()<PlanV<N, 2>>();
    } else {
        f.template operator()<PlanV<N, 0>>();
    }
}

#define FMVD_BATCH_ROW_SIZES(X) X(128) X(256) X(512) X(768)
#define FMVD_BATCH_COL_SIZES(X) X(256) X(512) X(1024) X(1536)

template <typename F> static void batch_dispatch_row(int N, int v, F&& f)
{
    switch (N) {
#define FMVD_CASE(n)                  \
    case n:                           \
        batch_with_variant<n>(v, f);  \
        break;
        FMVD_BATCH_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("batched RL: unsupported row FFT length");
    }
}
template <typename F> static void batch_dispatch_col(int H, int v, F&& f)
{
    switch (H) {
#define FMVD_CASE(n)                  \
    case n:                           \
        batch_with_variant<n>(v, f);  \
        break;
        FMVD_BATCH_COL_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default:
        throw_msg("batched RL: unsupported column FFT length");
    }
}

bool batch_extent_supported(int fftH, int fftW)
{
    bool r = false, c = false;
    switch (fftW / 2) {
#define FMVD_CASE(n) case n: r = true; break;
        FMVD_BATCH_ROW_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default: break;
    }
    switch (fftH) {
#define FMVD_CASE(n) case n: c = true; break;
        FMVD_BATCH_COL_SIZES(FMVD_CASE)
#undef FMVD_CASE
    default: break;
    }
    return r && c;
}

int batch_row_rows_per_block(int N, int variant)
{
    int G = 0;
    batch_dispatch_row(N, variant, [&]<typename PL>() { G = RowCfg<PL>::G; });
    return G;
}
int batch_col_seqs_per_block(int H, int variant)
{
    int SQ = 0;
    batch_dispatch_col(H, variant, [&]<typename PL>() { SQ = ColCfg<PL>::SQ; });
    return SQ;
}

template <typename SAMPLE>
void launch_batch_ratio(const RowGeom& g, const BatchSeg* segs, const int* list, int n, int nblocks, int view,
                        float2* XT, long xstride, int variant, hipStream_t s)
{
    if (nblocks <= 0) return;
    batch_dispatch_row(g.N, variant, [&]<typename PL>() {
        batch_ratio_kernel<SAMPLE, PL><<<nblocks, RowCfg<PL>::THREADS, 0, s>>>(g, segs, list, n, view, XT, xstride);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_batch_update(const RowGeom& g, const BatchSeg* segs, const int* list, int n, int nblocks, int view,
                         float2* XE, const float2* XT, float* ES, long xstride, long estride, int variant,
                         hipStream_t s)
{
    if (nblocks <= 0) return;
    batch_dispatch_row(g.N, variant, [&]<typename PL>() {
        batch_update_kernel<PL><<<nblocks, RowCfg<PL>::THREADS, 0, s>>>(g, segs, list, n, view, XE, XT, ES, xstride,
                                                                        estride);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_batch_col(const ColGeom& g, const BatchSeg* segs, const int* list, int n, int nblocks, int view, int hat,
                      const float2* Xin, float2* Xout, long xstride, int variant, hipStream_t s)
{
    if (nblocks <= 0) return;
    batch_dispatch_col(g.H, variant, [&]<typename PL>() {
        batch_col_kernel<PL><<<nblocks, ColCfg<PL>::THREADS, 0, s>>>(g, segs, list, n, view, hat, Xin, Xout, xstride);
    });
    FMVD_LAUNCH_CHECK();
}

void launch_batch_const_init(const BatchSeg* segs, const int* list, int n, int nplanes, float2* XE, float* ES,
                             long xstride, long estride, int N, int H, hipStream_t s)
{
    if (nplanes <= 0) return;
    const long per_plane = (long)N * H;
    dim3 grid((unsigned)((per_plane + 255) / 256), (unsigned)nplanes);
    batch_const_init_kernel<<<grid, 256, 0, s>>>(segs, list, n, XE, ES, xstride, estride, per_plane, H);
    FMVD_LAUNCH_CHECK();
}

template <typename SAMPLE>
void launch_batch_final(const BatchSeg* segs, const int* list, int n, int nplanes, int maxH, int maxW,
                        const float* ES, long estride, int W, hipStream_t s)
{
    if (nplanes <= 0) return;
    dim3 grid((unsigned)((maxW + 255) / 256), (unsigned)maxH, (unsigned)nplanes);
    batch_final_kernel<SAMPLE><<<grid, 256, 0, s>>>(segs, list, n, ES, estride, W);
    FMVD_LAUNCH_CHECK();
}

template void launch_batch_ratio<uint16_t>(const RowGeom&, const BatchSeg*, const int*, int, int, int, float2*, long,
                                           int, hipStream_t);
template void launch_batch_ratio<uint8_t>(const RowGeom&, const BatchSeg*, const int*, int, int, int, float2*, long,
                                          int, hipStream_t);
template void launch_batch_final<uint16_t>(const BatchSeg*, const int*, int, int, int, int, const float*, long, int,
                                           hipStream_t);
template void launch_batch_final<uint8_t>(const BatchSeg*, const int*, int, int, int, int, const float*, long, int,
                                          hipStream_t);

}  // namespace fmvd
