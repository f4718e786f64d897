// Affine view transform and border-weight mask (gfx950).
//
// Reference: fmvd_transform.impl.cu:24-56 (transform_plane_kernel, a CUDA
// 3D-texture trilinear fetch at r + 0.5 with border address mode and
// normalised-float read), :77-224 (transform_cuda driver) and
// fmvd_transform.cu:19-89 (create_transformed_mask).
//
// Arithmetic: source coordinates and interpolation weights are evaluated in
// FP64 with a fixed operation order and no contraction, so every output
// sample equals the FP64 definition bit for bit:
//   r = ((m0 x + m1 y) + m2 z) + m3 per axis, i = floor(r), a = r - i,
//   nested lerps v*(1-a) + v'*a in x, then y, then z (texels outside -> 0),
//   out = clamp(floor(interp / (2^B-1) * 2^B + 0.5), 0, 2^B-1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/fmvd_transform.h"
#include "hip_check.hpp"
#include "timing.hpp"
#include "transform.hpp"

namespace fmvd {

struct Mat12 {
    double m[12];
};

// One output sample at (x, y, z); fetch(ix, iy, iz) returns the texel or 0
// outside the volume.  Fixed FP64 operation order, no contraction.
template <typename S, typename Fetch>
__device__ __forceinline__ S transform_sample(int w, int h, int d, const Mat12& M, int x, int y, int z, Fetch fetch)
{
#pragma clang fp contract(off)
    const double xd = (double)x, yd = (double)y, zd = (double)z;
    const double rx = ((M.m[0] * xd + M.m[1] * yd) + M.m[2] * zd) + M.m[3];
    const double ry = ((M.m[4] * xd + M.m[5] * yd) + M.m[6] * zd) + M.m[7];
    const double rz = ((M.m[8] * xd + M.m[9] * yd) + M.m[10] * zd) + M.m[11];
    double val = 0.0;
    if (rx > -2.0 && ry > -2.0 && rz > -2.0 && rx < (double)w + 1.0 && ry < (double)h + 1.0 &&
        rz < (double)d + 1.0) {
        const double fx = floor(rx), fy = floor(ry), fz = floor(rz);
        const double ax = rx - fx, ay = ry - fy, az = rz - fz;
        const double bx = 1.0 - ax, by = 1.0 - ay, bz = 1.0 - az;
        const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
        const double c00 = fetch(ix, iy, iz) * bx + fetch(ix + 1, iy, iz) * ax;
        const double c10 = fetch(ix, iy + 1, iz) * bx + fetch(ix + 1, iy + 1, iz) * ax;
        const double c01 = fetch(ix, iy, iz + 1) * bx + fetch(ix + 1, iy, iz + 1) * ax;
        const double c11 = fetch(ix, iy + 1, iz + 1) * bx + fetch(ix + 1, iy + 1, iz + 1) * ax;
        const double c0 = c00 * by + c10 * ay;
        const double c1 = c01 * by + c11 * ay;
        val = c0 * bz + c1 * az;
    }
    constexpr double maxv = (double)((1u << (8 * sizeof(S))) - 1u);
    constexpr double scale = (double)(1u << (8 * sizeof(S)));
    // yv = (val / maxv) * scale + 0.5 decides the sample only through its
    // integer part and the bounds 0 and maxv (integers).  val * (1 / maxv)
    // stays within 3 ulp of val / maxv, so the reciprocal form moves yv
    // (< 2^16) by less than 1e-10: when it lies farther than 1e-9 from an
    // integer both forms give the same sample, otherwise the division
    // decides.
    constexpr double inv_maxv = 1.0 / maxv;
    double yv = (val * inv_maxv) * scale + 0.5;
    const double fr = yv - floor(yv);
    if (!(fr > 1e-9 && fr < 1.0 - 1e-9)) yv = (val / maxv) * scale + 0.5;
    if (yv >= maxv) return (S)maxv;
    if (yv > 0.0) return (S)floor(yv);
    return (S)0;
}

// Output tile of a workgroup: 16 x 16 x 8 voxels, thread (tx, ty) computes
// the 8 voxels of its column.  The tile's source footprint is the bounding
// box of its 8 corner images (an affine map takes its extremes at corners),
// widened by one texel for rounding and clamped to [-3, extent + 2], which
// holds every texel an in-range sample can read.  When the box fits the
// LDS brick it is staged with row-contiguous (coalesced) loads, zero outside
// the volume, and all trilinear fetches come from LDS; otherwise the
// workgroup gathers from global memory.  The brick (dynamic LDS, `cap`
// texels) is sized on the host from the map: every tile's box fits it, and a
// smaller brick lets more workgroups share a CU.
constexpr int kTX = 16, kTY = 16, kTZ = 8;
constexpr int kBrickMaxBytes = 48 * 1024;

template <typename S>
__global__ __launch_bounds__(256) void transform_kernel(const S* __restrict__ src, int w, int h, int d,
                                                        S* __restrict__ dst, int tw, int th, int z0, int nz,
                                                        Mat12 M, int cap)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char brick_lds[];
    S* brick = reinterpret_cast<S*>(brick_lds);
    const int CAP = cap;
    const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
    const int x0 = blockIdx.x * kTX, y0 = blockIdx.y * kTY, zt0 = z0 + blockIdx.z * kTZ;
    const int x1 = min(x0 + kTX, tw) - 1, y1 = min(y0 + kTY, th) - 1, zt1 = min(zt0 + kTZ, z0 + nz) - 1;
    // bounding box of the tile's source footprint (block-uniform)
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int c = 0; c < 8; ++c) {
        const double xd = (double)((c & 1) ? x1 : x0), yd = (double)((c & 2) ? y1 : y0),
                     zd = (double)((c & 4) ? zt1 : zt0);
        for (int a = 0; a < 3; ++a) {
            const double r = ((M.m[4 * a] * xd + M.m[4 * a + 1] * yd) + M.m[4 * a + 2] * zd) + M.m[4 * a + 3];
            lo[a] = fmin(lo[a], r);
            hi[a] = fmax(hi[a], r);
        }
    }
    const int ext[3] = {w, h, d};
    int b0[3], bn[3];
    bool outside = false;
    long cells = 1;
    for (int a = 0; a < 3; ++a) {
        // (a margin keeps interior points that round past a corner value in the staged path; NaN -> outside)
        if (!(hi[a] > -2.0 - 1e-6) || !(lo[a] < (double)ext[a] + 1.0 + 1e-6)) outside = true;
        const double l = fmax(floor(lo[a]) - 1.0, -3.0), u = fmin(floor(hi[a]) + 2.0, (double)ext[a] + 2.0);
        b0[a] = outside ? 0 : (int)l;
        bn[a] = outside ? 1 : max(1, (int)u - (int)l + 1);
        cells *= bn[a];
    }
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x <= x1 && y <= y1;
    if (outside) {  // every sample of the tile is outside the source: zeros
        if (active)
            for (int z = zt0; z <= zt1; ++z) dst[((long)(z - z0) * th + y) * tw + x] = (S)0;
        return;
    }
    if (cells <= CAP) {
        const int BX = bn[0], BY = bn[1];
        // 16 texels per thread in flight before their LDS writes.  The brick
        // index e splits into (x, y, z) with float reciprocals and one
        // correction step (e < 2^16: the quotient is off by at most one).
        constexpr int kU = 16;
        const float inv_bx = 1.0f / (float)BX, inv_by = 1.0f / (float)BY;
        for (int e0 = threadIdx.x; e0 < (int)cells; e0 += 256 * kU) {
            S v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int e = e0 + 256 * u;
                int q = (int)((float)e * inv_bx), ex = e - q * BX;
                if (ex < 0) {
                    --q;
                    ex += BX;
                } else if (ex >= BX) {
                    ++q;
                    ex -= BX;
                }
                int ez = (int)((float)q * inv_by), ey = q - ez * BY;
                if (ey < 0) {
                    --ez;
                    ey += BY;
                } else if (ey >= BY) {
                    ++ez;
                    ey -= BY;
                }
                const int gx = b0[0] + ex, gy = b0[1] + ey, gz = b0[2] + ez;
                const bool in = e < (int)cells && gx >= 0 && gy >= 0 && gz >= 0 && gx < w && gy < h && gz < d;
                v[u] = in ? src[((long)gz * h + gy) * w + gx] : (S)0;
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (e0 + 256 * u < (int)cells) brick[e0 + 256 * u] = v[u];
        }
        __syncthreads();
        if (!active) return;
        auto fetch = [&](int ix, int iy, int iz) -> double {
            return (double)brick[((iz - b0[2]) * BY + (iy - b0[1])) * BX + (ix - b0[0])];
        };
        for (int z = zt0; z <= zt1; ++z)
            dst[((long)(z - z0) * th + y) * tw + x] = transform_sample<S>(w, h, d, M, x, y, z, fetch);
    } else {
        if (!active) return;
        auto fetch = [&](int ix, int iy, int iz) -> double {
            if (ix < 0 || iy < 0 || iz < 0 || ix >= w || iy >= h || iz >= d) return 0.0;
            return (double)src[((long)iz * h + iy) * w + ix];
        };
        for (int z = zt0; z <= zt1; ++z)
            dst[((long)(z - z0) * th + y) * tw + x] = transform_sample<S>(w, h, d, M, x, y, z, fetch);
    }
}

__global__ __launch_bounds__(256) void mask_kernel(float* __restrict__ out, int z, int w, int h, int d,
                                                   double zspacing, int tw, int th, double border,
                                                   Mat12 M)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= tw || y >= th) return;
    const double xd = (double)x, yd = (double)y, zd = (double)z;
    const double rx = ((M.m[0] * xd + M.m[1] * yd) + M.m[2] * zd) + M.m[3];
    const double ry = ((M.m[4] * xd + M.m[5] * yd) + M.m[6] * zd) + M.m[7];
    const double rz = ((M.m[8] * xd + M.m[9] * yd) + M.m[10] * zd) + M.m[11];
    float r = 0.0f;
    if (!(rx < 0 || rx >= w || ry < 0 || ry >= h || rz < 0 || rz >= d)) {
        // fmvd_transform.cu:46-64 (w/2 etc. are integer divisions)
        double dx = rx < (double)(w / 2) ? rx : (double)w - rx;
        double dy = ry < (double)(h / 2) ? ry : (double)h - ry;
        double dz = rz < (double)(d / 2) ? rz : (double)d - rz;
        dz *= zspacing;
        double v = 1.0;
        if (dx < border) v *= 0.5 * (1.0 - cos(dx / border * M_PI));
        if (dy < border) v *= 0.5 * (1.0 - cos(dy / border * M_PI));
        if (dz < border) v *= 0.5 * (1.0 - cos(dz / border * M_PI));
        r = (float)v;
    }
    out[(long)y * tw + x] = r;
}

static Mat12 to_mat(const float* m)
{
    Mat12 M;
    for (int i = 0; i < 12; ++i) M.m[i] = (double)m[i];
    return M;
}

template <typename S>
static void launch_transform(const S* src, int w, int h, int d, S* dst, int tw, int th, int z0, int nz,
                             const Mat12& M, hipStream_t s)
{
    if (nz <= 0 || tw <= 0 || th <= 0) return;
    dim3 grid((unsigned)((tw + kTX - 1) / kTX), (unsigned)((th + kTY - 1) / kTY), (unsigned)((nz + kTZ - 1) / kTZ));
    // Box of a tile's source footprint per axis: the corner images span
    // sum_j |m_aj| (tile extent_j - 1); flooring and the one-texel widening
    // on both sides add at most 4 cells (+1 for rounding of the span).
    const int ext[3] = {kTX, kTY, kTZ};
    double cells = 1.0;
    for (int a = 0; a < 3; ++a) {
        double span = 0.0;
        for (int j = 0; j < 3; ++j) span += std::fabs(M.m[4 * a + j]) * (double)(ext[j] - 1);
        cells *= std::ceil(span) + 5.0;
    }
    const double maxb = (double)kBrickMaxBytes;
    const int bytes = cells * sizeof(S) >= maxb ? kBrickMaxBytes
                                                : (int)((cells * sizeof(S) + 255.0) / 256.0) * 256;
    TimedScope ts(kTimeTransform, s);
    launch(transform_kernel<S>, grid, dim3(256), (unsigned)bytes, s, src, w, h, d, dst, tw, th, z0, nz, M,
           bytes / (int)sizeof(S));
    FMVD_LAUNCH_CHECK();
}

static void launch_mask(float* out, int w, int h, int d, int tw, int th, int td, const Mat12& M,
                        int border, float zspacing, hipStream_t s)
{
    dim3 block(64, 4);
    dim3 grid((unsigned)((tw + 63) / 64), (unsigned)((th + 3) / 4));
    mask_kernel<<<grid, block, 0, s>>>(out, td / 2, w, h, d, (double)zspacing, tw, th, (double)border, M);
    FMVD_LAUNCH_CHECK();
}

static void check_dims(int w, int h, int d, int tw, int th, int td)
{
    if (w < 1 || h < 1 || d < 1 || tw < 1 || th < 1 || td < 1)
        throw_msg("invalid transform dimensions (all extents must be >= 1)");
}

// fmvd_transform.impl.cu:77-224: whole source volume resident on the device,
// target planes computed in chunks and written to `outfile` in z order while
// the next chunk computes.
template <typename S>
static void transform_files(S** h_data, int w, int h, int d, int tw, int th, int td, const float* inv,
                            const char* outfile, int createMasks, int border, float zspacing,
                            const char* maskfile)
{
    check_dims(w, h, d, tw, th, td);
    if (!h_data || !inv || !outfile) throw_msg("transform: NULL argument");
    const Mat12 M = to_mat(inv);
    const size_t plane = (size_t)w * h, tplane = (size_t)tw * th;
    S* dsrc = nullptr;
    S* ddst = nullptr;
    S* hdst = nullptr;
    float* dmask = nullptr;
    float* hmask = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    FILE* out = nullptr;
    auto cleanup = [&] {
        if (s) (void)hipStreamSynchronize(s);
        if (dsrc) (void)hipFree(dsrc);
        if (ddst) (void)hipFree(ddst);
        if (hdst) (void)hipHostFree(hdst);
        if (dmask) (void)hipFree(dmask);
        if (hmask) (void)hipHostFree(hmask);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
        if (out) fclose(out);
    };
    try {
        FMVD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        FMVD_HIP(hipMalloc(&dsrc, sizeof(S) * plane * d));
        for (int z = 0; z < d; ++z)
            FMVD_HIP(hipMemcpy(dsrc + plane * z, h_data[z], sizeof(S) * plane, hipMemcpyHostToDevice));
        if (createMasks) {
            if (!maskfile) throw_msg("transform: mask requested without a mask file");
            FMVD_HIP(hipMalloc(&dmask, sizeof(float) * tplane));
            FMVD_HIP(hipHostMalloc(&hmask, sizeof(float) * tplane, hipHostMallocDefault));
            launch_mask(dmask, w, h, d, tw, th, td, M, border, zspacing, s);
            FMVD_HIP(hipMemcpyAsync(hmask, dmask, sizeof(float) * tplane, hipMemcpyDeviceToHost, s));
            FMVD_HIP(hipStreamSynchronize(s));
            FILE* mf = fopen(maskfile, "wb");
            if (!mf) throw_msg(std::string("cannot open mask file for writing: ") + maskfile);
            const size_t wr = fwrite(hmask, sizeof(float), tplane, mf);
            fclose(mf);
            if (wr != tplane) throw_msg("short write to mask file");
        }
        out = fopen(outfile, "wb");
        if (!out) throw_msg(std::string("cannot open output file for writing: ") + outfile);
        // chunk of target planes per launch (~64 MiB of output)
        int zc = (int)std::max<size_t>(1, ((size_t)64 << 20) / (sizeof(S) * tplane));
        zc = std::min(zc, td);
        FMVD_HIP(hipMalloc(&ddst, sizeof(S) * tplane * zc * 2));
        FMVD_HIP(hipHostMalloc(&hdst, sizeof(S) * tplane * zc * 2, hipHostMallocDefault));
        for (auto& e : ev) FMVD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        int pend_z = -1, pend_n = 0, pend_slot = 0;
        auto flush = [&](int slot, int n) {
            FMVD_HIP(hipEventSynchronize(ev[slot]));
            const size_t cnt = tplane * n;
            if (fwrite(hdst + tplane * zc * slot, sizeof(S), cnt, out) != cnt)
                throw_msg("short write to output file");
        };
        int slot = 0;
        for (int z0 = 0; z0 < td; z0 += zc) {
            const int n = std::min(zc, td - z0);
            S* dd = ddst + tplane * zc * slot;
            launch_transform<S>(dsrc, w, h, d, dd, tw, th, z0, n, M, s);
            FMVD_HIP(hipMemcpyAsync(hdst + tplane * zc * slot, dd, sizeof(S) * tplane * n,
                                    hipMemcpyDeviceToHost, s));
            FMVD_HIP(hipEventRecord(ev[slot], s));
            if (pend_z >= 0) flush(pend_slot, pend_n);
            pend_z = z0;
            pend_n = n;
            pend_slot = slot;
            slot ^= 1;
        }
        if (pend_z >= 0) flush(pend_slot, pend_n);
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

void transform_volume_u16(const uint16_t* src, int w, int h, int d, uint16_t* dst, int tw, int th, int td,
                          const float* inv, hipStream_t s)
{
    check_dims(w, h, d, tw, th, td);
    if (!src || !dst || !inv) throw_msg("transform: NULL argument");
    launch_transform<uint16_t>(src, w, h, d, dst, tw, th, 0, td, to_mat(inv), s);
}

void transform_volume_u8(const uint8_t* src, int w, int h, int d, uint8_t* dst, int tw, int th, int td,
                         const float* inv, hipStream_t s)
{
    check_dims(w, h, d, tw, th, td);
    if (!src || !dst || !inv) throw_msg("transform: NULL argument");
    launch_transform<uint8_t>(src, w, h, d, dst, tw, th, 0, td, to_mat(inv), s);
}

void transform_mask(int w, int h, int d, int tw, int th, int td, const float* inv, int border, float zspacing,
                    float* out, hipStream_t s)
{
    check_dims(w, h, d, tw, th, td);
    if (!out || !inv) throw_msg("mask: NULL argument");
    launch_mask(out, w, h, d, tw, th, td, to_mat(inv), border, zspacing, s);
}

}  // namespace fmvd

extern "C" {

void transform_cuda_8(unsigned char** h_data, int w, int h, int d, int tw, int th, int td,
                      float* h_inverse, const char* outfile, int createTransformedMasks, int border,
                      float zspacing, const char* maskfile)
{
    FMVD_API_GUARD(, {
        fmvd::transform_files<uint8_t>(h_data, w, h, d, tw, th, td, h_inverse, outfile,
                                       createTransformedMasks, border, zspacing, maskfile);
    })
}

void transform_cuda_16(unsigned short** h_data, int w, int h, int d, int tw, int th, int td,
                       float* h_inverse, const char* outfile, int createTransformedMasks, int border,
                       float zspacing, const char* maskfile)
{
    FMVD_API_GUARD(, {
        fmvd::transform_files<uint16_t>(h_data, w, h, d, tw, th, td, h_inverse, outfile,
                                        createTransformedMasks, border, zspacing, maskfile);
    })
}

int fmvd_transform_device_8(const unsigned char* d_src, int w, int h, int d, unsigned char* d_dst,
                            int tw, int th, int td, const float* h_inverse, void* stream)
{
    FMVD_API_GUARD(-1, {
        fmvd::transform_volume_u8(d_src, w, h, d, d_dst, tw, th, td, h_inverse, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_transform_device_16(const unsigned short* d_src, int w, int h, int d, unsigned short* d_dst,
                             int tw, int th, int td, const float* h_inverse, void* stream)
{
    FMVD_API_GUARD(-1, {
        fmvd::transform_volume_u16(d_src, w, h, d, d_dst, tw, th, td, h_inverse, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_transform_mask_device(int w, int h, int d, int tw, int th, int td, const float* h_inverse,
                               int border, float zspacing, float* d_mask, void* stream)
{
    FMVD_API_GUARD(-1, {
        fmvd::transform_mask(w, h, d, tw, th, td, h_inverse, border, zspacing, d_mask, (hipStream_t)stream);
        return 0;
    })
}

}  // extern "C"
