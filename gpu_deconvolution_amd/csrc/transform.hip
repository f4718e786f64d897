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

// Two horizontally adjacent samples (x, x + 1) of one source row as one
// 4-byte (16-bit samples) or 2-byte (8-bit samples) load; x need not be even.
template <typename S>
__device__ __forceinline__ void load_pair(const S* row, int x, double& t0, double& t1)
{
    if constexpr (sizeof(S) == 2) {
        uint32_t u;
        __builtin_memcpy(&u, row + x, 4);
        t0 = (double)(u & 0xffffu);
        t1 = (double)(u >> 16);
    } else {
        uint16_t u;
        __builtin_memcpy(&u, row + x, 2);
        t0 = (double)(u & 0xffu);
        t1 = (double)(u >> 8);
    }
}

// Samples (ix, iy, iz) and (ix + 1, iy, iz), zero outside the volume.
template <typename S>
__device__ __forceinline__ void fetch_pair(const S* __restrict__ src, int w, int h, int d, int ix, int iy, int iz,
                                           double& t0, double& t1)
{
    if (iy < 0 || iy >= h || iz < 0 || iz >= d) {
        t0 = t1 = 0.0;
        return;
    }
    const S* row = src + ((size_t)iz * h + iy) * w;
    if (ix >= 0 && ix + 1 < w) {
        load_pair<S>(row, ix, t0, t1);
    } else {
        t0 = (ix >= 0 && ix < w) ? (double)row[ix] : 0.0;
        t1 = (ix + 1 >= 0 && ix + 1 < w) ? (double)row[ix + 1] : 0.0;
    }
}

// One output sample at plane z of target column (x, y); axy[a] holds
// m[4a] x + m[4a+1] y (the first sum of the fixed operation order).  FP64
// with no contraction:
//   r_a = ((m[4a] x + m[4a+1] y) + m[4a+2] z) + m[4a+3], i = floor(r), f = r - i,
//   nested lerps v (1 - f) + v' f in x, then y, then z (texels outside -> 0),
//   out = clamp(floor(interp / (2^B-1) * 2^B + 0.5), 0, 2^B-1)
// (fmvd_transform.impl.cu:41-53 with FP64 coordinates and weights).
template <typename S>
__device__ __forceinline__ S transform_sample(const S* __restrict__ src, int w, int h, int d, const Mat12& M,
                                              const double* axy, int z)
{
#pragma clang fp contract(off)
    const double zd = (double)z;
    const double rx = (axy[0] + M.m[2] * zd) + M.m[3];
    const double ry = (axy[1] + M.m[6] * zd) + M.m[7];
    const double rz = (axy[2] + M.m[10] * zd) + M.m[11];
    double val = 0.0;
    if (rx > -2.0 && ry > -2.0 && rz > -2.0 && rx < (double)w + 1.0 && ry < (double)h + 1.0 &&
        rz < (double)d + 1.0) {
        const double fx = floor(rx), fy = floor(ry), fz = floor(rz);
        const double ax = rx - fx, ay = ry - fy, az = rz - fz;
        const double bx = 1.0 - ax, by = 1.0 - ay, bz = 1.0 - az;
        const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
        double t000, t100, t010, t110, t001, t101, t011, t111;
        fetch_pair<S>(src, w, h, d, ix, iy, iz, t000, t100);
        fetch_pair<S>(src, w, h, d, ix, iy + 1, iz, t010, t110);
        fetch_pair<S>(src, w, h, d, ix, iy, iz + 1, t001, t101);
        fetch_pair<S>(src, w, h, d, ix, iy + 1, iz + 1, t011, t111);
        const double c00 = t000 * bx + t100 * ax;
        const double c10 = t010 * bx + t110 * ax;
        const double c01 = t001 * bx + t101 * ax;
        const double c11 = t011 * bx + t111 * ax;
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

// A thread computes kGZ consecutive target planes of one (x, y) column; a
// workgroup holds 64 x 4 columns.  The trilinear fetches read the source
// straight from global memory (two samples per load).  Under a rotation the
// 64 lanes of a wavefront walk diagonally through the source, so a source
// cache line is used by outputs that are neighbours in 3D, not along one
// target row: workgroups are therefore taken in supertiles of kSX x kSY x kSZ
// workgroups (128 x 32 x 32 target voxels, about 1.5 MB of source), numbered
// x fastest, then y, then z, with each XCD given a contiguous range, so the
// supertiles in flight on one XCD keep their source rows in its L2.
constexpr int kGX = 64, kGY = 4, kGZ = 8;
constexpr unsigned kSX = 2, kSY = 8, kSZ = 4, kSuper = kSX * kSY * kSZ;

__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblocks)
{
    const unsigned q = nblocks >> 3, rem = nblocks & 7u;
    const unsigned x = bid & 7u, k = bid >> 3;
    return x * q + (x < rem ? x : rem) + k;
}

template <typename S>
__global__ __launch_bounds__(256) void transform_kernel(const S* __restrict__ src, int w, int h, int d,
                                                        S* __restrict__ dst, int tw, int th, int z0, int nz,
                                                        Mat12 M, unsigned nbx, unsigned nby, unsigned nsx,
                                                        unsigned nsy)
{
#pragma clang fp contract(off)
    const unsigned lid = xcd_block(blockIdx.x, gridDim.x);
    const unsigned st = lid / kSuper, wi = lid % kSuper;
    const unsigned bx = (st % nsx) * kSX + wi % kSX;
    const unsigned by = ((st / nsx) % nsy) * kSY + (wi / kSX) % kSY;
    const unsigned bz = (st / (nsx * nsy)) * kSZ + wi / (kSX * kSY);
    if (bx >= nbx || by >= nby) return;
    const int x = (int)bx * kGX + (int)(threadIdx.x % kGX);
    const int y = (int)by * kGY + (int)(threadIdx.x / kGX);
    if (x >= tw || y >= th) return;
    const double xd = (double)x, yd = (double)y;
    const double axy[3] = {M.m[0] * xd + M.m[1] * yd, M.m[4] * xd + M.m[5] * yd, M.m[8] * xd + M.m[9] * yd};
    const int zb = z0 + (int)bz * kGZ, ze = min(zb + kGZ, z0 + nz);
    if (zb >= ze) return;
    S* out = dst + ((size_t)(zb - z0) * th + y) * tw + x;
    const size_t pstride = (size_t)th * tw;
    for (int z = zb; z < ze; ++z, out += pstride) *out = transform_sample<S>(src, w, h, d, M, axy, z);
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
    const unsigned nbx = (unsigned)((tw + kGX - 1) / kGX), nby = (unsigned)((th + kGY - 1) / kGY);
    const unsigned nbz = (unsigned)((nz + kGZ - 1) / kGZ);
    const unsigned nsx = (nbx + kSX - 1) / kSX, nsy = (nby + kSY - 1) / kSY, nsz = (nbz + kSZ - 1) / kSZ;
    const unsigned long long nblocks = (unsigned long long)nsx * nsy * nsz * kSuper;
    if (nblocks > 0x7fffffffull) throw_msg("transform: target volume too large for one launch");
    TimedScope ts(kTimeTransform, s);
    launch(transform_kernel<S>, dim3((unsigned)nblocks), dim3(kGX * kGY), 0, s, src, w, h, d, dst, tw, th, z0, nz,
           M, nbx, nby, nsx, nsy);
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
