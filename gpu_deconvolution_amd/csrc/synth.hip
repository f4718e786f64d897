// Synthetic SPIM volumes on the GPU for the benchmark workloads
// (libfmvd_synth.so; not part of the reconstruction library).
//
// Ground truth: background plus Gaussian beads (atomic accumulation, one
// workgroup per bead, rendered within +-ceil(4 sigma)).  Blur: separable
// Gaussian along one axis with clamp-to-edge borders.  Noise: Poisson
// samples from a counter-based generator (SplitMix64 of seed and voxel
// index), by multiplication of uniforms for lambda < 10 and by Hoermann's
// transformed rejection with squeeze (PTRS, W. Hoermann, "The transformed
// rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12, 1993) above; clamped to the sample range.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "hip_check.hpp"

namespace {

struct Bead {
    float cz, cy, cx, sigma, amp;
};

// Element-wise kernels walk the volume grid-stride: a launch may hold at
// most 2^32 work-items per grid dimension, below a 2048 x 2048 x 1024 volume.
__global__ void fill_kernel(float* __restrict__ v, long n, float value)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        v[i] = value;
}

__global__ void beads_kernel(float* __restrict__ vol, int Z, int H, int W, const Bead* __restrict__ beads)
{
    const Bead b = beads[blockIdx.x];
    const int r = (int)ceilf(4.0f * b.sigma);
    const int z0 = max(0, (int)b.cz - r), z1 = min(Z, (int)b.cz + r + 1);
    const int y0 = max(0, (int)b.cy - r), y1 = min(H, (int)b.cy + r + 1);
    const int x0 = max(0, (int)b.cx - r), x1 = min(W, (int)b.cx + r + 1);
    if (z0 >= z1 || y0 >= y1 || x0 >= x1) return;
    const int nx = x1 - x0, ny = y1 - y0, nz = z1 - z0;
    const long total = (long)nx * ny * nz;
    const float inv2s2 = 1.0f / (2.0f * b.sigma * b.sigma);
    for (long i = threadIdx.x; i < total; i += blockDim.x) {
        const int x = x0 + (int)(i % nx);
        const int y = y0 + (int)((i / nx) % ny);
        const int z = z0 + (int)(i / ((long)nx * ny));
        const float dz = z - b.cz, dy = y - b.cy, dx = x - b.cx;
        atomicAdd(&vol[((long)z * H + y) * W + x], b.amp * expf(-(dx * dx + dy * dy + dz * dz) * inv2s2));
    }
}

struct Taps {
    int r;
    float g[65];
};

__global__ void blur_kernel(const float* __restrict__ in, float* __restrict__ out, int Z, int H, int W, int axis,
                            Taps t)
{
    const long n = (long)Z * H * W;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % W);
    const int y = (int)((i / W) % H);
    const int z = (int)(i / ((long)W * H));
    const int len = axis == 0 ? W : (axis == 1 ? H : Z);
    const int c = axis == 0 ? x : (axis == 1 ? y : z);
    const long stride = axis == 0 ? 1 : (axis == 1 ? (long)W : (long)W * H);
    const long base = i - (long)c * stride;
    float acc = 0.0f;
    for (int k = -t.r; k <= t.r; ++k) {
        int cc = c + k;
        cc = cc < 0 ? 0 : (cc >= len ? len - 1 : cc);
        acc += t.g[k + t.r] * in[base + (long)cc * stride];
    }
    out[i] = acc;
    }
}

__device__ __forceinline__ uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

__device__ double poisson(double lam, uint64_t& s)
{
    if (lam <= 0.0) return 0.0;
    if (lam < 10.0) {
        const double L = exp(-lam);
        double p = 1.0;
        int k = 0;
        do {
            ++k;
            p *= uniform(s);
        } while (p > L && k < 1000);
        return (double)(k - 1);
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 256; ++it) {
        const double U = uniform(s) - 0.5;
        const double V = uniform(s);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
    return floor(lam + 0.5);
}

template <typename S>
__global__ void poisson_kernel(const float* __restrict__ lam, S* __restrict__ out, long n, uint64_t seed, double maxv)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        uint64_t s = seed ^ ((uint64_t)i * 0xD1342543DE82EF95ull);
        (void)splitmix(s);
        const double k = poisson((double)lam[i], s);
        out[i] = (S)(k > maxv ? maxv : k);
    }
}

// Plain streaming copy (16 bytes per thread): the known-size reference kernel
// for HBM byte counters (reads n bytes, writes n bytes).
__global__ void copy_kernel(const uint4* __restrict__ in, uint4* __restrict__ out, long n16)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) out[i] = in[i];
}

inline unsigned blocks(long n, int t) { return (unsigned)((n + t - 1) / t); }
// grid of the grid-stride kernels
inline unsigned strided_blocks(long n, int t)
{
    const long b = (n + t - 1) / t;
    return (unsigned)(b < (1L << 20) ? b : (1L << 20));
}

}  // namespace

extern "C" {

int fmvd_synth_truth(float* vol, int Z, int H, int W, const float* beads5, int nbeads, float background,
                     void* stream)
{
    FMVD_API_GUARD(-1, {
        hipStream_t s = (hipStream_t)stream;
        const long n = (long)Z * H * W;
        fill_kernel<<<strided_blocks(n, 256), 256, 0, s>>>(vol, n, background);
        FMVD_LAUNCH_CHECK();
        if (nbeads > 0) {
            Bead* d = nullptr;
            FMVD_HIP(hipMalloc(&d, sizeof(Bead) * nbeads));
            FMVD_HIP(hipMemcpy(d, beads5, sizeof(Bead) * nbeads, hipMemcpyHostToDevice));
            beads_kernel<<<(unsigned)nbeads, 256, 0, s>>>(vol, Z, H, W, d);
            FMVD_LAUNCH_CHECK();
            FMVD_HIP(hipStreamSynchronize(s));
            FMVD_HIP(hipFree(d));
        }
        return 0;
    })
}

// axis 0 = x, 1 = y, 2 = z; sigma <= 8
int fmvd_synth_blur(const float* in, float* out, int Z, int H, int W, int axis, float sigma, void* stream)
{
    FMVD_API_GUARD(-1, {
        Taps t;
        t.r = (int)std::ceil(4.0 * sigma);
        if (t.r > 32) fmvd::throw_msg("blur sigma too large");
        double sum = 0;
        std::vector<double> g(2 * t.r + 1);
        for (int k = -t.r; k <= t.r; ++k) sum += g[k + t.r] = std::exp(-0.5 * (k / (double)sigma) * (k / (double)sigma));
        for (int k = 0; k <= 2 * t.r; ++k) t.g[k] = (float)(g[k] / sum);
        const long n = (long)Z * H * W;
        blur_kernel<<<strided_blocks(n, 256), 256, 0, (hipStream_t)stream>>>(in, out, Z, H, W, axis, t);
        FMVD_LAUNCH_CHECK();
        return 0;
    })
}

int fmvd_synth_poisson(const float* lam, void* out, long n, unsigned long long seed, int bits, void* stream)
{
    FMVD_API_GUARD(-1, {
        hipStream_t s = (hipStream_t)stream;
        if (bits == 16)
            poisson_kernel<uint16_t><<<strided_blocks(n, 256), 256, 0, s>>>(lam, (uint16_t*)out, n, seed, 65535.0);
        else
            poisson_kernel<uint8_t><<<strided_blocks(n, 256), 256, 0, s>>>(lam, (uint8_t*)out, n, seed, 255.0);
        FMVD_LAUNCH_CHECK();
        return 0;
    })
}

// bytes must be a multiple of 16; both pointers 16-byte aligned.
int fmvd_synth_copy(const void* in, void* out, long bytes, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (bytes < 0 || bytes % 16) fmvd::throw_msg("fmvd_synth_copy: bytes must be a non-negative multiple of 16");
        const long n16 = bytes / 16;
        if (n16) copy_kernel<<<blocks(n16, 256), 256, 0, (hipStream_t)stream>>>((const uint4*)in, (uint4*)out, n16);
        FMVD_LAUNCH_CHECK();
        return 0;
    })
}

}  // extern "C"
