// Warm-start network kernels (gfx950): a 6-layer 3-D residual conv net with
// 32 channels and 3x3x3 kernels, bfloat16 weights and activations, FP32
// accumulation.  No counterpart in the reference.
//
// Layout: activations are voxel-major, channel-fastest: act[v][32] bf16 with
// v = ((n * D + z) * H + y) * W + x (n = sample of a batch).  Weights of a
// 32->32 layer are [tap][cout][cin] bf16 (tap = (dz*3 + dy)*3 + dx reads
// voxel (z+dz-1, y+dy-1, x+dx-1); zero padding at the volume faces).
//
// The four 32->32 layers (98% of the work) run on MFMA: one wavefront
// computes a 32-voxel x 32-channel output tile as an implicit GEMM over
// K = 27 taps x 32 input channels with v_mfma_f32_32x32x16_bf16 (A = 32
// consecutive voxels x 16 channels loaded straight from the activation
// layout, B = 16 channels x 32 outputs from the block's LDS copy of the
// layer weights).  The weight gradient of those layers is the same GEMM with
// voxels as the reduction dimension, staged through LDS in channel-major
// order.  The 2->32 input layer and the 32->1 output layer have too few
// channels for MFMA tiles and run on the VALU, still with FP32 accumulation.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "hip_check.hpp"
#include "net.hpp"

namespace fmvd {
namespace net {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct Geo {
    int N, D, H, W;  // batch, depth, height, width
    __device__ __forceinline__ long vox(int n, int z, int y, int x) const
    {
        return (((long)n * D + z) * H + y) * W + x;
    }
};

// ------------------------------------------------------------ MFMA check
// D = A (32 x 16) * B (16 x 32) for a unit test of the operand layout:
// lane l supplies A[l%32][8(l/32)+j] and B[8(l/32)+j][l%32] (j < 8) and
// receives D[8(r/4) + 4(l/32) + r%4][l%32] in element r (r < 16).
__global__ void mfma_layout_kernel(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                   float* __restrict__ Dm)
{
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int j = 0; j < 8; ++j) {
        const uint16_t av = A[(l % 32) * 16 + 8 * (l / 32) + j];
        const uint16_t bv = B[(8 * (l / 32) + j) * 32 + (l % 32)];
        a[j] = __builtin_bit_cast(__bf16, av);
        b[j] = __builtin_bit_cast(__bf16, bv);
    }
    f32x16 acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 16; ++r) Dm[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + (l % 32)] = acc[r];
}

// ------------------------------------------------------------ input layer
// views (u16, V volumes of D x H x W per sample) -> x (bf16, 2 channels):
// mean of the even-index views, mean of the odd-index views, over scale.
__global__ void prep_input_kernel(const uint16_t* const* __restrict__ views, int nviews, Geo g,
                                  const float* __restrict__ inv_scale, uint16_t* __restrict__ x)
{
    const long per = (long)g.D * g.H * g.W;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= per) return;
    float se = 0.f, so = 0.f;
    int ne = 0, no = 0;
    for (int v = 0; v < nviews; ++v) {
        const float s = (float)views[n * nviews + v][i];
        if (v & 1) {
            so += s;
            ++no;
        } else {
            se += s;
            ++ne;
        }
    }
    const float e = se / (float)ne, o = no ? so / (float)no : e;
    const float k = inv_scale[n];
    x[2 * (n * per + i)] = f2bf(e * k);
    x[2 * (n * per + i) + 1] = f2bf(o * k);
}

// Layer 1 (2 -> 32), VALU: h1 = relu(conv(x, W1) + b1); W1 is FP32 [27][32][2]
// (bf16-rounded values).  One thread per voxel.
__global__ void conv_in_kernel(const uint16_t* __restrict__ x, Geo g, const float* __restrict__ w,
                               const float* __restrict__ b, uint16_t* __restrict__ h)
{
    __shared__ float ws[27 * 32 * 2 + 32];
    for (int i = threadIdx.x; i < 27 * 64 + 32; i += blockDim.x) ws[i] = i < 27 * 64 ? w[i] : b[i - 27 * 64];
    __syncthreads();
    const long per = (long)g.D * g.H * g.W;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= per) return;
    const int xx = (int)(i % g.W), yy = (int)((i / g.W) % g.H), zz = (int)(i / ((long)g.W * g.H));
    float acc[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) acc[c] = ws[27 * 64 + c];
    for (int k = 0; k < 27; ++k) {
        const int z = zz + k / 9 - 1, y = yy + (k / 3) % 3 - 1, xi = xx + k % 3 - 1;
        if (z < 0 || z >= g.D || y < 0 || y >= g.H || xi < 0 || xi >= g.W) continue;
        const long v = g.vox(n, z, y, xi);
        const float x0 = bf2f(x[2 * v]), x1 = bf2f(x[2 * v + 1]);
#pragma unroll
        for (int c = 0; c < 32; ++c) acc[c] += x0 * ws[(k * 32 + c) * 2] + x1 * ws[(k * 32 + c) * 2 + 1];
    }
    uint16_t* o = h + (n * per + i) * 32;
#pragma unroll
    for (int c = 0; c < 32; ++c) o[c] = f2bf(fmaxf(acc[c], 0.f));
}

// Layer 6 (32 -> 1), VALU: y = conv(h5, W6) + b6 + (x0 + x1) / 2 (FP32 out).
__global__ void conv_out_kernel(const uint16_t* __restrict__ h, const uint16_t* __restrict__ x, Geo g,
                                const float* __restrict__ w, const float* __restrict__ b, float* __restrict__ y)
{
    __shared__ float ws[27 * 32];
    for (int i = threadIdx.x; i < 27 * 32; i += blockDim.x) ws[i] = w[i];
    __syncthreads();
    const long per = (long)g.D * g.H * g.W;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= per) return;
    const int xx = (int)(i % g.W), yy = (int)((i / g.W) % g.H), zz = (int)(i / ((long)g.W * g.H));
    float acc = b[0];
    for (int k = 0; k < 27; ++k) {
        const int z = zz + k / 9 - 1, yv = yy + (k / 3) % 3 - 1, xi = xx + k % 3 - 1;
        if (z < 0 || z >= g.D || yv < 0 || yv >= g.H || xi < 0 || xi >= g.W) continue;
        const uint4* hv = reinterpret_cast<const uint4*>(h + g.vox(n, z, yv, xi) * 32);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 u = hv[q];
            const uint32_t wds[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = q * 8 + 2 * j;
                acc += bf2f((uint16_t)(wds[j] & 0xFFFFu)) * ws[k * 32 + c];
                acc += bf2f((uint16_t)(wds[j] >> 16)) * ws[k * 32 + c + 1];
            }
        }
    }
    const long v = n * per + i;
    y[v] = acc + 0.5f * (bf2f(x[2 * v]) + bf2f(x[2 * v + 1]));
}

// --------------------------------------------------------- MFMA conv 32->32
// Block: 4 wavefronts; wave w computes row (n, z, y0 + w) of 32-voxel tiles
// x0 = 32 t.  Weights [27][32 cout][32 cin] bf16 are staged in LDS.
// MODE 0 (forward):  r = relu(acc + b);  hout = hin + r  (bf16 stores of r, hout)
// MODE 1 (dgrad):    dh[v] += acc  (FP32, in place; weights given transposed)
template <int MODE>
__global__ __launch_bounds__(256) void conv32_mfma_kernel(const uint16_t* __restrict__ in, Geo g,
                                                          const uint16_t* __restrict__ wt, const float* __restrict__ b,
                                                          const uint16_t* __restrict__ hin, uint16_t* __restrict__ rout,
                                                          uint16_t* __restrict__ hout, float* __restrict__ dh)
{
    __shared__ uint4 wl[27 * 32 * 32 * 2 / 16];  // 27 taps x 32 x 32 bf16
    {
        const uint4* src = reinterpret_cast<const uint4*>(wt);
        for (int i = threadIdx.x; i < 27 * 32 * 32 * 2 / 16; i += 256) wl[i] = src[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rows = (g.H + 3) / 4;
    const int n = blockIdx.x / (g.D * rows);
    const int rem = blockIdx.x % (g.D * rows);
    const int z = rem / rows, y = (rem % rows) * 4 + wave;
    if (y >= g.H) return;  // wave-uniform; no block barrier follows
    const int m = lane & 31, kh = lane >> 5;  // tile voxel / K half of this lane
    const uint16_t* wl16 = reinterpret_cast<const uint16_t*>(wl);
    for (int x0 = 0; x0 < g.W; x0 += 32) {
        f32x16 acc = {};
        for (int k = 0; k < 27; ++k) {
            const int zz = z + k / 9 - 1, yy = y + (k / 3) % 3 - 1, xx = x0 + m + k % 3 - 1;
            const bool ok = zz >= 0 && zz < g.D && yy >= 0 && yy < g.H && xx >= 0 && xx < g.W;
            const uint16_t* ip = in + g.vox(n, ok ? zz : 0, ok ? yy : 0, ok ? xx : 0) * 32;
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) {
                const int c0 = 16 * kc + 8 * kh;
                uint4 av = ok ? *reinterpret_cast<const uint4*>(ip + c0) : make_uint4(0, 0, 0, 0);
                const uint4 bv = *reinterpret_cast<const uint4*>(wl16 + (k * 32 + m) * 32 + c0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                              __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
            }
        }
        // lane holds D[voxel mv][channel m] for mv = 8(r/4) + 4 kh + r%4
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int mv = 8 * (r / 4) + 4 * kh + (r % 4);
            const int xx = x0 + mv;
            if (xx >= g.W) continue;
            const long o = g.vox(n, z, y, xx) * 32 + m;
            if constexpr (MODE == 0) {
                const float rv = fmaxf(acc[r] + b[m], 0.f);
                const uint16_t rb = f2bf(rv);
                rout[o] = rb;
                hout[o] = f2bf(bf2f(hin[o]) + bf2f(rb));
            } else {
                dh[o] += acc[r];
            }
        }
    }
}

// ------------------------------------------------------- backward helpers
// loss partials (FP64) and dy = 2 (y - t) / count
__global__ void loss_kernel(const float* __restrict__ y, const float* __restrict__ t, long count,
                            float* __restrict__ dy, double* __restrict__ partial)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
        const double d = (double)y[i] - (double)t[i];
        s += d * d;
        dy[i] = (float)(2.0 * d / (double)count);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// dh5[v][c] = sum_k dy[v - off(k)] W6[k][c]  (FP32 out)
__global__ void out_dgrad_kernel(const float* __restrict__ dy, Geo g, const float* __restrict__ w,
                                 float* __restrict__ dh)
{
    __shared__ float ws[27 * 32];
    for (int i = threadIdx.x; i < 27 * 32; i += blockDim.x) ws[i] = w[i];
    __syncthreads();
    const long per = (long)g.D * g.H * g.W;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= per) return;
    const int xx = (int)(i % g.W), yy = (int)((i / g.W) % g.H), zz = (int)(i / ((long)g.W * g.H));
    float acc[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) acc[c] = 0.f;
    for (int k = 0; k < 27; ++k) {
        const int z = zz - (k / 9 - 1), y = yy - ((k / 3) % 3 - 1), xi = xx - (k % 3 - 1);
        if (z < 0 || z >= g.D || y < 0 || y >= g.H || xi < 0 || xi >= g.W) continue;
        const float d = dy[g.vox(n, z, y, xi)];
#pragma unroll
        for (int c = 0; c < 32; ++c) acc[c] += d * ws[k * 32 + c];
    }
    float* o = dh + (n * per + i) * 32;
#pragma unroll
    for (int c = 0; c < 32; ++c) o[c] = acc[c];
}

// g = bf16(dh * (r > 0))
__global__ void mask_kernel(const float* __restrict__ dh, const uint16_t* __restrict__ r, long count,
                            uint16_t* __restrict__ gout)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    gout[i] = (r[i] & 0x7FFFu) != 0 && !(r[i] & 0x8000u) ? f2bf(dh[i]) : (uint16_t)0;
}

// Weight gradient of the two small layers on the VALU: one block per (n, z)
// plane.  Per output row y the block stages the row's gradients (W x COUT,
// FP32) in LDS, then, one at a time, each of the 9 input rows
// (z + dz - 1, y + dy - 1) over x = -1 .. W (FP32); thread t accumulates the
// products e = t, t + 256, ... of dW[k][co][ci] = sum_v g[v][co] a[v + off(k)][ci]
// whose tap uses the staged row, plus the bias sums.  Layer 1: a = x
// (CIN 2), g bf16 (COUT 32).  Layer 6: a = h5 (CIN 32), g = dy FP32 (COUT 1).
// Requires W <= kMaxW.
constexpr int kMaxW = 256;
template <int CIN, int COUT, bool GF32>
__global__ __launch_bounds__(256) void wgrad_small_kernel(const void* __restrict__ gv, const uint16_t* __restrict__ a,
                                                          Geo g, float* __restrict__ partial)
{
    constexpr int NW = 27 * COUT * CIN, NP = NW + COUT, PER = (NP + 255) / 256;
    __shared__ float gs[kMaxW * COUT];          // [x][co]
    __shared__ float as[(kMaxW + 2) * CIN];     // [x + 1][ci]
    const int n = blockIdx.x / g.D, z = blockIdx.x % g.D;
    float acc[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) acc[j] = 0.f;
    for (int y = 0; y < g.H; ++y) {
        __syncthreads();
        for (int e = threadIdx.x; e < g.W * COUT; e += 256) {
            const long v = g.vox(n, z, y, e / COUT);
            gs[e] = GF32 ? static_cast<const float*>(gv)[v * COUT + e % COUT]
                         : bf2f(static_cast<const uint16_t*>(gv)[v * COUT + e % COUT]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {  // bias sums need only the gradients
            const int e = threadIdx.x + 256 * j;
            if (e >= NW && e < NP) {
                float s = 0.f;
                for (int x = 0; x < g.W; ++x) s += gs[x * COUT + (e - NW)];
                acc[j] += s;
            }
        }
        for (int row = 0; row < 9; ++row) {
            const int zz = z + row / 3 - 1, yy = y + row % 3 - 1;
            const bool rowok = zz >= 0 && zz < g.D && yy >= 0 && yy < g.H;  // block-uniform
            if (!rowok) continue;
            __syncthreads();
            for (int e = threadIdx.x; e < (g.W + 2) * CIN; e += 256) {
                const int xx = e / CIN - 1, ci = e % CIN;
                as[e] = (xx >= 0 && xx < g.W) ? bf2f(a[g.vox(n, zz, yy, xx) * CIN + ci]) : 0.f;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int e = threadIdx.x + 256 * j;
                if (e < NW) {
                    const int k = e / (COUT * CIN), co = (e / CIN) % COUT, ci = e % CIN;
                    if (k / 3 != row) continue;
                    const float* ar = as + (k % 3) * CIN + ci;  // voxel x + dx - 1 -> ar[x * CIN]
                    float s = 0.f;
                    for (int x = 0; x < g.W; ++x) s += gs[x * COUT + co] * ar[x * CIN];
                    acc[j] += s;
                }
            }
        }
    }
    float* p = partial + (long)blockIdx.x * NP;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int e = threadIdx.x + 256 * j;
        if (e < NP) p[e] = acc[j];
    }
}

// MFMA weight gradient of a 32 -> 32 layer: one block per (n, z) plane, 4
// waves; wave w accumulates taps k = w, w+4, ... (<= 7 taps x 16 floats).
// Per 16-voxel chunk of an output row the block stages g (16 vox x 32 co)
// and the 9 shifted input rows (18 vox x 32 ci, x0-1 .. x0+16) in LDS in
// channel-major order, so each lane's A (8 voxels of one co) and B (8 voxels
// of one ci) operands are contiguous.  Bias sums come from the same tiles.
__global__ __launch_bounds__(256) void wgrad_mfma_kernel(const uint16_t* __restrict__ gin,
                                                         const uint16_t* __restrict__ a, Geo g,
                                                         float* __restrict__ partial)
{
    __shared__ __attribute__((aligned(16))) uint16_t gt[32][16];  // [co][voxel]
    __shared__ uint16_t at[9][32][24];       // [row][ci][x - x0 + 1], 18 used (+ pad)
    __shared__ float bias[4][32];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = blockIdx.x / g.D, z = blockIdx.x % g.D;
    const int m = lane & 31, kh = lane >> 5;
    f32x16 acc[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) acc[j] = f32x16{};
    float bsum = 0.f;
    for (int y = 0; y < g.H; ++y)
        for (int x0 = 0; x0 < g.W; x0 += 16) {
            __syncthreads();
            // stage g: 16 voxels x 32 co, transposed
            for (int e = threadIdx.x; e < 16 * 32; e += 256) {
                const int vx = e / 32, co = e % 32;
                const int xx = x0 + vx;
                gt[co][vx] = xx < g.W ? gin[g.vox(n, z, y, xx) * 32 + co] : (uint16_t)0;
            }
            // stage the 9 shifted input rows: 18 voxels x 32 ci each, transposed
            for (int e = threadIdx.x; e < 9 * 18 * 32; e += 256) {
                const int row = e / (18 * 32), vx = (e / 32) % 18, ci = e % 32;
                const int zz = z + row / 3 - 1, yy = y + row % 3 - 1, xx = x0 + vx - 1;
                const bool ok = zz >= 0 && zz < g.D && yy >= 0 && yy < g.H && xx >= 0 && xx < g.W;
                at[row][ci][vx] = ok ? a[g.vox(n, zz, yy, xx) * 32 + ci] : (uint16_t)0;
            }
            __syncthreads();
            // A = g^T: lane (m = co, kh) holds voxels 8 kh .. 8 kh + 7
            const bf16x8 av = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&gt[m][8 * kh]));
            if (wave == 0 && kh == 0) {
#pragma unroll
                for (int j = 0; j < 16; ++j) bsum += bf2f(gt[m][j]);
            }
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int k = wave + 4 * j;
                if (k < 27) {
                    // B = shifted input: lane (n = ci, kh) holds voxels x0 + 8 kh + j' + dx - 1
                    const int row = k / 3, dx = k % 3;
                    bf16x8 bv;
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) bv[jj] = __builtin_bit_cast(__bf16, at[row][m][8 * kh + jj + dx]);
                    acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[j], 0, 0, 0);
                }
            }
        }
    // D[co][ci]: lane holds ci = m, co = 8(r/4) + 4 kh + r%4
    float* p = partial + (long)blockIdx.x * (27 * 32 * 32 + 32);
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const int k = wave + 4 * j;
        if (k < 27) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int co = 8 * (r / 4) + 4 * kh + (r % 4);
                p[(k * 32 + co) * 32 + m] = acc[j][r];
            }
        }
    }
    if (wave == 0 && kh == 0) bias[0][m] = bsum;
    __syncthreads();
    if (threadIdx.x < 32) p[27 * 32 * 32 + threadIdx.x] = bias[0][threadIdx.x];
}

// sum partials[nb][count] -> out[count] (FP64 accumulation)
__global__ void reduce_partials_kernel(const float* __restrict__ partial, int nb, int count, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += (double)partial[(long)b * count + i];
    out[i] = (float)s;
}

// Adam on FP32 master parameters; also refreshes the bf16 copies used by
// the kernels.  Bias-corrected (Kingma & Ba 2015).
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                            float* __restrict__ v, int count, float lr, float b1, float b2, float eps, float c1,
                            float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gg = gr[i];
    const float mm = b1 * m[i] + (1.f - b1) * gg;
    const float vv = b2 * v[i] + (1.f - b2) * gg * gg;
    m[i] = mm;
    v[i] = vv;
    p[i] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// bf16 weight copies of the 32->32 layers: forward [k][co][ci] and the
// transposed, tap-flipped copy for the input gradient [26-k][ci][co]; FP32
// bf16-rounded copies of the small layers.
__global__ void pack_weights_kernel(const float* __restrict__ p, uint16_t* __restrict__ wf, uint16_t* __restrict__ wt,
                                    float* __restrict__ small)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = 27 * 32 * 32;
    if (i < 4 * per) {
        const int l = i / per, e = i % per;
        const int k = e / 1024, co = (e / 32) % 32, ci = e % 32;
        const uint16_t b = f2bf(p[off_w(l + 1) + e]);
        wf[i] = b;
        wt[l * per + ((26 - k) * 32 + ci) * 32 + co] = b;
    }
    if (i < 27 * 32 * 2) small[i] = bf2f(f2bf(p[off_w(0) + i]));
    if (i < 27 * 32) small[27 * 64 + i] = bf2f(f2bf(p[off_w(5) + i]));
}

// Network input of a sub-volume [z0, z0+r.D) x [y0, ...) x [x0, ...) of one
// sample's views (full volumes D x H x W).
__global__ void prep_region_kernel(const uint16_t* const* __restrict__ views, int nviews, int D, int H, int W, int z0,
                                   int y0, int x0, Geo r, float inv_scale, uint16_t* __restrict__ x)
{
    const long per = (long)r.D * r.H * r.W;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per) return;
    const int xx = (int)(i % r.W), yy = (int)((i / r.W) % r.H), zz = (int)(i / ((long)r.W * r.H));
    const long src = ((long)(z0 + zz) * H + (y0 + yy)) * W + (x0 + xx);
    float se = 0.f, so = 0.f;
    int ne = 0, no = 0;
    for (int v = 0; v < nviews; ++v) {
        const float s = (float)views[v][src];
        if (v & 1) {
            so += s;
            ++no;
        } else {
            se += s;
            ++ne;
        }
    }
    const float e = se / (float)ne, o = no ? so / (float)no : e;
    x[2 * i] = f2bf(e * inv_scale);
    x[2 * i + 1] = f2bf(o * inv_scale);
}

// out[z0+z][y0+y][x0+x] = max(y_region[iz0+z][iy0+y][ix0+x] * scale, floor)
__global__ void scale_out_kernel(const float* __restrict__ yr, Geo r, int H, int W, int z0, int y0, int x0, int iz0,
                                 int iy0, int ix0, int nz, int ny, int nx, float scale, float floor_value,
                                 float* __restrict__ out)
{
    const long per = (long)nz * ny * nx;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per) return;
    const int xx = (int)(i % nx), yy = (int)((i / nx) % ny), zz = (int)(i / ((long)nx * ny));
    const float v = yr[r.vox(0, iz0 + zz, iy0 + yy, ix0 + xx)] * scale;
    out[((long)(z0 + zz) * H + (y0 + yy)) * W + (x0 + xx)] = fmaxf(v, floor_value);
}

// ------------------------------------------------------------ launchers
static inline Geo geo(Dims d) { return Geo{d.N, d.D, d.H, d.W}; }
static inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_mfma_layout_test(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s)
{
    mfma_layout_kernel<<<1, 64, 0, s>>>(A, B, D);
    FMVD_LAUNCH_CHECK();
}
void launch_prep_input(const uint16_t* const* d_views, int nviews, Dims d, const float* d_inv_scale, uint16_t* x,
                       hipStream_t s)
{
    const long per = (long)d.D * d.H * d.W;
    prep_input_kernel<<<dim3(nblk(per, 256), d.N), 256, 0, s>>>(d_views, nviews, geo(d), d_inv_scale, x);
    FMVD_LAUNCH_CHECK();
}
void launch_prep_region(const uint16_t* const* d_views, int nviews, int D, int H, int W, int z0, int y0, int x0,
                        Dims r, float inv_scale, uint16_t* x, hipStream_t s)
{
    prep_region_kernel<<<nblk(r.voxels(), 256), 256, 0, s>>>(d_views, nviews, D, H, W, z0, y0, x0, geo(r), inv_scale,
                                                              x);
    FMVD_LAUNCH_CHECK();
}
void launch_pack_weights(const float* params, uint16_t* wf, uint16_t* wt, float* small, hipStream_t s)
{
    pack_weights_kernel<<<nblk(4 * kConv32, 256), 256, 0, s>>>(params, wf, wt, small);
    FMVD_LAUNCH_CHECK();
}
void launch_conv_in(const uint16_t* x, Dims d, const float* small, const float* params, uint16_t* h, hipStream_t s)
{
    const long per = (long)d.D * d.H * d.W;
    conv_in_kernel<<<dim3(nblk(per, 256), d.N), 256, 0, s>>>(x, geo(d), small, params + off_b(0), h);
    FMVD_LAUNCH_CHECK();
}
static inline unsigned conv32_blocks(Dims d) { return (unsigned)((long)d.N * d.D * ((d.H + 3) / 4)); }
void launch_conv32_fwd(const uint16_t* in, Dims d, const uint16_t* wf_layer, const float* bias, uint16_t* rout,
                       uint16_t* hout, hipStream_t s)
{
    conv32_mfma_kernel<0><<<conv32_blocks(d), 256, 0, s>>>(in, geo(d), wf_layer, bias, in, rout, hout, nullptr);
    FMVD_LAUNCH_CHECK();
}
void launch_conv32_dgrad(const uint16_t* g, Dims d, const uint16_t* wt_layer, float* dh, hipStream_t s)
{
    conv32_mfma_kernel<1><<<conv32_blocks(d), 256, 0, s>>>(g, geo(d), wt_layer, nullptr, nullptr, nullptr, nullptr,
                                                           dh);
    FMVD_LAUNCH_CHECK();
}
void launch_conv_out(const uint16_t* h, const uint16_t* x, Dims d, const float* small, const float* params, float* y,
                     hipStream_t s)
{
    const long per = (long)d.D * d.H * d.W;
    conv_out_kernel<<<dim3(nblk(per, 256), d.N), 256, 0, s>>>(h, x, geo(d), small + 27 * 64, params + off_b(5), y);
    FMVD_LAUNCH_CHECK();
}
int launch_loss(const float* y, const float* t, long count, float* dy, double* partial, hipStream_t s)
{
    const int nb = (int)std::min<long>(1024, nblk(count, 256));
    loss_kernel<<<nb, 256, 0, s>>>(y, t, count, dy, partial);
    FMVD_LAUNCH_CHECK();
    return nb;
}
void launch_out_dgrad(const float* dy, Dims d, const float* small, float* dh, hipStream_t s)
{
    const long per = (long)d.D * d.H * d.W;
    out_dgrad_kernel<<<dim3(nblk(per, 256), d.N), 256, 0, s>>>(dy, geo(d), small + 27 * 64, dh);
    FMVD_LAUNCH_CHECK();
}
void launch_mask(const float* dh, const uint16_t* r, long count, uint16_t* g, hipStream_t s)
{
    mask_kernel<<<nblk(count, 256), 256, 0, s>>>(dh, r, count, g);
    FMVD_LAUNCH_CHECK();
}
void launch_wgrad_in(const uint16_t* g, const uint16_t* x, Dims d, float* partial, hipStream_t s)
{
    if (d.W > kMaxW) throw_msg("network: volume width above 256");
    wgrad_small_kernel<2, 32, false><<<d.N * d.D, 256, 0, s>>>(g, x, geo(d), partial);
    FMVD_LAUNCH_CHECK();
}
void launch_wgrad_out(const float* dy, const uint16_t* h5, Dims d, float* partial, hipStream_t s)
{
    if (d.W > kMaxW) throw_msg("network: volume width above 256");
    wgrad_small_kernel<32, 1, true><<<d.N * d.D, 256, 0, s>>>(dy, h5, geo(d), partial);
    FMVD_LAUNCH_CHECK();
}
void launch_wgrad32(const uint16_t* g, const uint16_t* a, Dims d, float* partial, hipStream_t s)
{
    wgrad_mfma_kernel<<<d.N * d.D, 256, 0, s>>>(g, a, geo(d), partial);
    FMVD_LAUNCH_CHECK();
}
void launch_reduce_partials(const float* partial, int nb, int count, float* out, hipStream_t s)
{
    reduce_partials_kernel<<<nblk(count, 256), 256, 0, s>>>(partial, nb, count, out);
    FMVD_LAUNCH_CHECK();
}
void launch_adam(float* p, const float* g, float* m, float* v, int count, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 - std::pow(0.9, step)), c2 = (float)(1.0 - std::pow(0.999, step));
    adam_kernel<<<nblk(count, 256), 256, 0, s>>>(p, g, m, v, count, lr, b1, b2, 1e-8f, c1, c2);
    FMVD_LAUNCH_CHECK();
}
void launch_scale_out(const float* y, Dims r, int D, int H, int W, int z0, int y0, int x0, int iz0, int iy0,
                      int ix0, int nz, int ny, int nx, float scale, float floor_value, float* out, hipStream_t s)
{
    (void)D;
    scale_out_kernel<<<nblk((long)nz * ny * nx, 256), 256, 0, s>>>(y, geo(r), H, W, z0, y0, x0, iz0, iy0, ix0, nz,
                                                                   ny, nx, scale, floor_value, out);
    FMVD_LAUNCH_CHECK();
}

}  // namespace net
}  // namespace fmvd
