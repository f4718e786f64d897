// Warm-start network: parameter layout and kernel launchers.
//
// Flat FP32 parameter vector, layer by layer: W1 [27][32][2], b1 [32],
// W2..W5 [27][32][32], b2..b5 [32], W6 [27][1][32], b6 [1]  (tap-major,
// then output channel, then input channel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fmvd {
namespace net {

constexpr int kC = 32;
constexpr int kLayers = 6;
constexpr int kConv32 = 27 * kC * kC;  // weights of one 32 -> 32 layer
__host__ __device__ constexpr int w_size(int l) { return l == 0 ? 27 * kC * 2 : (l == 5 ? 27 * kC : kConv32); }
__host__ __device__ constexpr int b_size(int l) { return l == 5 ? 1 : kC; }
__host__ __device__ constexpr int off_w(int l)
{
    int o = 0;
    for (int i = 0; i < l; ++i) o += w_size(i) + b_size(i);
    return o;
}
__host__ __device__ constexpr int off_b(int l) { return off_w(l) + w_size(l); }
constexpr int kParams = off_w(kLayers);
constexpr int kSmall = 27 * kC * 2 + 27 * kC;  // bf16-rounded FP32 copies of W1 and W6

struct Dims {
    int N, D, H, W;
    long voxels() const { return (long)N * D * H * W; }
};

// forward / backward building blocks (net_kernels.hip); all enqueue on s
void launch_mfma_layout_test(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s);
void launch_prep_input(const uint16_t* const* d_views, int nviews, Dims d, const float* d_inv_scale, uint16_t* x,
                       hipStream_t s);
void launch_prep_region(const uint16_t* const* d_views, int nviews, int D, int H, int W, int z0, int y0, int x0,
                        Dims r, float inv_scale, uint16_t* x, hipStream_t s);
void launch_pack_weights(const float* params, uint16_t* wf, uint16_t* wt, float* small, hipStream_t s);
void launch_conv_in(const uint16_t* x, Dims d, const float* small, const float* params, uint16_t* h, hipStream_t s);
void launch_conv32_fwd(const uint16_t* in, Dims d, const uint16_t* wf_layer, const float* bias, uint16_t* rout,
                       uint16_t* hout, hipStream_t s);
void launch_conv32_dgrad(const uint16_t* g, Dims d, const uint16_t* wt_layer, float* dh, hipStream_t s);
void launch_conv_out(const uint16_t* h, const uint16_t* x, Dims d, const float* small, const float* params, float* y,
                     hipStream_t s);
// loss partial sums (FP64, one per block of the returned count) and dy
int launch_loss(const float* y, const float* t, long count, float* dy, double* partial, hipStream_t s);
void launch_out_dgrad(const float* dy, Dims d, const float* small, float* dh, hipStream_t s);
void launch_mask(const float* dh, const uint16_t* r, long count, uint16_t* g, hipStream_t s);
// partial weight gradients, one row of (weights + biases) per (n, z) plane
void launch_wgrad_in(const uint16_t* g, const uint16_t* x, Dims d, float* partial, hipStream_t s);
void launch_wgrad_out(const float* dy, const uint16_t* h5, Dims d, float* partial, hipStream_t s);
void launch_wgrad32(const uint16_t* g, const uint16_t* a, Dims d, float* partial, hipStream_t s);
void launch_reduce_partials(const float* partial, int nb, int count, float* out, hipStream_t s);
void launch_adam(float* p, const float* g, float* m, float* v, int count, float lr, int step, hipStream_t s);
void launch_scale_out(const float* y, Dims r, int D, int H, int W, int z0, int y0, int x0, int iz0, int iy0,
                      int ix0, int nz, int ny, int nx, float scale, float floor_value, float* out, hipStream_t s);

}  // namespace net
}  // namespace fmvd
