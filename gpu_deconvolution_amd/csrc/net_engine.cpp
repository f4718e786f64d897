// Warm-start network engine: parameters, Adam state, training step and
// tile-wise inference (C ABI of fmvd_net.h).  No counterpart in the
// reference.  The network's definition is fixed in net.hpp / net_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <random>
#include <vector>

#include "../../include/fmvd_cuda_utils.h"
#include "../../include/fmvd_net.h"
#include "hip_check.hpp"
#include "host_buffers.hpp"
#include "net.hpp"

namespace fmvd {
namespace net {

struct Engine {
    int device = 0;
    DeviceBuffer params, grads, m, v, wf, wt, small;
    DeviceBuffer partial, lossPart;
    // activations for the current shape
    Dims cur{0, 0, 0, 0};
    DeviceBuffer x, y, dy, dh, g;
    DeviceBuffer h[5], r[5];
    int step = 0;
    bool packed = false;

    void init(unsigned long long seed)
    {
        FMVD_HIP(hipGetDevice(&device));
        params.alloc(sizeof(float) * kParams);
        grads.alloc(sizeof(float) * kParams);
        m.alloc(sizeof(float) * kParams);
        v.alloc(sizeof(float) * kParams);
        wf.alloc(sizeof(uint16_t) * 4 * kConv32);
        wt.alloc(sizeof(uint16_t) * 4 * kConv32);
        small.alloc(sizeof(float) * kSmall);
        // He-style normal init (scaled down in the residual blocks, small in
        // the output layer), biases zero
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        std::vector<float> p(kParams, 0.f);
        for (int l = 0; l < kLayers; ++l) {
            const int cin = l == 0 ? 2 : kC;
            double std = std::sqrt(2.0 / (27.0 * cin));
            if (l >= 1 && l <= 4) std *= 0.5;
            if (l == 5) std *= 0.1;
            for (int i = 0; i < w_size(l); ++i) p[off_w(l) + i] = (float)(std * nd(rng));
        }
        set_params(p.data());
    }

    void set_params(const float* host)
    {
        FMVD_HIP(hipMemcpy(params.p, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
        FMVD_HIP(hipMemset(m.p, 0, sizeof(float) * kParams));
        FMVD_HIP(hipMemset(v.p, 0, sizeof(float) * kParams));
        FMVD_HIP(hipMemset(grads.p, 0, sizeof(float) * kParams));
        step = 0;
        packed = false;
    }

    void pack(hipStream_t s)
    {
        if (packed) return;
        launch_pack_weights(params.as<float>(), wf.as<uint16_t>(), wt.as<uint16_t>(), small.as<float>(), s);
        packed = true;
    }

    void shape(Dims d, bool training)
    {
        if (d.N < 1 || d.D < 1 || d.H < 1 || d.W < 1) throw_msg("network: invalid volume shape");
        if (training && d.W > 256) throw_msg("network: training volume width above 256");
        const size_t nv = (size_t)d.voxels();
        x.alloc(sizeof(uint16_t) * 2 * nv);
        y.alloc(sizeof(float) * nv);
        const int nh = training ? 5 : 2;
        for (int i = 0; i < nh; ++i) h[i].alloc(sizeof(uint16_t) * kC * nv);
        if (training) {
            for (int i = 0; i < 5; ++i) r[i].alloc(sizeof(uint16_t) * kC * nv);
            dy.alloc(sizeof(float) * nv);
            dh.alloc(sizeof(float) * kC * nv);
            g.alloc(sizeof(uint16_t) * kC * nv);
            partial.alloc(sizeof(float) * (size_t)d.N * d.D * (kConv32 + kC));
            lossPart.alloc(sizeof(double) * 1024);
        } else {
            r[0].alloc(sizeof(uint16_t) * kC * nv);  // relu scratch of the forward kernels
        }
        cur = d;
    }

    // Forward pass on x (already in this->x); training keeps every layer.
    void forward(Dims d, bool training, hipStream_t s)
    {
        pack(s);
        const float* P = params.as<float>();
        const uint16_t* X = x.as<uint16_t>();
        launch_conv_in(X, d, small.as<float>(), P, h[0].as<uint16_t>(), s);
        for (int l = 1; l <= 4; ++l) {
            const uint16_t* in = training ? h[l - 1].as<uint16_t>() : h[(l - 1) & 1].as<uint16_t>();
            uint16_t* out = training ? h[l].as<uint16_t>() : h[l & 1].as<uint16_t>();
            uint16_t* rr = training ? r[l].as<uint16_t>() : r[0].as<uint16_t>();
            launch_conv32_fwd(in, d, wf.as<uint16_t>() + (size_t)(l - 1) * kConv32, P + off_b(l), rr, out, s);
        }
        const uint16_t* h5 = training ? h[4].as<uint16_t>() : h[0].as<uint16_t>();
        launch_conv_out(h5, X, d, small.as<float>(), P, y.as<float>(), s);
    }

    // Gradients of mean((y - t)^2); returns the loss (FP64 sums on the GPU).
    double backward(Dims d, const float* t, hipStream_t s)
    {
        const long nv = d.voxels();
        const int nb = launch_loss(y.as<float>(), t, nv, dy.as<float>(), lossPart.as<double>(), s);
        float* G = grads.as<float>();
        const int planes = d.N * d.D;
        // layer 6
        launch_wgrad_out(dy.as<float>(), h[4].as<uint16_t>(), d, partial.as<float>(), s);
        launch_reduce_partials(partial.as<float>(), planes, w_size(5) + b_size(5), G + off_w(5), s);
        launch_out_dgrad(dy.as<float>(), d, small.as<float>(), dh.as<float>(), s);
        // layers 5 .. 2
        for (int l = 4; l >= 1; --l) {
            launch_mask(dh.as<float>(), r[l].as<uint16_t>(), nv * kC, g.as<uint16_t>(), s);
            launch_wgrad32(g.as<uint16_t>(), h[l - 1].as<uint16_t>(), d, partial.as<float>(), s);
            launch_reduce_partials(partial.as<float>(), planes, kConv32 + kC, G + off_w(l), s);
            launch_conv32_dgrad(g.as<uint16_t>(), d, wt.as<uint16_t>() + (size_t)(l - 1) * kConv32, dh.as<float>(),
                                s);
        }
        // layer 1 (mask: h1 = relu(a1))
        launch_mask(dh.as<float>(), h[0].as<uint16_t>(), nv * kC, g.as<uint16_t>(), s);
        launch_wgrad_in(g.as<uint16_t>(), x.as<uint16_t>(), d, partial.as<float>(), s);
        launch_reduce_partials(partial.as<float>(), planes, w_size(0) + b_size(0), G + off_w(0), s);
        std::vector<double> part(nb);
        FMVD_HIP(hipMemcpyAsync(part.data(), lossPart.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        FMVD_HIP(hipStreamSynchronize(s));
        double sum = 0.0;
        for (double p : part) sum += p;
        return sum / (double)nv;
    }

    void adam(float lr, hipStream_t s)
    {
        ++step;
        launch_adam(params.as<float>(), grads.as<float>(), m.as<float>(), v.as<float>(), kParams, lr, step, s);
        packed = false;
    }

    // Tile-wise inference: interior tiles of at most (tz, ty, tx) with a halo
    // of `halo` voxels on every side that lies inside the volume.  The
    // receptive field of the network is 6 voxels, so with halo >= 6 the
    // result equals whole-volume inference.
    void infer(const uint16_t* const* d_views, int nviews, int D, int H, int W, float scale, float floor_value,
               float* out, hipStream_t s, int tz = 128, int ty = 224, int tx = 224, int halo = 16)
    {
        const float inv = 1.0f / scale;
        for (int z0 = 0; z0 < D; z0 += tz)
            for (int y0 = 0; y0 < H; y0 += ty)
                for (int x0 = 0; x0 < W; x0 += tx) {
                    const int z1 = std::min(D, z0 + tz), y1 = std::min(H, y0 + ty), x1 = std::min(W, x0 + tx);
                    const int rz0 = std::max(0, z0 - halo), ry0 = std::max(0, y0 - halo), rx0 = std::max(0, x0 - halo);
                    const int rz1 = std::min(D, z1 + halo), ry1 = std::min(H, y1 + halo), rx1 = std::min(W, x1 + halo);
                    const Dims r{1, rz1 - rz0, ry1 - ry0, rx1 - rx0};
                    if (r.voxels() > cur.voxels() || cur.N == 0) shape(r, false);
                    launch_prep_region(d_views, nviews, D, H, W, rz0, ry0, rx0, r, inv, x.as<uint16_t>(), s);
                    forward(r, false, s);
                    launch_scale_out(y.as<float>(), r, D, H, W, z0, y0, x0, z0 - rz0, y0 - ry0, x0 - rx0, z1 - z0,
                                     y1 - y0, x1 - x0, scale, floor_value, out, s);
                }
    }
};

}  // namespace net
}  // namespace fmvd

struct fmvd_net {
    fmvd::net::Engine impl;
};

using fmvd::net::Dims;

extern "C" {

size_t fmvd_net_param_count(void) { return (size_t)fmvd::net::kParams; }

struct fmvd_net* fmvd_net_create(unsigned long long seed)
{
    FMVD_API_GUARD(nullptr, {
        std::unique_ptr<fmvd_net> n(new fmvd_net());
        n->impl.init(seed);
        return n.release();
    })
}

void fmvd_net_destroy(struct fmvd_net* n) { delete n; }

int fmvd_net_get_params(const struct fmvd_net* n, float* host)
{
    FMVD_API_GUARD(-1, {
        if (!n || !host) fmvd::throw_msg("fmvd_net_get_params: null argument");
        FMVD_HIP(hipMemcpy(host, n->impl.params.p, sizeof(float) * fmvd::net::kParams, hipMemcpyDeviceToHost));
        return 0;
    })
}

int fmvd_net_set_params(struct fmvd_net* n, const float* host)
{
    FMVD_API_GUARD(-1, {
        if (!n || !host) fmvd::throw_msg("fmvd_net_set_params: null argument");
        n->impl.set_params(host);
        return 0;
    })
}

int fmvd_net_get_grads(const struct fmvd_net* n, float* host)
{
    FMVD_API_GUARD(-1, {
        if (!n || !host) fmvd::throw_msg("fmvd_net_get_grads: null argument");
        FMVD_HIP(hipMemcpy(host, n->impl.grads.p, sizeof(float) * fmvd::net::kParams, hipMemcpyDeviceToHost));
        return 0;
    })
}

int fmvd_net_forward(struct fmvd_net* n, const void* d_x, int N, int D, int H, int W, float* d_y, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!n || !d_x || !d_y) fmvd::throw_msg("fmvd_net_forward: null argument");
        const Dims d{N, D, H, W};
        hipStream_t s = (hipStream_t)stream;
        n->impl.shape(d, true);
        FMVD_HIP(hipMemcpyAsync(n->impl.x.p, d_x, sizeof(uint16_t) * 2 * d.voxels(), hipMemcpyDeviceToDevice, s));
        n->impl.forward(d, true, s);
        FMVD_HIP(hipMemcpyAsync(d_y, n->impl.y.p, sizeof(float) * d.voxels(), hipMemcpyDeviceToDevice, s));
        return 0;
    })
}

int fmvd_net_backward(struct fmvd_net* n, const void* d_x, const float* d_t, int N, int D, int H, int W,
                      double* loss, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!n || !d_x || !d_t) fmvd::throw_msg("fmvd_net_backward: null argument");
        const Dims d{N, D, H, W};
        hipStream_t s = (hipStream_t)stream;
        n->impl.shape(d, true);
        FMVD_HIP(hipMemcpyAsync(n->impl.x.p, d_x, sizeof(uint16_t) * 2 * d.voxels(), hipMemcpyDeviceToDevice, s));
        n->impl.forward(d, true, s);
        const double l = n->impl.backward(d, d_t, s);
        if (loss) *loss = l;
        return 0;
    })
}

int fmvd_net_adam(struct fmvd_net* n, float lr, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!n) fmvd::throw_msg("fmvd_net_adam: null network");
        n->impl.adam(lr, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_net_prep_input(const unsigned short* const* d_views, int nviews, int N, int D, int H, int W,
                        const float* d_inv_scales, void* d_x, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!d_views || !d_inv_scales || !d_x || nviews < 1) fmvd::throw_msg("fmvd_net_prep_input: bad argument");
        fmvd::net::launch_prep_input(reinterpret_cast<const uint16_t* const*>(d_views), nviews, Dims{N, D, H, W},
                                     d_inv_scales, static_cast<uint16_t*>(d_x), (hipStream_t)stream);
        return 0;
    })
}

int fmvd_net_infer(struct fmvd_net* n, const unsigned short* const* d_views, int nviews, int D, int H, int W,
                   float scale, float floor_value, float* d_out, void* stream)
{
    return fmvd_net_infer_tiled(n, d_views, nviews, D, H, W, scale, floor_value, 128, 224, 224, 16, d_out, stream);
}

int fmvd_net_infer_tiled(struct fmvd_net* n, const unsigned short* const* d_views, int nviews, int D, int H, int W,
                         float scale, float floor_value, int tile_z, int tile_y, int tile_x, int halo, float* d_out,
                         void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!n || !d_views || !d_out || nviews < 1 || !(scale > 0.f)) fmvd::throw_msg("fmvd_net_infer: bad argument");
        if (D < 1 || H < 1 || W < 1 || tile_z < 1 || tile_y < 1 || tile_x < 1 || halo < 6)
            fmvd::throw_msg("fmvd_net_infer: bad shape, tile or halo (halo must cover the 6-voxel receptive field)");
        n->impl.infer(reinterpret_cast<const uint16_t* const*>(d_views), nviews, D, H, W, scale, floor_value, d_out,
                      (hipStream_t)stream, tile_z, tile_y, tile_x, halo);
        return 0;
    })
}

int fmvd_net_mfma_layout_test(const unsigned short* d_a, const unsigned short* d_b, float* d_d)
{
    FMVD_API_GUARD(-1, {
        fmvd::net::launch_mfma_layout_test(d_a, d_b, d_d, nullptr);
        FMVD_HIP(hipDeviceSynchronize());
        return 0;
    })
}

}  // extern "C"
