// Plane-wise multi-view RL engine: plan, batched device execution, plane
// streaming and the C ABI of fmvd_deconvolve.h.
//
// Reference: fmvd_deconvolve.impl.cpp:42-336 (prepareKernels, prepareWeights,
// fmvd_initialize_cuda, fmvd_deconvolve_planes_cuda, fmvd_destroy_cuda).
//
// Per view and iteration the device runs four launches over a chunk of
// planes (the reference runs eight per plane), all in place on one spectrum
// buffer X per plane:
//   A  column pass   X = colconv(X, K_v / H)             X: XE -> spectrum of Q
//   B  row pass      X = R2C(Q > 1 ? D_v / Q : Q), Q = C2R(X)
//   C  column pass   X = colconv(X, Khat_v / (H W))      X: spectrum of C
//   D  row pass      C = C2R(X), E += w_v (E C - E), X = XE = R2C(E)/W
// When Khat_v is K_v flipped about its centre, pass C uses conj(K_v / H) and
// pass B's R2C carries the remaining 1/W.
// E (ES) is the spatial estimate on the padded plane, as in the reference; XE
// is its row spectrum scaled by 1/W, kept so that pass A needs no row FFT.
// XE is dead from pass A's read to pass D's write, so it shares X with the
// intermediate spectra.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fmvd_cuda_utils.h"
#include "../../include/fmvd_deconvolve.h"
#include "hip_check.hpp"
#include "host_buffers.hpp"
#include "host_psf.hpp"
#include "rl_launch.hpp"
#include "timing.hpp"
#include "transform.hpp"
#include "noalloc_scratch.hpp"
#include "autotune_cache.hpp"

namespace fmvd {

static size_t env_bytes(const char* name, size_t dflt)
{
    const char* s = getenv(name);
    if (!s || !*s) return dflt;
    return (size_t)strtoull(s, nullptr, 10);
}

static bool env_flag(const char* name)
{
    const char* s = getenv(name);
    return s && *s && strcmp(s, "0") != 0;
}

// ------------------------------------------------------------------ plan
template <typename S> struct Plan {
    using source_t = int (*)(S**, int);
    using sink_t = void (*)(S*);

    int dataH = 0, dataW = 0, kernelH = 0, kernelW = 0, nViews = 0, nStreams = 0;
    int fftH = 0, fftW = 0, N = 0;
    int device = 0;
    source_t source = nullptr;
    sink_t sink = nullptr;
    fmvd_estimate_sink_t esink = nullptr;
    void* euser = nullptr;
    int batchCap = 0, chunkCap = 0;
    // > 0: deconvolve_planes keeps at most this many planes pulled from the
    // datasource ahead of the datasink, with the reference's call order
    int lookahead = 0;
    int cus = 256;  // compute units of the plan's device
    Variants var;

    DeviceBuffer twW, twRow[3], twCol[3];
    std::vector<DeviceBuffer> K, Kh, Wn;
    // Kernel-hat spectrum = conj(kernel spectrum): independent PSFs (the hat is
    // the kernel flipped about its centre) of odd width and height.  Pass C
    // then reads K with conjugation and Kh is not built; the ratio pass folds
    // in the 1/W that Kh would carry.
    bool khatConj = false;
    DeviceBuffer XE, ES;  // XE: the spectrum buffer X of every pass
    std::unique_ptr<Stream> compute, copy;

    size_t plane() const { return (size_t)dataH * dataW; }
    const float2* khat(int v) const { return (khatConj ? K[v] : Kh[v]).as<float2>(); }
    float ratio_scale() const { return khatConj ? 1.0f / (float)fftW : 1.0f; }
    size_t xplane() const { return (size_t)N * fftH; }  // complex per plane
    size_t eplane() const { return (size_t)fftW * fftH; }  // floats per plane

    RowGeom row_geom() const
    {
        RowGeom g;
        g.H = fftH;
        g.W = fftW;
        g.N = N;
        g.dataH = dataH;
        g.dataW = dataW;
        // data pad uses swapped half-widths (fmvd_deconvolve.impl.cpp:36-37
        // passes (kernelH, kernelW) into parameters named (kernelW, kernelH),
        // fmvd_deconvolve_cuda.impl.cu:76-77,87-88)
        g.dkY = kernelW / 2;
        g.dkX = kernelH / 2;
        g.wkY = kernelH / 2;
        g.wkX = kernelW / 2;
        for (int v = 0; v < 3; ++v) g.twP[v] = twRow[v].as<float2>();
        g.twW = twW.as<float2>();
        return g;
    }
    ColGeom col_geom() const
    {
        ColGeom g;
        g.H = fftH;
        g.N = N;
        for (int v = 0; v < 3; ++v) g.twP[v] = twCol[v].as<float2>();
        return g;
    }

    int batch_planes() const
    {
        if (batchCap > 0) return batchCap;
        const size_t budget = env_bytes("FMVD_BATCH_BYTES", (size_t)2 << 30);
        const size_t per = xplane() * sizeof(float2) + eplane() * sizeof(float);
        size_t nb = budget / (per ? per : 1);
        nb = std::max<size_t>(1, std::min<size_t>(nb, 4096));
        return (int)std::max<size_t>(nb, (size_t)std::max(1, nStreams));
    }

    // Work buffers of the running call: the plan's own (XE, ES), or, in the
    // allocation-free entry point, slices of the device scratch array.
    float2* workX = nullptr;
    float* workE = nullptr;
    float2* xbase() const { return workX ? workX : XE.as<float2>(); }
    float* ebase() const { return workX ? workE : ES.as<float>(); }

    void ensure_work(int planes)
    {
        if (workX) return;  // scratch slices are sized by the caller
        XE.alloc(xplane() * sizeof(float2) * (size_t)planes);
        ES.alloc(eplane() * sizeof(float) * (size_t)planes);
    }

    void init(int dH, int dW, const float* const* h_Weights, float** h_Kernel, int kH, int kW,
              int psf, int nV, int nS, source_t src, sink_t snk)
    {
        if (dH < 1 || dW < 1 || kH < 1 || kW < 1 || nV < 1)
            throw_msg("invalid plan dimensions (dataH, dataW, kernelH, kernelW, nViews must be >= 1)");
        if (psf < 0 || psf > 3) throw_msg("invalid PSF type (expected 0..3)");
        if (!h_Kernel) throw_msg("kernels must not be NULL");
        dataH = dH;
        dataW = dW;
        kernelH = kH;
        kernelW = kW;
        nViews = nV;
        nStreams = nS < 1 ? 1 : nS;
        source = src;
        sink = snk;
        fftH = snap_transform_size(dataH + kernelH - 1);
        fftW = snap_transform_size(dataW + kernelW - 1);
        N = fftW / 2;
        if (!row_size_supported(N) || !col_size_supported(fftH) || fftW < 16 || fftH < 16) {
            char msg[160];
            snprintf(msg, sizeof msg, "unsupported FFT extent %d x %d (data %d x %d, kernel %d x %d)",
                     fftH, fftW, dataH, dataW, kernelH, kernelW);
            throw_msg(msg);
        }
        FMVD_HIP(hipGetDevice(&device));
        FMVD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        compute.reset(new Stream());
        copy.reset(new Stream());

        // fmvd_deconvolve.impl.cpp:61-64: kernel-hat from the raw kernels,
        // then the caller's kernels are normalised in place.
        const int len = kernelH * kernelW;
        std::vector<double> khat = compute_kernel_hat(h_Kernel, kernelW, kernelH, nViews, psf);
        khatConj = psf == 0 && (kernelW % 2) == 1 && (kernelH % 2) == 1 && !env_flag("FMVD_EXPLICIT_KHAT");
        for (int v = 0; v < nViews; ++v) normalize_kernel_f32(h_Kernel[v], len);

        upload(twW, twiddles(fftW, N));
        for (int v = 0; v < 3; ++v) {
            // variants that do not exist fall back to plan 0 (same table)
            upload(twRow[v], row_pass_twiddles(N, v));
            upload(twCol[v], col_pass_twiddles(fftH, v));
        }

        K.resize(nViews);
        Kh.resize(nViews);
        Wn.resize(nViews);
        std::vector<double> kv(len);
        for (int v = 0; v < nViews; ++v) {
            for (int i = 0; i < len; ++i) kv[i] = (double)h_Kernel[v][i];
            // pass A consumes XE = R2C(E)/W: fold 1/H here; pass C consumes
            // R2C(R): fold 1/(H W) (modulateAndNormalize's c, fmvd_deconvolve_cuda.cu:369)
            upload(K[v], kernel_spectrum_transposed(kv.data(), kernelW, kernelH, fftH, fftW, 1.0 / fftH));
            if (!khatConj)
                upload(Kh[v], kernel_spectrum_transposed(khat.data() + (size_t)v * len, kernelW, kernelH,
                                                         fftH, fftW, 1.0 / ((double)fftH * fftW)));
        }
        // NULL weights (an extension): every view weighs 1 until
        // set_weights_device replaces them
        std::vector<float> ones;
        std::vector<const float*> wp(nViews);
        if (!h_Weights) ones.assign(plane(), 1.0f);
        for (int v = 0; v < nViews; ++v) {
            wp[v] = h_Weights ? h_Weights[v] : ones.data();
            if (!wp[v]) throw_msg("weight image must not be NULL");
        }
        std::vector<float> wn = normalized_weights(wp.data(), dataW, dataH, nViews);
        for (int v = 0; v < nViews; ++v) {
            std::vector<float> one(wn.begin() + (size_t)v * plane(), wn.begin() + (size_t)(v + 1) * plane());
            upload(Wn[v], one);
        }
    }

    // RL on `np` resident planes; views[v] points at the first plane of view v.
    // Optional convergence trace: when set, run_batch adds after ingest and
    // after every iteration k the sums (E - ref)^2 and ref^2 of its planes
    // into errAcc[2k], errAcc[2k + 1] (device FP64).
    const float* errRef = nullptr;
    double* errAcc = nullptr;

    // Cache-resident chunks.  Every plane's RL is independent, so run_batch
    // splits its planes into chunks and runs each chunk's whole iteration
    // chain before the next.  One iteration of a chunk touches its X and ES
    // (19 MB per plane at FFT 1536^2), plus every view's kernel spectra (one
    // when the hat is the conjugate, else two) and weights; the samples are
    // streamed with non-temporal loads and stay out of the cache.  The chunk
    // is the power of two of planes for which this fits a 200 MiB budget in
    // the 256 MB memory-side cache (8 planes at FFT 1536^2 with 4 views).  A
    // launch must still fill the chip: `fill` planes give one residency of
    // the column pass (12 sequences per CU).  When the cache cannot hold that
    // many planes, the chunk is 4 x fill planes, so that every launch runs
    // several residency rounds (16 planes at FFT 2560^2).
    int chunk_planes() const
    {
        if (chunkCap > 0) return chunkCap;
        const size_t budget = env_bytes("FMVD_CHUNK_BYTES", (size_t)200 << 20);
        const size_t spectra = khatConj ? 1 : 2;
        const size_t shared = (size_t)nViews * (spectra * xplane() * sizeof(float2) + plane() * sizeof(float));
        const size_t per = xplane() * sizeof(float2) + eplane() * sizeof(float);
        const size_t fit = budget > shared ? (budget - shared) / per : 0;
        auto pow2_floor = [](size_t n) {
            size_t p = 1;
            while (p * 2 <= n) p *= 2;
            return p;
        };
        const size_t seqs = (size_t)std::max(1, cus) * 12, want = (seqs + (size_t)N - 1) / (size_t)N;
        size_t fill = 1;
        while (fill < want) fill *= 2;
        const size_t n = fit >= fill ? pow2_floor(std::min<size_t>(fit, 4096)) : std::min<size_t>(4 * fill, 4096);
        return (int)n;
    }

    void run_batch(const S* const* views, long dstride, int np, int iterations, const float* init,
                   long istride, S* out, float* est, long ostride, hipStream_t s, long refOffset = 0)
    {
        ensure_work(np);
        const int c = std::max(1, std::min(np, chunk_planes()));
        const int lanes = (errAcc || workX) ? 1 : (int)std::max<size_t>(1, std::min<size_t>(2, env_bytes("FMVD_LANES", 1)));
        if (lanes > 1 && np > c) {
            if (!lane2) {
                lane2.reset(new Stream());
                laneFork.reset(new Event());
                laneJoin.reset(new Event());
            }
            FMVD_HIP(hipEventRecord(laneFork->e, s));
            FMVD_HIP(hipStreamWaitEvent(lane2->s, laneFork->e, 0));
        }
        for (int start = 0, i = 0; start < np; start += c, ++i)
            run_chunk(views, dstride, start, std::min(c, np - start), iterations, init, istride, out, est, ostride,
                      (lanes > 1 && (i & 1)) ? lane2->s : s, refOffset, !workX && i % kTimingStride == 0);
        if (lanes > 1 && np > c) {
            FMVD_HIP(hipEventRecord(laneJoin->e, lane2->s));
            FMVD_HIP(hipStreamWaitEvent(s, laneJoin->e, 0));
        }
    }
    std::unique_ptr<Stream> lane2;
    std::unique_ptr<Event> laneFork, laneJoin;
    // With kernel timing on, every kTimingStride-th chunk's launches are timed.
    static constexpr int kTimingStride = 8;

    // The iteration chain on planes [start, start + np) of a batch.
    void run_chunk(const S* const* views, long dstride, int start, int np, int iterations, const float* init,
                   long istride, S* out, float* est, long ostride, hipStream_t s, long refOffset, bool timed)
    {
        const RowGeom rg = row_geom();
        const ColGeom cg = col_geom();
        const long xs = (long)xplane();
        const long esd = (long)eplane();
        float2* xe = xbase() + (size_t)start * xs;
        float2* xt = xe;  // passes A-C write the spectrum in place
        float* es = ebase() + (size_t)start * esd;
        const float* ref = errRef ? errRef + refOffset + (size_t)start * plane() : nullptr;
        if (init) init += (size_t)start * istride;
        if (out) out += (size_t)start * ostride;
        if (est) est += (size_t)start * ostride;
        {
            TimedScope ts(kTimeIngest, s, timed);
            if (init) launch_row_init(rg, xe, es, xs, esd, init, istride, np, var.ingest, s);
            else launch_const_init(rg, xe, es, xs, esd, 10.0f * (float)nViews, np, s);
        }
        if (errAcc) launch_error_sums(rg, es, esd, ref, (long)plane(), np, errAcc, s);
        for (int it = 0; it < iterations; ++it) {
            for (int v = 0; v < nViews; ++v) {
                const S* dv = views[v] + (size_t)start * dstride;
                {
                    TimedScope ts(kTimeColumn, s, timed);
                    launch_col_conv(cg, xe, xt, xs, K[v].as<float2>(), np, var.column, s);
                }
                {
                    TimedScope ts(kTimeRatio, s, timed);
                    launch_row_ratio<S>(rg, xt, xs, dv, dstride, np, var.ratio, s, ratio_scale());
                }
                {
                    TimedScope ts(kTimeColumn, s, timed);
                    launch_col_conv(cg, xt, xt, xs, khat(v), np, var.column, s, khatConj);
                }
                {
                    TimedScope ts(kTimeUpdate, s, timed);
                    launch_row_update(rg, xe, xt, es, xs, esd, Wn[v].as<float>(), np, var.update, s);
                }
            }
            if (errAcc) launch_error_sums(rg, es, esd, ref, (long)plane(), np, errAcc + 2 * (it + 1), s);
        }
        TimedScope ts(kTimeEgress, s, timed);
        launch_final<S>(rg, es, esd, out, est, ostride, np, s);
    }

    // Autotuner: times every compiled FFT-plan variant of each kernel kind on
    // this plan's own buffers (planes planes, synthetic content) and keeps the
    // fastest.  times[kind*3 + variant] receives the mean launch time in ms
    // (negative where the variant does not exist).
    // Key of this plan's tuning in the autotuner cache file: device, extents,
    // sample type, kernel-hat form and the planes the launches cover.
    std::string tune_key(int planes) const
    {
        hipDeviceProp_t prop;
        FMVD_HIP(hipGetDeviceProperties(&prop, device));
        char key[256];
        snprintf(key, sizeof key, "%s|%d|%dx%d|u%d|c%d|p%d", prop.gcnArchName, cus, fftH, fftW, (int)(8 * sizeof(S)),
                 khatConj ? 1 : 0, planes);
        for (char* c = key; *c; ++c)
            if (*c == ' ') *c = '_';
        return key;
    }

    void autotune(int planes, float* times)
    {
        FMVD_HIP(hipSetDevice(device));
        planes = std::max(1, std::min({planes > 0 ? planes : 32, batch_planes(), chunk_planes()}));
        const std::string key = tune_key(planes);
        TuneRecord rec;
        if (autotune_cache_find(key, rec)) {
            int v[5];
            std::copy(rec.chosen, rec.chosen + 5, v);
            if (valid_variants(v)) {
                var = Variants{v[0], v[1], v[2], v[3], v[4]};
                if (times) std::copy(rec.times, rec.times + 15, times);
                return;
            }
        }
        float measured[15];
        autotune_measure(planes, measured);
        rec.chosen[0] = var.column;
        rec.chosen[1] = var.ratio;
        rec.chosen[2] = var.update;
        rec.chosen[3] = var.ingest;
        rec.chosen[4] = var.egress;
        std::copy(measured, measured + 15, rec.times);
        autotune_cache_store(key, rec);
        if (times) std::copy(measured, measured + 15, times);
    }

    bool valid_variants(const int* v) const
    {
        const int nc = col_variants(fftH), nr = row_variants(N);
        for (int k = 0; k < 4; ++k)
            if (v[k] < 0 || v[k] >= (k == 0 ? nc : nr)) return false;
        return v[4] == 0;
    }

    void autotune_measure(int planes, float* times)
    {
        ensure_work(planes);
        const size_t pl = plane();
        DeviceBuffer data, out, e0;
        data.alloc(sizeof(S) * pl * planes);
        out.alloc(sizeof(S) * pl * planes);
        e0.alloc(sizeof(float) * pl * planes);
        hipStream_t s = compute->s;
        FMVD_HIP(hipMemsetAsync(data.p, 0x3c, data.bytes, s));
        FMVD_HIP(hipMemsetAsync(e0.p, 0x42, e0.bytes, s));  // ~ 48.6f
        const RowGeom rg = row_geom();
        const ColGeom cg = col_geom();
        const long xs = (long)xplane();
        float2* xe = XE.as<float2>();
        float2* xt = xe;
        float* es = ES.as<float>();
        const long esd = (long)eplane();
        launch_const_init(rg, xe, es, xs, esd, 10.0f * (float)nViews, planes, s);
        launch_col_conv(cg, xe, xt, xs, K[0].as<float2>(), planes, 0, s);
        hipEvent_t a, b;
        FMVD_HIP(hipEventCreate(&a));
        FMVD_HIP(hipEventCreate(&b));
        auto time_it = [&](auto&& launch) -> float {
            launch();  // warm-up
            FMVD_HIP(hipEventRecord(a, s));
            const int reps = 10;
            for (int r = 0; r < reps; ++r) launch();
            FMVD_HIP(hipEventRecord(b, s));
            FMVD_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            FMVD_HIP(hipEventElapsedTime(&ms, a, b));
            return ms / reps;
        };
        const int nc = col_variants(fftH), nr = row_variants(N);
        int* slots[5] = {&var.column, &var.ratio, &var.update, &var.ingest, &var.egress};
        for (int k = 0; k < 5; ++k) {
            const int nv = k == 0 ? nc : (k == 4 ? 1 : nr);
            float best = 0.f;
            int bestv = 0;
            for (int v = 0; v < 3; ++v) {
                float ms = -1.f;
                if (v < nv) {
                    switch (k) {
                    case 0: ms = time_it([&] { launch_col_conv(cg, xt, xt, xs, khat(0), planes, v, s, khatConj); }); break;
                    case 1: ms = time_it([&] { launch_row_ratio<S>(rg, xt, xs, data.as<S>(), (long)pl, planes, v, s, ratio_scale()); }); break;
                    case 2: ms = time_it([&] { launch_row_update(rg, xe, xt, es, xs, esd, Wn[0].as<float>(), planes, v, s); }); break;
                    case 3: ms = time_it([&] { launch_row_init(rg, xt, es, xs, esd, e0.as<float>(), (long)pl, planes, v, s); }); break;
                    case 4: ms = time_it([&] { launch_final<S>(rg, es, esd, out.as<S>(), nullptr, (long)pl, planes, s); }); break;
                    }
                    if (v == 0 || ms < best) {
                        best = ms;
                        bestv = v;
                    }
                }
                if (times) times[k * 3 + v] = ms;
            }
            *slots[k] = bestv;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        FMVD_HIP(hipStreamSynchronize(s));
    }

    // Replaces the plan's weights with nViews device weight images (dataH x
    // dataW floats each), normalised over views on the device in the host
    // rule's float operations.  Enqueued on s: later runs on s see them.
    DeviceBuffer wsum;
    void set_weights_device(const float* const* dw, hipStream_t s)
    {
        if (!dw) throw_msg("d_weights must not be NULL");
        for (int v = 0; v < nViews; ++v)
            if (!dw[v]) throw_msg("weight image must not be NULL");
        wsum.alloc(sizeof(float) * plane());
        const long n = (long)plane();
        for (int v = 0; v < nViews; ++v) launch_weight_accumulate(wsum.as<float>(), dw[v], n, v == 0, s);
        for (int v = 0; v < nViews; ++v) launch_weight_normalize(Wn[v].as<float>(), dw[v], wsum.as<float>(), n, s);
    }

    // GPU-resident chain: view v's acquired volume (w x h x d, device) is
    // transformed with inv[v] into registered[v] (nplanes planes of this
    // plan's dataH x dataW); the border mask of every view becomes the
    // plan's weights; then RL as run_volume.  Nothing returns to the host.
    DeviceBuffer masks;
    void reconstruct_device(const S* const* acquired, int w, int h, int d, const float* const* inv, int border,
                            float zspacing, S* const* registered, int nplanes, int iterations, S* out, float* est,
                            hipStream_t s)
    {
        if (!acquired || !inv || !registered) throw_msg("reconstruct: NULL argument");
        if (nplanes < 1) throw_msg("reconstruct: nPlanes must be >= 1");
        masks.alloc(sizeof(float) * plane() * nViews);
        std::vector<const float*> mp(nViews);
        std::vector<const S*> rp(nViews);
        for (int v = 0; v < nViews; ++v) {
            if (!acquired[v] || !inv[v] || !registered[v]) throw_msg("reconstruct: NULL view");
            transform_volume<S>(acquired[v], w, h, d, registered[v], dataW, dataH, nplanes, inv[v], s);
            float* m = masks.as<float>() + (size_t)v * plane();
            transform_mask(w, h, d, dataW, dataH, nplanes, inv[v], border, zspacing, m, s);
            mp[v] = m;
            rp[v] = registered[v];
        }
        set_weights_device(mp.data(), s);
        run_volume(rp.data(), nplanes, iterations, nullptr, out, est, s);
    }

    void run_volume(const S* const* dviews, int nplanes, int iterations, const float* init, S* out,
                    float* est, hipStream_t s)
    {
        if (nplanes < 0 || iterations < 0) throw_msg("negative plane or iteration count");
        if (!dviews) throw_msg("d_views must not be NULL");
        const int nb = batch_planes();
        const size_t pl = plane();
        std::vector<const S*> vv(nViews);
        for (int start = 0; start < nplanes; start += nb) {
            const int n = std::min(nb, nplanes - start);
            for (int v = 0; v < nViews; ++v) vv[v] = dviews[v] + (size_t)start * pl;
            run_batch(vv.data(), (long)pl, n, iterations, init ? init + (size_t)start * pl : nullptr,
                      (long)pl, out ? out + (size_t)start * pl : nullptr,
                      est ? est + (size_t)start * pl : nullptr, (long)pl, s, (long)start * (long)pl);
        }
    }

    // run_volume with the work buffers in the device scratch array instead of
    // the plan's own: no device memory is allocated or freed.  Planes run in
    // batches of as many as the scratch holds.  Every call on one device
    // shares that device's scratch, so calls that overlap on one GPU race.
    static constexpr int kNoAllocMaxViews = 64;
    void run_volume_noalloc(const S* const* dviews, int nplanes, int iterations, const float* init, S* out,
                            float* est, hipStream_t s)
    {
        if (nplanes < 0 || iterations < 0) throw_msg("negative plane or iteration count");
        if (!dviews) throw_msg("d_views must not be NULL");
        if (nViews > kNoAllocMaxViews) throw_msg("the allocation-free entry point takes at most 64 views");
        const size_t xb = xplane() * sizeof(float2), eb = eplane() * sizeof(float);
        const size_t cap = kNoAllocScratchBytes / (xb + eb);
        if (cap < 1) {
            char msg[200];
            snprintf(msg, sizeof msg,
                     "FFT extent %d x %d needs %zu bytes of scratch per plane; the allocation-free entry point "
                     "has %zu", fftH, fftW, xb + eb, kNoAllocScratchBytes);
            throw_msg(msg);
        }
        unsigned char* base = static_cast<unsigned char*>(noalloc_scratch());
        const int nb = (int)std::min<size_t>((size_t)batch_planes(), cap);
        struct Restore {
            Plan* p;
            ~Restore() { p->workX = nullptr; p->workE = nullptr; }
        } restore{this};
        workX = reinterpret_cast<float2*>(base);
        workE = reinterpret_cast<float*>(base + (size_t)nb * xb);
        const size_t pl = plane();
        const S* vv[kNoAllocMaxViews];
        for (int start = 0; start < nplanes; start += nb) {
            const int n = std::min(nb, nplanes - start);
            for (int v = 0; v < nViews; ++v) vv[v] = dviews[v] + (size_t)start * pl;
            run_batch(vv, (long)pl, n, iterations, init ? init + (size_t)start * pl : nullptr, (long)pl,
                      out ? out + (size_t)start * pl : nullptr, est ? est + (size_t)start * pl : nullptr, (long)pl,
                      s, (long)start * (long)pl);
        }
    }

        // RL on a GPU-resident volume recording, after ingest and after every
    // iteration, the relative L2 distance of the estimate (data region) to
    // ref; errs receives iterations + 1 values.
    void convergence(const S* const* dviews, int nplanes, int iterations, const float* init, const float* ref,
                     double* errs, hipStream_t s)
    {
        if (!ref || !errs) throw_msg("convergence trace needs a reference volume and an output array");
        DeviceBuffer acc;
        acc.alloc(sizeof(double) * 2 * (size_t)(iterations + 1));
        FMVD_HIP(hipMemsetAsync(acc.p, 0, acc.bytes, s));
        errRef = ref;
        errAcc = acc.as<double>();
        try {
            run_volume(dviews, nplanes, iterations, init, nullptr, nullptr, s);
        } catch (...) {
            errRef = nullptr;
            errAcc = nullptr;
            throw;
        }
        errRef = nullptr;
        errAcc = nullptr;
        std::vector<double> h(2 * (size_t)(iterations + 1));
        FMVD_HIP(hipMemcpyAsync(h.data(), acc.p, acc.bytes, hipMemcpyDeviceToHost, s));
        FMVD_HIP(hipStreamSynchronize(s));
        for (int k = 0; k <= iterations; ++k) errs[k] = h[2 * k + 1] > 0 ? std::sqrt(h[2 * k] / h[2 * k + 1]) : 0.0;
    }

    // fmvd_deconvolve.impl.cpp:205-300, as a double-buffered batch pipeline:
    // while batch i computes on the compute stream, the host pulls batch i+1
    // from the datasource and uploads it on the copy stream, and pushes batch
    // i-1's results to the datasink.  Planes are pulled and pushed in z order.
    void deconvolve_planes(int iterations)
    {
        if (!source || !sink) throw_msg("datasource and datasink must be set");
        if (iterations < 0) throw_msg("negative iteration count");
        FMVD_HIP(hipSetDevice(device));
        if (lookahead > 0) return deconvolve_planes_bounded(iterations);
        const int nb = batch_planes();
        const size_t pl = plane();
        const size_t inBytes = sizeof(S) * pl * nViews * (size_t)nb;
        DeviceBuffer din[2], dout, dest;
        PinnedBuffer hin[2], hout[2], hest[2];
        Event up[2], done[2];
        for (int k = 0; k < 2; ++k) {
            din[k].alloc(inBytes);
            hin[k].alloc(inBytes);
            hout[k].alloc(sizeof(S) * pl * nb);
            if (esink) hest[k].alloc(sizeof(float) * pl * nb);
        }
        dout.alloc(sizeof(S) * pl * nb);
        if (esink) dest.alloc(sizeof(float) * pl * nb);
        ensure_work(nb);

        std::vector<S*> hviews[2];
        for (int k = 0; k < 2; ++k)
            for (int v = 0; v < nViews; ++v) hviews[k].push_back(hin[k].as<S>() + (size_t)v * nb * pl);
        bool exhausted = false;
        auto fill = [&](int slot) -> int {
            int n = 0;
            while (n < nb && !exhausted) {
                if (!source(hviews[slot].data(), (int)(n * pl))) {
                    exhausted = true;
                    break;
                }
                ++n;
            }
            return n;
        };
        auto send = [&](int slot, int n) {
            for (int v = 0; v < nViews; ++v)
                FMVD_HIP(hipMemcpyAsync(din[slot].as<S>() + (size_t)v * nb * pl, hviews[slot][v],
                                        sizeof(S) * pl * n, hipMemcpyHostToDevice, copy->s));
            FMVD_HIP(hipEventRecord(up[slot].e, copy->s));
        };
        int zout = 0;
        auto drain = [&](int slot, int n) {
            FMVD_HIP(hipEventSynchronize(done[slot].e));
            for (int j = 0; j < n; ++j) {
                if (esink) esink(hest[slot].as<float>() + (size_t)j * pl, zout, euser);
                sink(hout[slot].as<S>() + (size_t)j * pl);
                ++zout;
            }
        };

        int cnt[2] = {0, 0};
        int slot = 0;
        cnt[0] = fill(0);
        if (cnt[0] == 0) return;
        send(0, cnt[0]);
        bool prev = false;
        std::vector<const S*> vv(nViews);
        for (;;) {
            hipStream_t cs = compute->s;
            FMVD_HIP(hipStreamWaitEvent(cs, up[slot].e, 0));
            for (int v = 0; v < nViews; ++v) vv[v] = din[slot].as<S>() + (size_t)v * nb * pl;
            run_batch(vv.data(), (long)pl, cnt[slot], iterations, nullptr, 0, dout.as<S>(),
                      esink ? dest.as<float>() : nullptr, (long)pl, cs);
            FMVD_HIP(hipMemcpyAsync(hout[slot].p, dout.p, sizeof(S) * pl * cnt[slot],
                                    hipMemcpyDeviceToHost, cs));
            if (esink)
                FMVD_HIP(hipMemcpyAsync(hest[slot].p, dest.p, sizeof(float) * pl * cnt[slot],
                                        hipMemcpyDeviceToHost, cs));
            FMVD_HIP(hipEventRecord(done[slot].e, cs));
            const int ns = slot ^ 1;
            if (prev) drain(ns, cnt[ns]);
            cnt[ns] = exhausted ? 0 : fill(ns);
            if (cnt[ns] > 0) send(ns, cnt[ns]);
            if (cnt[ns] == 0) {
                drain(slot, cnt[slot]);
                break;
            }
            prev = true;
            slot = ns;
        }
    }

    // fmvd_deconvolve.impl.cpp:205-300 call order with at most L = lookahead
    // planes pulled ahead of the sink: L source calls, then per plane a sink
    // call followed by a source call (until one returns 0), then the sinks
    // of the remaining planes.  Planes run in batches of L: the batch's
    // results are sunk one by one, and each sink makes room for the next
    // plane's source call.  A source that runs dry during the first L calls
    // ends the stream there (the reference kept calling it).
    void deconvolve_planes_bounded(int iterations)
    {
        const int nb = lookahead;
        const size_t pl = plane();
        DeviceBuffer din, dout, dest;
        PinnedBuffer hin, hout, hest;
        din.alloc(sizeof(S) * pl * nViews * (size_t)nb);
        hin.alloc(sizeof(S) * pl * nViews * (size_t)nb);
        dout.alloc(sizeof(S) * pl * nb);
        hout.alloc(sizeof(S) * pl * nb);
        if (esink) {
            dest.alloc(sizeof(float) * pl * nb);
            hest.alloc(sizeof(float) * pl * nb);
        }
        ensure_work(nb);
        std::vector<S*> hviews(nViews);
        std::vector<const S*> dviews(nViews);
        for (int v = 0; v < nViews; ++v) {
            hviews[v] = hin.as<S>() + (size_t)v * nb * pl;
            dviews[v] = din.as<S>() + (size_t)v * nb * pl;
        }
        bool exhausted = false;
        int n = 0;
        while (n < nb && source(hviews.data(), (int)(n * pl))) ++n;
        exhausted = n < nb;
        hipStream_t cs = compute->s;
        int zout = 0;
        while (n > 0) {
            for (int v = 0; v < nViews; ++v)
                FMVD_HIP(hipMemcpyAsync(const_cast<S*>(dviews[v]), hviews[v], sizeof(S) * pl * n,
                                        hipMemcpyHostToDevice, cs));
            run_batch(dviews.data(), (long)pl, n, iterations, nullptr, 0, dout.as<S>(),
                      esink ? dest.as<float>() : nullptr, (long)pl, cs);
            FMVD_HIP(hipMemcpyAsync(hout.p, dout.p, sizeof(S) * pl * n, hipMemcpyDeviceToHost, cs));
            if (esink)
                FMVD_HIP(hipMemcpyAsync(hest.p, dest.p, sizeof(float) * pl * n, hipMemcpyDeviceToHost, cs));
            FMVD_HIP(hipStreamSynchronize(cs));
            int m = 0;
            for (int j = 0; j < n; ++j) {
                if (esink) esink(hest.as<float>() + (size_t)j * pl, zout, euser);
                sink(hout.as<S>() + (size_t)j * pl);
                ++zout;
                if (!exhausted) {
                    if (source(hviews.data(), (int)(m * pl))) ++m;
                    else exhausted = true;
                }
            }
            n = m;
        }
    }
};

}  // namespace fmvd

// ---------------------------------------------------------------- C ABI
struct fmvd_plan_cuda_16 {
    fmvd::Plan<uint16_t> impl;
};
struct fmvd_plan_cuda_8 {
    fmvd::Plan<uint8_t> impl;
};

namespace {

template <typename PlanT, typename S, typename Src, typename Snk>
PlanT* initialize(int dataH, int dataW, float const* const* h_Weights, float** h_Kernel, int kernelH,
                  int kernelW, fmvd_psf_type t, int nViews, int nstreams, Src src, Snk snk)
{
    FMVD_API_GUARD(nullptr, {
        std::unique_ptr<PlanT> p(new PlanT());
        p->impl.init(dataH, dataW, h_Weights, h_Kernel, kernelH, kernelW, (int)t, nViews, nstreams,
                     reinterpret_cast<int (*)(S**, int)>(src), reinterpret_cast<void (*)(S*)>(snk));
        return p.release();
    })
}

template <typename PlanT> int plan_info(const PlanT* p, int* dims)
{
    if (!p || !dims) return -1;
    dims[0] = p->impl.dataH;
    dims[1] = p->impl.dataW;
    dims[2] = p->impl.fftH;
    dims[3] = p->impl.fftW;
    dims[4] = p->impl.nViews;
    dims[5] = p->impl.batch_planes();
    dims[6] = p->impl.chunk_planes();
    return 0;
}

template <typename PlanT> static int autotune_plan(PlanT* plan, int planes, int* chosen, float* times_ms)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        plan->impl.autotune(planes, times_ms);
        if (chosen) {
            const fmvd::Variants& v = plan->impl.var;
            chosen[0] = v.column;
            chosen[1] = v.ratio;
            chosen[2] = v.update;
            chosen[3] = v.ingest;
            chosen[4] = v.egress;
        }
        return 0;
    })
}

template <typename PlanT> static int set_variants(PlanT* plan, const int* v)
{
    if (!plan || !v) return -1;
    if (!plan->impl.valid_variants(v)) return -1;  // (the egress kernel has no FFT plan: v[4] must be 0)
    fmvd::Variants& x = plan->impl.var;
    x.column = v[0];
    x.ratio = v[1];
    x.update = v[2];
    x.ingest = v[3];
    x.egress = v[4];
    return 0;
}

}  // namespace

extern "C" {

struct fmvd_plan_cuda_16* fmvd_initialize_cuda_16(int dataH, int dataW, float const* const* h_Weights,
                                                  float** h_Kernel, int kernelH, int kernelW,
                                                  fmvd_psf_type iteration_type, int nViews,
                                                  int nstreams, datasource_t_16 get_next_plane,
                                                  datasink_t_16 return_next_plane)
{
    return initialize<fmvd_plan_cuda_16, uint16_t>(dataH, dataW, h_Weights, h_Kernel, kernelH,
                                                   kernelW, iteration_type, nViews, nstreams,
                                                   get_next_plane, return_next_plane);
}

struct fmvd_plan_cuda_8* fmvd_initialize_cuda_8(int dataH, int dataW, float const* const* h_Weights,
                                                float** h_Kernel, int kernelH, int kernelW,
                                                fmvd_psf_type iteration_type, int nViews,
                                                int nstreams, datasource_t_8 get_next_plane,
                                                datasink_t_8 return_next_plane)
{
    return initialize<fmvd_plan_cuda_8, uint8_t>(dataH, dataW, h_Weights, h_Kernel, kernelH,
                                                 kernelW, iteration_type, nViews, nstreams,
                                                 get_next_plane, return_next_plane);
}

void fmvd_deconvolve_planes_cuda_16(const struct fmvd_plan_cuda_16* plan, int iterations)
{
    FMVD_API_GUARD(, {
        if (!plan) fmvd::throw_msg("NULL plan");
        const_cast<fmvd_plan_cuda_16*>(plan)->impl.deconvolve_planes(iterations);
    })
}

void fmvd_deconvolve_planes_cuda_8(const struct fmvd_plan_cuda_8* plan, int iterations)
{
    FMVD_API_GUARD(, {
        if (!plan) fmvd::throw_msg("NULL plan");
        const_cast<fmvd_plan_cuda_8*>(plan)->impl.deconvolve_planes(iterations);
    })
}

void fmvd_destroy_cuda_16(struct fmvd_plan_cuda_16* plan)
{
    if (plan) (void)hipDeviceSynchronize();
    delete plan;
}

void fmvd_destroy_cuda_8(struct fmvd_plan_cuda_8* plan)
{
    if (plan) (void)hipDeviceSynchronize();
    delete plan;
}

int fmvd_set_estimate_sink_cuda_16(struct fmvd_plan_cuda_16* plan, fmvd_estimate_sink_t sink, void* user)
{
    if (!plan) return -1;
    plan->impl.esink = sink;
    plan->impl.euser = user;
    return 0;
}

int fmvd_set_estimate_sink_cuda_8(struct fmvd_plan_cuda_8* plan, fmvd_estimate_sink_t sink, void* user)
{
    if (!plan) return -1;
    plan->impl.esink = sink;
    plan->impl.euser = user;
    return 0;
}

int fmvd_set_batch_planes_cuda_16(struct fmvd_plan_cuda_16* plan, int planes)
{
    if (!plan) return -1;
    plan->impl.batchCap = planes > 0 ? planes : 0;
    return plan->impl.batch_planes();
}

int fmvd_set_batch_planes_cuda_8(struct fmvd_plan_cuda_8* plan, int planes)
{
    if (!plan) return -1;
    plan->impl.batchCap = planes > 0 ? planes : 0;
    return plan->impl.batch_planes();
}

int fmvd_set_chunk_planes_cuda_16(struct fmvd_plan_cuda_16* plan, int planes)
{
    if (!plan) return -1;
    plan->impl.chunkCap = planes > 0 ? planes : 0;
    return plan->impl.chunk_planes();
}

int fmvd_set_chunk_planes_cuda_8(struct fmvd_plan_cuda_8* plan, int planes)
{
    if (!plan) return -1;
    plan->impl.chunkCap = planes > 0 ? planes : 0;
    return plan->impl.chunk_planes();
}

int fmvd_set_lookahead_cuda_16(struct fmvd_plan_cuda_16* plan, int planes)
{
    if (!plan) return -1;
    plan->impl.lookahead = planes > 0 ? planes : 0;
    return plan->impl.lookahead;
}

int fmvd_set_lookahead_cuda_8(struct fmvd_plan_cuda_8* plan, int planes)
{
    if (!plan) return -1;
    plan->impl.lookahead = planes > 0 ? planes : 0;
    return plan->impl.lookahead;
}

int fmvd_set_weights_device_cuda_16(struct fmvd_plan_cuda_16* plan, const float* const* d_weights, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        plan->impl.set_weights_device(d_weights, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_set_weights_device_cuda_8(struct fmvd_plan_cuda_8* plan, const float* const* d_weights, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        plan->impl.set_weights_device(d_weights, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_reconstruct_device_cuda_16(struct fmvd_plan_cuda_16* plan, const unsigned short* const* d_acquired, int w,
                                    int h, int d, const float* const* h_inverse, int border, float zspacing,
                                    unsigned short* const* d_registered, int nPlanes, int iterations,
                                    unsigned short* d_out, float* d_estimate, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        plan->impl.reconstruct_device(d_acquired, w, h, d, h_inverse, border, zspacing, d_registered, nPlanes,
                                      iterations, d_out, d_estimate, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_reconstruct_device_cuda_8(struct fmvd_plan_cuda_8* plan, const unsigned char* const* d_acquired, int w,
                                   int h, int d, const float* const* h_inverse, int border, float zspacing,
                                   unsigned char* const* d_registered, int nPlanes, int iterations,
                                   unsigned char* d_out, float* d_estimate, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        plan->impl.reconstruct_device(d_acquired, w, h, d, h_inverse, border, zspacing, d_registered, nPlanes,
                                      iterations, d_out, d_estimate, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_deconvolve_volume_device_cuda_16(const struct fmvd_plan_cuda_16* plan,
                                          const unsigned short* const* d_views, int nPlanes,
                                          int iterations, const float* d_init, unsigned short* d_out,
                                          float* d_estimate, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        const_cast<fmvd_plan_cuda_16*>(plan)->impl.run_volume(d_views, nPlanes, iterations, d_init, d_out,
                                                              d_estimate, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_deconvolve_volume_device_noalloc_cuda_16(const struct fmvd_plan_cuda_16* plan,
                                                  const unsigned short* const* d_views, int nPlanes, int iterations,
                                                  const float* d_init, unsigned short* d_out, float* d_estimate,
                                                  void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        const_cast<fmvd_plan_cuda_16*>(plan)->impl.run_volume_noalloc(d_views, nPlanes, iterations, d_init, d_out,
                                                                      d_estimate, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_deconvolve_volume_device_noalloc_cuda_8(const struct fmvd_plan_cuda_8* plan,
                                                 const unsigned char* const* d_views, int nPlanes, int iterations,
                                                 const float* d_init, unsigned char* d_out, float* d_estimate,
                                                 void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        const_cast<fmvd_plan_cuda_8*>(plan)->impl.run_volume_noalloc(d_views, nPlanes, iterations, d_init, d_out,
                                                                     d_estimate, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_convergence_cuda_16(const struct fmvd_plan_cuda_16* plan, const unsigned short* const* d_views,
                             int nPlanes, int iterations, const float* d_init, const float* d_ref, double* errs,
                             void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        const_cast<fmvd_plan_cuda_16*>(plan)->impl.convergence(d_views, nPlanes, iterations, d_init, d_ref, errs,
                                                               (hipStream_t)stream);
        return 0;
    })
}

int fmvd_convergence_cuda_8(const struct fmvd_plan_cuda_8* plan, const unsigned char* const* d_views, int nPlanes,
                            int iterations, const float* d_init, const float* d_ref, double* errs, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        const_cast<fmvd_plan_cuda_8*>(plan)->impl.convergence(d_views, nPlanes, iterations, d_init, d_ref, errs,
                                                              (hipStream_t)stream);
        return 0;
    })
}

int fmvd_deconvolve_volume_device_cuda_8(const struct fmvd_plan_cuda_8* plan,
                                         const unsigned char* const* d_views, int nPlanes,
                                         int iterations, const float* d_init, unsigned char* d_out,
                                         float* d_estimate, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!plan) fmvd::throw_msg("NULL plan");
        const_cast<fmvd_plan_cuda_8*>(plan)->impl.run_volume(d_views, nPlanes, iterations, d_init, d_out,
                                                             d_estimate, (hipStream_t)stream);
        return 0;
    })
}

int fmvd_autotune_cuda_16(struct fmvd_plan_cuda_16* plan, int planes, int* chosen, float* times_ms)
{
    return autotune_plan(plan, planes, chosen, times_ms);
}
int fmvd_autotune_cuda_8(struct fmvd_plan_cuda_8* plan, int planes, int* chosen, float* times_ms)
{
    return autotune_plan(plan, planes, chosen, times_ms);
}
int fmvd_set_variants_cuda_16(struct fmvd_plan_cuda_16* plan, const int* variants) { return set_variants(plan, variants); }
int fmvd_set_variants_cuda_8(struct fmvd_plan_cuda_8* plan, const int* variants) { return set_variants(plan, variants); }

int fmvd_plan_info_cuda_16(const struct fmvd_plan_cuda_16* plan, int* dims) { return plan_info(plan, dims); }
int fmvd_plan_info_cuda_8(const struct fmvd_plan_cuda_8* plan, int* dims) { return plan_info(plan, dims); }

int fmvd_snap_transform_size(int n) { return fmvd::snap_transform_size(n); }

int fmvd_fft_extent_supported(int fftH, int fftW)
{
    return fftW >= 16 && fftH >= 16 && (fftW % 2 == 0) && fmvd::row_size_supported(fftW / 2) &&
           fmvd::col_size_supported(fftH);
}

}  // extern "C"
