// Batched plane-wise multi-view RL over many independent volumes of uneven
// size, with one RL iteration captured as a hipGraph (C ABI of fmvd_batch.h).
//
// No reference counterpart: the reference reconstructs one volume per plan
// (fmvd_deconvolve.impl.cpp:130-300).  Each volume here is prepared exactly
// as a single-volume plan prepares its operands (PSF preparation, kernel
// spectra, normalised weights), and each pass runs the single-volume tile code,
// so a batched volume's result is bit-identical to running it alone with the
// same FFT plan variants.
//
// Volumes are grouped by padded FFT extent into classes.  A class's planes
// are split into cache-resident chunks (as the single-volume plan does): a
// chunk holds whole volumes or plane ranges of one (segments of the device
// table), sized so that one iteration's working set fits a 200 MiB budget in
// the memory-side cache.  Per iteration of a chunk the device runs, for each
// view index v, the four passes (column, ratio, column, update) over every
// plane of every chunk segment with more than v views: one launch per pass
// and view, captured into one graph per chunk that is replayed once per
// iteration before the next chunk starts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "../../include/fmvd_batch.h"
#include "../../include/fmvd_cuda_utils.h"
#include "hip_check.hpp"
#include "host_buffers.hpp"
#include "host_psf.hpp"
#include "rl_batch.hpp"
#include "rl_launch.hpp"

namespace fmvd {

template <typename S> struct Batch {
    struct Vol {
        int dataH = 0, dataW = 0, nplanes = 0, nViews = 0, kH = 0, kW = 0;
        int cls = 0, plane0 = 0;
        bool khconj = false;  // kernel-hat spectrum = conj(K) (odd independent PSFs): Kh not built
        float rscale = 1.0f;
        std::vector<DeviceBuffer> K, Kh, Wn;
        const S* views[kBatchMaxViews] = {};
        S* out = nullptr;
        float* est = nullptr;
    };
    struct Launch {
        int off = 0, n = 0, blocks = 0;  // list offset (ints), entries, total blocks
    };
    struct Cls {
        int fftH = 0, fftW = 0, N = 0, nplanes = 0, maxV = 0, maxH = 0, maxW = 0;
        std::vector<int> vols;
        std::vector<int> segs;   // device segment table entries of the class
        std::vector<int> chunks;
        DeviceBuffer XE, ES, twW, twRow[3], twCol[3];  // XE: the in-place spectrum of every pass
        Variants var;
        Launch planes;
        long xs() const { return (long)N * fftH; }
        long es() const { return (long)fftW * fftH; }
    };
    struct Chunk {
        int cls = 0;
        std::vector<int> segs;
        std::vector<Launch> col, ratio, update;  // per view index
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };

    int device = 0;
    std::vector<Vol> vols;
    std::vector<Cls> cls;
    std::vector<BatchSeg> table;  // host copy of the device segment table
    std::vector<Chunk> chunks;
    DeviceBuffer segs, lists;
    std::unique_ptr<Stream> cap;

    ~Batch() { drop_graph(); }

    void drop_graph()
    {
        for (Chunk& ch : chunks) {
            if (ch.exec) (void)hipGraphExecDestroy(ch.exec);
            if (ch.graph) (void)hipGraphDestroy(ch.graph);
            ch.exec = nullptr;
            ch.graph = nullptr;
        }
    }

    static bool explicit_khat()
    {
        const char* e = getenv("FMVD_EXPLICIT_KHAT");
        return e && *e && strcmp(e, "0") != 0;
    }

    static size_t chunk_budget()
    {
        const char* e = getenv("FMVD_CHUNK_BYTES");
        return (e && *e) ? (size_t)strtoull(e, nullptr, 10) : ((size_t)200 << 20);
    }

    // The segment-table entry of a whole volume (init narrows it to planes).
    BatchSeg volume_seg(const Vol& v) const
    {
        BatchSeg s{};
        s.dataH = v.dataH;
        s.dataW = v.dataW;
        // data pad half-widths swapped, as the reference
        // (fmvd_deconvolve.impl.cpp:36-37, fmvd_deconvolve_cuda.impl.cu:76-77)
        s.dkY = v.kW / 2;
        s.dkX = v.kH / 2;
        s.wkY = v.kH / 2;
        s.wkX = v.kW / 2;
        s.nplanes = v.nplanes;
        s.plane0 = v.plane0;
        s.nViews = v.nViews;
        s.e0 = 10.0f * (float)v.nViews;
        s.dstride = (long)v.dataH * v.dataW;
        for (int k = 0; k < v.nViews; ++k) {
            s.data[k] = v.views[k];
            s.w[k] = v.Wn[k].template as<float>();
            s.K[k] = v.K[k].template as<float2>();
            s.Kh[k] = (v.khconj ? v.K[k] : v.Kh[k]).template as<float2>();
        }
        s.khconj = v.khconj ? 1 : 0;
        s.rscale = v.rscale;
        s.out = v.out;
        s.est = v.est;
        return s;
    }

    void init(int n, const fmvd_batch_volume* d)
    {
        if (n < 1 || !d) throw_msg("fmvd_batch_create: need at least one volume");
        for (int i = 0; i < n; ++i) {  // validate everything before touching the device
            const fmvd_batch_volume& a = d[i];
            if (a.dataH < 1 || a.dataW < 1 || a.nplanes < 1 || a.kernelH < 1 || a.kernelW < 1 || a.nViews < 1)
                throw_msg("fmvd_batch_create: invalid volume dimensions");
            if (a.nViews > kBatchMaxViews) throw_msg("fmvd_batch_create: at most 8 views per volume");
            if ((int)a.psf_type < 0 || (int)a.psf_type > 3) throw_msg("fmvd_batch_create: invalid PSF type");
            if (!a.h_weights || !a.h_kernels || !a.d_views) throw_msg("fmvd_batch_create: null operand arrays");
            for (int k = 0; k < a.nViews; ++k)
                if (!a.d_views[k] || !a.h_weights[k] || !a.h_kernels[k])
                    throw_msg("fmvd_batch_create: null view, weight or kernel pointer");
            const int fftH = snap_transform_size(a.dataH + a.kernelH - 1);
            const int fftW = snap_transform_size(a.dataW + a.kernelW - 1);
            if (!batch_extent_supported(fftH, fftW)) {
                char msg[160];
                snprintf(msg, sizeof msg, "fmvd_batch_create: unsupported FFT extent %d x %d for volume %d", fftH,
                         fftW, i);
                throw_msg(msg);
            }
        }
        FMVD_HIP(hipGetDevice(&device));
        vols.resize(n);
        std::map<std::pair<int, int>, int> classOf;
        for (int i = 0; i < n; ++i) {
            const fmvd_batch_volume& a = d[i];
            Vol& v = vols[i];
            v.dataH = a.dataH;
            v.dataW = a.dataW;
            v.nplanes = a.nplanes;
            v.nViews = a.nViews;
            v.kH = a.kernelH;
            v.kW = a.kernelW;
            for (int k = 0; k < a.nViews; ++k) v.views[k] = static_cast<const S*>(a.d_views[k]);
            v.out = static_cast<S*>(a.d_out);
            v.est = a.d_est;
            const int fftH = snap_transform_size(a.dataH + a.kernelH - 1);
            const int fftW = snap_transform_size(a.dataW + a.kernelW - 1);
            auto key = std::make_pair(fftH, fftW);
            if (!classOf.count(key)) {
                classOf[key] = (int)cls.size();
                Cls c;
                c.fftH = fftH;
                c.fftW = fftW;
                c.N = fftW / 2;
                cls.push_back(std::move(c));
            }
            v.cls = classOf[key];
            Cls& c = cls[v.cls];
            v.plane0 = c.nplanes;
            c.nplanes += v.nplanes;
            c.maxV = std::max(c.maxV, v.nViews);
            c.maxH = std::max(c.maxH, v.dataH);
            c.maxW = std::max(c.maxW, v.dataW);
            c.vols.push_back(i);

            // operands exactly as Plan::init prepares them
            // (fmvd_deconvolve.impl.cpp:61-64, 86-124)
            const int len = a.kernelH * a.kernelW;
            std::vector<double> khat = compute_kernel_hat(a.h_kernels, a.kernelW, a.kernelH, a.nViews, a.psf_type);
            for (int k = 0; k < a.nViews; ++k) normalize_kernel_f32(a.h_kernels[k], len);
            v.K.resize(a.nViews);
            v.Kh.resize(a.nViews);
            v.Wn.resize(a.nViews);
            // as Plan::init: an odd independent PSF's hat spectrum is conj(K)
            v.khconj = a.psf_type == 0 && a.kernelW % 2 == 1 && a.kernelH % 2 == 1 && !explicit_khat();
            v.rscale = v.khconj ? 1.0f / (float)fftW : 1.0f;
            std::vector<double> kv(len);
            for (int k = 0; k < a.nViews; ++k) {
                for (int j = 0; j < len; ++j) kv[j] = (double)a.h_kernels[k][j];
                upload(v.K[k], kernel_spectrum_transposed(kv.data(), a.kernelW, a.kernelH, fftH, fftW, 1.0 / fftH));
                if (!v.khconj)
                    upload(v.Kh[k], kernel_spectrum_transposed(khat.data() + (size_t)k * len, a.kernelW, a.kernelH,
                                                               fftH, fftW, 1.0 / ((double)fftH * fftW)));
            }
            const size_t plane = (size_t)a.dataH * a.dataW;
            std::vector<float> wn = normalized_weights(a.h_weights, a.dataW, a.dataH, a.nViews);
            for (int k = 0; k < a.nViews; ++k)
                upload(v.Wn[k], std::vector<float>(wn.begin() + (size_t)k * plane, wn.begin() + (size_t)(k + 1) * plane));
        }
        for (Cls& c : cls) {
            upload(c.twW, twiddles(c.fftW, c.N));
            for (int k = 0; k < 3; ++k) {
                upload(c.twRow[k], row_pass_twiddles(c.N, k));
                upload(c.twCol[k], col_pass_twiddles(c.fftH, k));
            }
            c.XE.alloc(sizeof(float2) * (size_t)c.xs() * c.nplanes);
            c.ES.alloc(sizeof(float) * (size_t)c.es() * c.nplanes);
        }
        // device segment table: every volume split at chunk boundaries
        const size_t budget = chunk_budget();
        for (size_t ci = 0; ci < cls.size(); ++ci) {
            Cls& c = cls[ci];
            size_t used = 0;
            auto open_chunk = [&] {
                Chunk ch;
                ch.cls = (int)ci;
                c.chunks.push_back((int)chunks.size());
                chunks.push_back(std::move(ch));
                used = 0;
            };
            open_chunk();
            for (int i : c.vols) {
                const Vol& v = vols[i];
                const size_t plane = (size_t)v.dataH * v.dataW;
                // per plane: spectrum, estimate (the samples are streamed with
                // non-temporal loads); per volume: every view's kernel spectra
                // (one or two) and weights
                const size_t pp = (size_t)c.xs() * sizeof(float2) + (size_t)c.es() * sizeof(float);
                const size_t nsp = v.khconj ? 1 : 2;
                const size_t pv = (size_t)v.nViews * (nsp * (size_t)c.xs() * sizeof(float2) + plane * sizeof(float));
                int first = 0;
                while (first < v.nplanes) {
                    const size_t need = pv + pp;
                    if (used + need > budget && !chunks[c.chunks.back()].segs.empty()) open_chunk();
                    const size_t room = budget > used + pv ? budget - used - pv : 0;
                    const int take = std::max(1, std::min(v.nplanes - first, (int)std::min<size_t>(room / pp, 1 << 20)));
                    BatchSeg sg = volume_seg(v);
                    sg.nplanes = take;
                    sg.plane0 = v.plane0 + first;
                    for (int k = 0; k < v.nViews; ++k)
                        sg.data[k] = static_cast<const S*>(v.views[k]) + (size_t)first * plane;
                    sg.out = v.out ? static_cast<void*>(v.out + (size_t)first * plane) : nullptr;
                    sg.est = v.est ? v.est + (size_t)first * plane : nullptr;
                    c.segs.push_back((int)table.size());
                    chunks[c.chunks.back()].segs.push_back((int)table.size());
                    table.push_back(sg);
                    used += pv + (size_t)take * pp;
                    first += take;
                }
            }
        }
        upload(segs, table);
        build_lists();
        cap.reset(new Stream());
    }

    // Launch lists for the current plan variants (block counts depend on them).
    void build_lists()
    {
        std::vector<int> L;
        auto add = [&](const std::vector<int>& ids_in, int minViews, auto blocksOf) {
            Launch l;
            l.off = (int)L.size();
            std::vector<int> ids, starts;
            int total = 0;
            for (int i : ids_in)
                if (table[i].nViews > minViews) {
                    ids.push_back(i);
                    starts.push_back(total);
                    total += blocksOf(table[i]);
                }
            starts.push_back(total);
            l.n = (int)ids.size();
            l.blocks = total;
            L.insert(L.end(), ids.begin(), ids.end());
            L.insert(L.end(), starts.begin(), starts.end());
            return l;
        };
        for (Cls& c : cls) {
            c.planes = add(c.segs, 0, [](const BatchSeg& g) { return g.nplanes; });
            const int Gr = batch_row_rows_per_block(c.N, c.var.ratio);
            const int Gu = batch_row_rows_per_block(c.N, c.var.update);
            const int SQ = batch_col_seqs_per_block(c.fftH, c.var.column);
            if (c.fftH % Gr || c.fftH % Gu || c.N % SQ) throw_msg("batched RL: tile shape does not divide the extent");
            for (int chi : c.chunks) {
                Chunk& ch = chunks[chi];
                ch.col.clear();
                ch.ratio.clear();
                ch.update.clear();
                for (int k = 0; k < c.maxV; ++k) {
                    ch.col.push_back(add(ch.segs, k, [&](const BatchSeg& g) { return g.nplanes * (c.N / SQ); }));
                    ch.ratio.push_back(add(ch.segs, k, [&](const BatchSeg& g) { return g.nplanes * (c.fftH / Gr); }));
                    ch.update.push_back(add(ch.segs, k, [&](const BatchSeg& g) { return g.nplanes * (c.fftH / Gu); }));
                }
            }
        }
        // the previous lists may still be read by queued launches
        FMVD_HIP(hipDeviceSynchronize());
        lists.release();
        upload(lists, L);
        drop_graph();
    }

    RowGeom row_geom(const Cls& c) const
    {
        RowGeom g{};
        g.H = c.fftH;
        g.W = c.fftW;
        g.N = c.N;
        for (int k = 0; k < 3; ++k) g.twP[k] = c.twRow[k].template as<float2>();
        g.twW = c.twW.template as<float2>();
        return g;
    }
    ColGeom col_geom(const Cls& c) const
    {
        ColGeom g{};
        g.H = c.fftH;
        g.N = c.N;
        for (int k = 0; k < 3; ++k) g.twP[k] = c.twCol[k].template as<float2>();
        return g;
    }

    const int* L(const Launch& l) const { return lists.as<int>() + l.off; }

    void enqueue_init(hipStream_t s)
    {
        const BatchSeg* sg = segs.as<BatchSeg>();
        for (Cls& c : cls)
            launch_batch_const_init(sg, L(c.planes), c.planes.n, c.nplanes, c.XE.template as<float2>(),
                                    c.ES.template as<float>(), c.xs(), c.es(), c.N, c.fftH, s);
    }
    void enqueue_iteration(const Chunk& ch, hipStream_t s)
    {
        const BatchSeg* sg = segs.as<BatchSeg>();
        Cls& c = cls[ch.cls];
        const RowGeom rg = row_geom(c);
        const ColGeom cg = col_geom(c);
        float2* xe = c.XE.template as<float2>();
        float2* xt = xe;
        float* es = c.ES.template as<float>();
        for (int v = 0; v < c.maxV; ++v) {
            if (ch.col[v].n == 0) continue;
            launch_batch_col(cg, sg, L(ch.col[v]), ch.col[v].n, ch.col[v].blocks, v, 0, xe, xt, c.xs(), c.var.column,
                             s);
            launch_batch_ratio<S>(rg, sg, L(ch.ratio[v]), ch.ratio[v].n, ch.ratio[v].blocks, v, xt, c.xs(),
                                  c.var.ratio, s);
            launch_batch_col(cg, sg, L(ch.col[v]), ch.col[v].n, ch.col[v].blocks, v, 1, xt, xt, c.xs(), c.var.column,
                             s);
            launch_batch_update(rg, sg, L(ch.update[v]), ch.update[v].n, ch.update[v].blocks, v, xe, xt, es, c.xs(),
                                c.es(), c.var.update, s);
        }
    }
    void enqueue_final(hipStream_t s)
    {
        const BatchSeg* sg = segs.as<BatchSeg>();
        for (Cls& c : cls)
            launch_batch_final<S>(sg, L(c.planes), c.planes.n, c.nplanes, c.maxH, c.maxW, c.ES.template as<float>(),
                                  c.es(), c.fftW, s);
    }

    void build_graph(Chunk& ch)
    {
        hipStream_t s = cap->s;
        FMVD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_iteration(ch, s);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        FMVD_HIP(hipStreamEndCapture(s, &ch.graph));
        FMVD_HIP(hipGraphInstantiate(&ch.exec, ch.graph, nullptr, nullptr, 0));
    }

    void run(int iterations, int use_graph, hipStream_t s)
    {
        if (iterations < 0) throw_msg("fmvd_batch_run: negative iteration count");
        FMVD_HIP(hipSetDevice(device));
        enqueue_init(s);
        for (Chunk& ch : chunks) {
            if (use_graph && iterations > 0) {
                if (!ch.exec) build_graph(ch);
                for (int it = 0; it < iterations; ++it) FMVD_HIP(hipGraphLaunch(ch.exec, s));
            } else {
                for (int it = 0; it < iterations; ++it) enqueue_iteration(ch, s);
            }
        }
        enqueue_final(s);
    }

    // Times each compiled plan variant of the column, ratio and update
    // passes on every class's own buffers (view 0, all planes) and keeps the
    // fastest.  chosen[3*class + kind] receives the pick (kind 0 column,
    // 1 ratio, 2 update).
    void autotune(int* chosen)
    {
        FMVD_HIP(hipSetDevice(device));
        hipStream_t s = cap->s;
        hipEvent_t a, b;
        FMVD_HIP(hipEventCreate(&a));
        FMVD_HIP(hipEventCreate(&b));
        enqueue_init(s);
        const BatchSeg* sg = segs.as<BatchSeg>();
        for (size_t ci = 0; ci < cls.size(); ++ci) {
            Cls& c = cls[ci];
            int* slots[3] = {&c.var.column, &c.var.ratio, &c.var.update};
            const int nv[3] = {col_variants(c.fftH), row_variants(c.N), row_variants(c.N)};
            for (int kind = 0; kind < 3; ++kind) {
                float best = 0.f;
                int bestv = 0;
                for (int var = 0; var < nv[kind]; ++var) {
                    *slots[kind] = var;
                    build_lists();
                    // the class's first chunk, at the size the run launches
                    const Chunk& ch = chunks[c.chunks.front()];
                    auto launch = [&] {
                        const RowGeom rg = row_geom(c);
                        const ColGeom cg = col_geom(c);
                        float2* xe = c.XE.template as<float2>();
                        float2* xt = xe;
                        float* es = c.ES.template as<float>();
                        if (kind == 0)
                            launch_batch_col(cg, sg, L(ch.col[0]), ch.col[0].n, ch.col[0].blocks, 0, 0, xe, xt,
                                             c.xs(), var, s);
                        else if (kind == 1)
                            launch_batch_ratio<S>(rg, sg, L(ch.ratio[0]), ch.ratio[0].n, ch.ratio[0].blocks, 0, xt,
                                                  c.xs(), var, s);
                        else
                            launch_batch_update(rg, sg, L(ch.update[0]), ch.update[0].n, ch.update[0].blocks, 0, xe,
                                                xt, es, c.xs(), c.es(), var, s);
                    };
                    launch();
                    FMVD_HIP(hipEventRecord(a, s));
                    for (int r = 0; r < 3; ++r) launch();
                    FMVD_HIP(hipEventRecord(b, s));
                    FMVD_HIP(hipEventSynchronize(b));
                    float ms = 0.f;
                    FMVD_HIP(hipEventElapsedTime(&ms, a, b));
                    if (var == 0 || ms < best) {
                        best = ms;
                        bestv = var;
                    }
                }
                *slots[kind] = bestv;
                if (chosen) chosen[3 * ci + kind] = bestv;
            }
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        build_lists();
        FMVD_HIP(hipStreamSynchronize(s));
    }
};

}  // namespace fmvd

struct fmvd_batch_cuda_16 {
    fmvd::Batch<uint16_t> impl;
};
struct fmvd_batch_cuda_8 {
    fmvd::Batch<uint8_t> impl;
};

namespace {
template <typename B> B* create(int n, const fmvd_batch_volume* v)
{
    B* b = new B();
    try {
        b->impl.init(n, v);
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}
template <typename B> int info(const B* b, int* nclasses, int* extents, int* planes)
{
    if (!b) fmvd::throw_msg("fmvd_batch_info: null batch");
    if (nclasses) *nclasses = (int)b->impl.cls.size();
    for (size_t i = 0; i < b->impl.cls.size(); ++i) {
        if (extents) {
            extents[2 * i] = b->impl.cls[i].fftH;
            extents[2 * i + 1] = b->impl.cls[i].fftW;
        }
        if (planes) planes[i] = b->impl.cls[i].nplanes;
    }
    return 0;
}
}  // namespace

extern "C" {

struct fmvd_batch_cuda_16* fmvd_batch_create_cuda_16(int nvolumes, const struct fmvd_batch_volume* volumes)
{
    FMVD_API_GUARD(nullptr, { return create<fmvd_batch_cuda_16>(nvolumes, volumes); })
}
struct fmvd_batch_cuda_8* fmvd_batch_create_cuda_8(int nvolumes, const struct fmvd_batch_volume* volumes)
{
    FMVD_API_GUARD(nullptr, { return create<fmvd_batch_cuda_8>(nvolumes, volumes); })
}
int fmvd_batch_run_cuda_16(struct fmvd_batch_cuda_16* b, int iterations, int use_graph, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!b) fmvd::throw_msg("fmvd_batch_run: null batch");
        b->impl.run(iterations, use_graph, (hipStream_t)stream);
        return 0;
    })
}
int fmvd_batch_run_cuda_8(struct fmvd_batch_cuda_8* b, int iterations, int use_graph, void* stream)
{
    FMVD_API_GUARD(-1, {
        if (!b) fmvd::throw_msg("fmvd_batch_run: null batch");
        b->impl.run(iterations, use_graph, (hipStream_t)stream);
        return 0;
    })
}
int fmvd_batch_autotune_cuda_16(struct fmvd_batch_cuda_16* b, int* chosen)
{
    FMVD_API_GUARD(-1, {
        if (!b) fmvd::throw_msg("fmvd_batch_autotune: null batch");
        b->impl.autotune(chosen);
        return 0;
    })
}
int fmvd_batch_autotune_cuda_8(struct fmvd_batch_cuda_8* b, int* chosen)
{
    FMVD_API_GUARD(-1, {
        if (!b) fmvd::throw_msg("fmvd_batch_autotune: null batch");
        b->impl.autotune(chosen);
        return 0;
    })
}
int fmvd_batch_info_cuda_16(const struct fmvd_batch_cuda_16* b, int* nclasses, int* extents, int* planes)
{
    FMVD_API_GUARD(-1, { return info(b, nclasses, extents, planes); })
}
int fmvd_batch_info_cuda_8(const struct fmvd_batch_cuda_8* b, int* nclasses, int* extents, int* planes)
{
    FMVD_API_GUARD(-1, { return info(b, nclasses, extents, planes); })
}
void fmvd_batch_destroy_cuda_16(struct fmvd_batch_cuda_16* b) { delete b; }
void fmvd_batch_destroy_cuda_8(struct fmvd_batch_cuda_8* b) { delete b; }

}  // extern "C"
