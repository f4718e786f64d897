// Device-side description of a batch of independent RL volumes and the
// launchers of the batched passes (rl_batch_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "rl_launch.hpp"

namespace fmvd {

constexpr int kBatchMaxViews = 8;

// One volume of an extent class (an entry of the device segment table).
struct BatchSeg {
    int dataH, dataW;
    int dkY, dkX;            // data pad half-widths (swapped, as the reference)
    int wkY, wkX;            // weight pad half-widths
    int nplanes, plane0;     // planes, first plane in the class buffers
    int nViews;
    float e0;                // constant start 10 * nViews
    long dstride;            // dataH * dataW: elements between planes of data/out/est
    const void* data[kBatchMaxViews];   // views, plane-major SAMPLE volumes
    const float* w[kBatchMaxViews];     // normalised weights, dataH x dataW
    const float2* K[kBatchMaxViews];    // kernel spectra (transposed layout, pass A scaling)
    const float2* Kh[kBatchMaxViews];   // kernel-hat spectra (pass C scaling), or K when khconj
    int khconj;              // pass C multiplies by conj(Kh): the hat is the flipped kernel
    float rscale;            // ratio-pass output scale: 1, or 1/fftW when khconj
    void* out;               // SAMPLE result volume or null
    float* est;              // float estimate volume or null
};

bool batch_extent_supported(int fftH, int fftW);
int batch_row_rows_per_block(int N, int variant);
int batch_col_seqs_per_block(int H, int variant);

// list: ids[n] then block starts[n + 1] (see rl_batch_kernels.hip).
template <typename SAMPLE>
void launch_batch_ratio(const RowGeom& g, const BatchSeg* segs, const int* list, int n, int nblocks, int view,
                        float2* XT, long xstride, int variant, hipStream_t s);
void launch_batch_update(const RowGeom& g, const BatchSeg* segs, const int* list, int n, int nblocks, int view,
                         float2* XE, const float2* XT, float* ES, long xstride, long estride, int variant,
                         hipStream_t s);
void launch_batch_col(const ColGeom& g, const BatchSeg* segs, const int* list, int n, int nblocks, int view, int hat,
                      const float2* Xin, float2* Xout, long xstride, int variant, hipStream_t s);
// Plane lists: ids[n] then plane starts[n + 1].
void launch_batch_const_init(const BatchSeg* segs, const int* list, int n, int nplanes, float2* XE, float* ES,
                             long xstride, long estride, int N, int H, hipStream_t s);
template <typename SAMPLE>
void launch_batch_final(const BatchSeg* segs, const int* list, int n, int nplanes, int maxH, int maxW,
                        const float* ES, long estride, int W, hipStream_t s);

}  // namespace fmvd
