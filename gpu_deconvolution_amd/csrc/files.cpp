// File front end: fmvd_deconvolve_files_cuda_{16,8}.
// Reference: fmvd_deconvolve_hdd.impl.cpp:16-76.  Raw planes (dataW*dataH
// samples, host endian, no header) are pulled from one file per view and the
// results appended to resultFile in z order.  The datasource/datasink
// callbacks of the plan API carry no user pointer, so - as in the reference -
// the file state is process-global and the function is not reentrant.
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/fmvd_deconvolve.h"
#include "hip_check.hpp"

namespace {

struct IoState {
    FILE** dataFiles = nullptr;
    FILE* resultFile = nullptr;
    size_t planeSamples = 0;
    int plane = 0, nPlanes = 0, nViews = 0;
    std::string error;
};

IoState* g_io = nullptr;

template <typename S> int get_next_plane(S** data, int offset)
{
    IoState* io = g_io;
    if (!io || io->plane >= io->nPlanes || !io->error.empty()) return 0;
    for (int v = 0; v < io->nViews; ++v) {
        const size_t got = fread(data[v] + offset, sizeof(S), io->planeSamples, io->dataFiles[v]);
        if (got != io->planeSamples) {
            io->error = "short read from view file " + std::to_string(v) + " at plane " +
                        std::to_string(io->plane);
            return 0;
        }
    }
    ++io->plane;
    return 1;
}

template <typename S> void return_next_plane(S* data)
{
    IoState* io = g_io;
    if (!io || !io->error.empty()) return;
    if (fwrite(data, sizeof(S), io->planeSamples, io->resultFile) != io->planeSamples)
        io->error = "short write to result file";
}

template <typename S, typename PlanT, typename InitF, typename RunF, typename DestroyF>
void run_files(FILE** dataFiles, FILE* resultFile, int dataW, int dataH, int dataD, float** h_Weights,
               float** h_Kernel, int kernelH, int kernelW, fmvd_psf_type t, int nViews, int iterations,
               InitF init, RunF run, DestroyF destroy)
{
    if (!dataFiles || !resultFile) fmvd::throw_msg("files front end: NULL file handle");
    for (int v = 0; v < nViews; ++v)
        if (!dataFiles[v]) fmvd::throw_msg("files front end: NULL view file");
    if (dataD < 0) fmvd::throw_msg("files front end: negative plane count");
    IoState io;
    io.dataFiles = dataFiles;
    io.resultFile = resultFile;
    io.planeSamples = (size_t)dataW * dataH;
    io.nPlanes = dataD;
    io.nViews = nViews;
    g_io = &io;
    // fmvd_deconvolve_hdd.impl.cpp:51: three planes in flight
    PlanT* plan = init(dataH, dataW, h_Weights, h_Kernel, kernelH, kernelW, t, nViews, 3,
                       get_next_plane<S>, return_next_plane<S>);
    if (!plan) {
        g_io = nullptr;
        return;  // already reported
    }
    run(plan, iterations);
    destroy(plan);
    g_io = nullptr;
    if (!io.error.empty()) fmvd::throw_msg(io.error);
}

}  // namespace

extern "C" {

void fmvd_deconvolve_files_cuda_16(FILE** dataFiles, FILE* resultFile, int dataW, int dataH, int dataD,
                                   float** h_Weights, float** h_Kernel, int kernelH, int kernelW,
                                   fmvd_psf_type iteration_type, int nViews, int iterations)
{
    FMVD_API_GUARD(, {
        run_files<uint16_t, fmvd_plan_cuda_16>(dataFiles, resultFile, dataW, dataH, dataD, h_Weights,
                                               h_Kernel, kernelH, kernelW, iteration_type, nViews,
                                               iterations, fmvd_initialize_cuda_16,
                                               fmvd_deconvolve_planes_cuda_16, fmvd_destroy_cuda_16);
    })
}

void fmvd_deconvolve_files_cuda_8(FILE** dataFiles, FILE* resultFile, int dataW, int dataH, int dataD,
                                  float** h_Weights, float** h_Kernel, int kernelH, int kernelW,
                                  fmvd_psf_type iteration_type, int nViews, int iterations)
{
    FMVD_API_GUARD(, {
        run_files<uint8_t, fmvd_plan_cuda_8>(dataFiles, resultFile, dataW, dataH, dataD, h_Weights,
                                             h_Kernel, kernelH, kernelW, iteration_type, nViews,
                                             iterations, fmvd_initialize_cuda_8,
                                             fmvd_deconvolve_planes_cuda_8, fmvd_destroy_cuda_8);
    })
}

}  // extern "C"
