// fmvd_deconvolve command line front end.
// Reference: fmvd_deconvolve_hdd.cpp:22-180 (same flags, same raw formats).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fmvd_cuda_utils.h"
#include "fmvd_deconvolve.h"

static void print_usage()
{
    printf("Usage:\n");
    printf("fmvd_deconvolve -v <nViews>\n");
    printf("                -i <nIterations>\n");
    printf("                -t <iteration type>\n");
    printf("                -o <output file>\n");
    printf("                -d <datafile view1> <datafile view2> ...\n");
    printf("                -w <weightfile view1> <weightfile view2> ...\n");
    printf("                -k <kernelfile view1> <kernelfile view2> ...\n");
    printf("                -dw <data width>\n");
    printf("                -dh <data height>\n");
    printf("                -dd <data depth>\n");
    printf("                -kw <kernel width>\n");
    printf("                -kh <kernel height>\n");
    printf("                [-8bit]\n");
}

static bool read_floats(const char* path, std::vector<float>& out, size_t n)
{
    FILE* f = fopen(path, "rb");
    if (!f) {
        printf("Could not open %s for reading\n", path);
        return false;
    }
    out.resize(n);
    const size_t got = fread(out.data(), sizeof(float), n, f);
    fclose(f);
    if (got != n) {
        printf("Short read from %s\n", path);
        return false;
    }
    return true;
}

int main(int argc, char** argv)
{
    int i = 1;
    if (argc < 3 || strcmp(argv[i++], "-v")) {
        print_usage();
        return -1;
    }
    const int nViews = atoi(argv[i++]);
    if (nViews < 1) {
        print_usage();
        return -1;
    }
    int nIterations = -1, w = -1, h = -1, d = -1, kw = -1, kh = -1, iterationType = 0, eightbit = 0;
    const char* outputfile = nullptr;
    std::vector<const char*> datafiles(nViews, nullptr), weightfiles(nViews, nullptr), kernelfiles(nViews, nullptr);
    auto need = [&](int k) {
        if (i + k >= argc) {
            print_usage();
            exit(-1);
        }
    };
    for (; i < argc; ++i) {
        if (!strcmp(argv[i], "-i")) { need(1); nIterations = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "-t")) { need(1); iterationType = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "-dw")) { need(1); w = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "-dh")) { need(1); h = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "-dd")) { need(1); d = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "-kw")) { need(1); kw = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "-kh")) { need(1); kh = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "-o")) { need(1); outputfile = argv[++i]; }
        else if (!strcmp(argv[i], "-d")) { need(nViews); for (int v = 0; v < nViews; ++v) datafiles[v] = argv[++i]; }
        else if (!strcmp(argv[i], "-w")) { need(nViews); for (int v = 0; v < nViews; ++v) weightfiles[v] = argv[++i]; }
        else if (!strcmp(argv[i], "-k")) { need(nViews); for (int v = 0; v < nViews; ++v) kernelfiles[v] = argv[++i]; }
        else if (!strcmp(argv[i], "-8bit")) eightbit = 1;
    }
    if (nIterations < 0 || w < 1 || h < 1 || d < 0 || kw < 1 || kh < 1 || !outputfile) {
        print_usage();
        return -1;
    }
    for (int v = 0; v < nViews; ++v)
        if (!datafiles[v] || !weightfiles[v] || !kernelfiles[v]) {
            print_usage();
            return -1;
        }

    printf("Starting plane-wise deconvolution\n");
    printf("\tnViews:        %d\n", nViews);
    printf("\tnIterations:   %d\n", nIterations);
    printf("\titerationType: %d\n", iterationType);
    printf("\tdimensions:    %d x %d x %d\n", w, h, d);
    printf("\tkernel dims:   %d x %d\n", kw, kh);

    std::vector<std::vector<float>> kernels(nViews), weights(nViews);
    std::vector<float*> kp(nViews), wp(nViews);
    for (int v = 0; v < nViews; ++v) {
        if (!read_floats(kernelfiles[v], kernels[v], (size_t)kw * kh)) return -1;
        if (!read_floats(weightfiles[v], weights[v], (size_t)w * h)) return -1;
        kp[v] = kernels[v].data();
        wp[v] = weights[v].data();
    }
    std::vector<FILE*> dataFiles(nViews);
    for (int v = 0; v < nViews; ++v) {
        dataFiles[v] = fopen(datafiles[v], "rb");
        if (!dataFiles[v]) {
            printf("Could not open %s for reading\n", datafiles[v]);
            return -1;
        }
    }
    FILE* resultFile = fopen(outputfile, "wb");
    if (!resultFile) {
        printf("Could not open %s for writing\n", outputfile);
        return -1;
    }
    const fmvd_psf_type t = (fmvd_psf_type)iterationType;
    if (eightbit)
        fmvd_deconvolve_files_cuda_8(dataFiles.data(), resultFile, w, h, d, wp.data(), kp.data(), kh, kw, t,
                                     nViews, nIterations);
    else
        fmvd_deconvolve_files_cuda_16(dataFiles.data(), resultFile, w, h, d, wp.data(), kp.data(), kh, kw, t,
                                      nViews, nIterations);
    for (FILE* f : dataFiles) fclose(f);
    fclose(resultFile);
    return EXIT_SUCCESS;
}
