// fftcc_host.cpp -- cuFFTCC2D's host-output API (InitializeFFTCC /
// ComputeFFTCC / DestroyFFTCC, TW/TW_Engine/TW_paDIC_cuFFTCC2D.cu:237-461)
// driven as a caller of the reference would: the outputs are hcreateptr
// arrays [NumPOIY][NumPOIX] the callee allocates, read through their row
// pointers.  Two constructors:
//   whole: cuFFTCC2D(ImgW, ImgH, ROIW, ROIH, StartX, StartY, ...)
//          (TW_paDIC_cuFFTCC2D.cu:192-210), POIs in image coordinates;
//   roi:   cuFFTCC2D(ROIW, ROIH, ...) (TW_paDIC_cuFFTCC2D.cu:178-190), the
//          images ARE the ROI and the POIs are ROI-relative (:254-266).
// Writes NumPOIX, NumPOIY (int32), the POI array (int32 {y, x} per POI), then
// per target frame u, v, zncc (float, row-major [NumPOIY][NumPOIX]).  Usage:
//   tw_fftcc_host whole W H roiX roiY roiW roiH S grid margin out ref.raw tar.raw [tar.raw ...]
//   tw_fftcc_host roi   W H S grid margin out ref.raw tar.raw [tar.raw ...]
// (in roi mode W x H is the ROI and the size of every image)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "tw/TW_paDIC_cuFFTCC2D.h"
#include "twdic.h"

using namespace TW;
using namespace TW::paDIC;

static std::vector<unsigned char> read_raw(const char* path, size_t n)
{
    std::vector<unsigned char> v(n);
    FILE* f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), 1, n, f) != n) throw std::runtime_error(std::string("cannot read ") + path);
    std::fclose(f);
    return v;
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s whole|roi ...\n", argv[0]);
        return 2;
    }
    const bool whole = std::strcmp(argv[1], "whole") == 0;
    const int need = whole ? 14 : 10;
    if (argc < need) {
        std::fprintf(stderr, "usage: see the header of fftcc_host.cpp\n");
        return 2;
    }
    try {
        int a = 2;
        const int W = std::atoi(argv[a++]), H = std::atoi(argv[a++]);
        int rx = 0, ry = 0, rw = W, rh = H;
        if (whole) {
            rx = std::atoi(argv[a++]); ry = std::atoi(argv[a++]);
            rw = std::atoi(argv[a++]); rh = std::atoi(argv[a++]);
        }
        const int S = std::atoi(argv[a++]), G = std::atoi(argv[a++]), M = std::atoi(argv[a++]);
        const char* out = argv[a++];
        auto ref = read_raw(argv[a++], (size_t)W * H);
        MatView refMat(ref.data(), H, W, W);
        cuFFTCC2D* cc = whole ? new cuFFTCC2D(W, H, rw, rh, rx, ry, S, S, G, G, M, M)
                              : new cuFFTCC2D(rw, rh, S, S, G, G, M, M);
        real_t **fU = nullptr, **fV = nullptr, **fZ = nullptr;
        cc->InitializeFFTCC(fU, fV, fZ, refMat);
        const int nx = cc->GetNumPOIsX(), ny = cc->GetNumPOIsY(), n = cc->GetNumPOIs();
        std::vector<int32_t> poi(2 * (size_t)n);
        if (twdic_memcpy(cc->context(), poi.data(), cc->g_cuHandle.m_d_iPOIXY, poi.size() * sizeof(int32_t), 1) !=
            TWDIC_OK)
            throw std::runtime_error("POI download");
        FILE* f = std::fopen(out, "wb");
        if (!f) throw std::runtime_error("cannot open output");
        std::fwrite(&nx, 4, 1, f);
        std::fwrite(&ny, 4, 1, f);
        std::fwrite(poi.data(), 4, poi.size(), f);
        std::vector<float> row((size_t)nx);
        for (; a < argc; ++a) {
            auto tar = read_raw(argv[a], (size_t)W * H);
            cc->ComputeFFTCC(fU, fV, fZ, MatView(tar.data(), H, W, W));
            for (real_t** arr : {fU, fV, fZ})
                for (int iy = 0; iy < ny; ++iy) {       // through the row pointers
                    for (int ix = 0; ix < nx; ++ix) row[ix] = arr[iy][ix];
                    std::fwrite(row.data(), 4, row.size(), f);
                }
        }
        std::fclose(f);
        cc->DestroyFFTCC(fU, fV, fZ);
        if (fU || fV || fZ) throw std::runtime_error("DestroyFFTCC left an output pointer set");
        delete cc;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
