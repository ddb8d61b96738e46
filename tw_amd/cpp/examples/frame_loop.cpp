// frame_loop.cpp -- the per-frame driver of TW_Core's real-time variant
// (FFTCCTWorkerThread ctor + processFrameFFTCC_ICGN,
// TW/TW_Core/fftcctworkerthread.cpp:9-103, 399-572) written against the
// drop-in classes: the same constructor calls, reference reset cadence, the
// g_cuHandle members handed from cuFFTCC2D to cuICGN2D, and the POI tracking
// around a reset (accumulated U/V, POIs advanced by the last integer shift).
// On a reset the current frame becomes the reference (TW_Core dequeues it
// from its reference buffer).  Reads raw 8-bit frames, writes per frame:
// u, v, zncc (float), iterations (int32), flags (uint8), p6 (float x 6),
// current POIs (int32 {y, x}), accumulated u, v (float).  Usage:
//   tw_frame_loop W H roiX roiY roiW roiH S grid margin N refPeriod out.bin ref.raw tar1.raw [tar2.raw ...]
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "tw/TW_paDIC_cuFFTCC2D.h"
#include "tw/TW_cuda_utils.h"
#include "tw/TW_paDIC_cuICGN2D.h"
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
    if (argc < 15) {
        std::fprintf(stderr, "usage: %s W H roiX roiY roiW roiH S grid margin N refPeriod out ref tar...\n", argv[0]);
        return 2;
    }
    try {
        const int W = std::atoi(argv[1]), H = std::atoi(argv[2]);
        const int rx = std::atoi(argv[3]), ry = std::atoi(argv[4]), rw = std::atoi(argv[5]), rh = std::atoi(argv[6]);
        const int S = std::atoi(argv[7]), G = std::atoi(argv[8]), M = std::atoi(argv[9]), N = std::atoi(argv[10]);
        const int period = std::atoi(argv[11]);
        const char* out = argv[12];
        auto ref = read_raw(argv[13], (size_t)W * H);
        MatView refMat(ref.data(), H, W, W);

        // fftcctworkerthread.cpp:51-57 (note the ROI size before the ROI start)
        cuFFTCC2D fftcc(W, H, rw, rh, rx, ry, S, S, G, G, M, M);
        if (N != 2 * S) fftcc.setFFTWindow(N);
        // fftcctworkerthread.cpp:66-74
        cuICGN2D icgn(W, H, rx, ry, rw, rh, S, S, fftcc.GetNumPOIsX(), fftcc.GetNumPOIsY(), 20, 0.001f,
                      ICGN2DInterpolationFLag::Bicubic);
        real_t *d_U = nullptr, *d_V = nullptr, *d_ZNCC = nullptr;
        fftcc.cuInitializeFFTCC(d_U, d_V, d_ZNCC, refMat);             // :89
        icgn.cuInitialize(fftcc.g_cuHandle.m_d_fRefImg);
        const int n = fftcc.GetNumPOIs();
        // fftcctworkerthread.cpp:79-90: accumulators zeroed, current POIs = POIs
        twdic_ctx* fc = fftcc.context();
        real_t *d_accU = nullptr, *d_accV = nullptr;
        int* d_curPOI = nullptr;
        twdic_malloc(fc, (void**)&d_accU, n * sizeof(real_t));
        twdic_malloc(fc, (void**)&d_accV, n * sizeof(real_t));
        twdic_malloc(fc, (void**)&d_curPOI, 2 * n * sizeof(int));
        twdic_memset(fc, d_accU, 0, n * sizeof(real_t));
        twdic_memset(fc, d_accV, 0, n * sizeof(real_t));
        twdic_memcpy(fc, d_curPOI, fftcc.g_cuHandle.m_d_iPOIXY, 2 * n * sizeof(int), 2);
        twdic_sync(fc);
        FILE* fo = std::fopen(out, "wb");
        if (!fo) throw std::runtime_error("cannot open output");
        std::vector<float> u(n), v(n), z(n), p6(6 * n);
        std::vector<int> it(n), cur(2 * n);
        std::vector<float> au(n), av(n);
        std::vector<unsigned char> fl(n);
        for (int k = 13 + 1, count = 1; k < argc; ++k, ++count) {
            auto tar = read_raw(argv[k], (size_t)W * H);
            MatView tarMat(tar.data(), H, W, W);
            if (count % period == 1 && count > 1) {                    // :408-443
                fftcc.ResetRefImg(tarMat);
                icgn.cuInitialize(fftcc.g_cuHandle.m_d_fRefImg);
                // :419-437 (count > period here): accumulate the last frame's
                // U/V, then advance the POIs by its integer part
                cuAccumulateUV(d_accU, d_accV, n, d_U, d_V);
                cuUpdatePOIpos(d_U, d_V, fftcc.GetNumPOIsX(), fftcc.GetNumPOIsY(), fftcc.g_cuHandle.m_d_iPOIXY);
            }
            fftcc.cuComputeFFTCC(d_U, d_V, d_ZNCC, tarMat);           // :449-453
            twdic_memcpy(fftcc.context(), z.data(), d_ZNCC, n * 4, 1);
            icgn.cuCompute(fftcc.g_cuHandle.m_d_fTarImg, fftcc.g_cuHandle.m_d_iPOIXY, d_U, d_V);   // :457-461
            cuAccumulatePOI(d_U, d_V, fftcc.g_cuHandle.m_d_iPOIXY, n, d_curPOI);   // :476-481
            twdic_ctx* c = icgn.context();
            twdic_memcpy(c, u.data(), d_U, n * 4, 1);
            twdic_memcpy(c, v.data(), d_V, n * 4, 1);
            twdic_memcpy(c, it.data(), icgn.g_cuHandleICGN.m_d_iIterationNums, n * 4, 1);
            twdic_memcpy(c, fl.data(), icgn.g_cuHandleICGN.m_d_iFlags, n, 1);
            twdic_memcpy(c, p6.data(), icgn.g_cuHandleICGN.m_d_dP, n * 24, 1);
            twdic_memcpy(c, cur.data(), d_curPOI, n * 8, 1);
            twdic_memcpy(c, au.data(), d_accU, n * 4, 1);
            twdic_memcpy(c, av.data(), d_accV, n * 4, 1);
            std::fwrite(u.data(), 4, n, fo);
            std::fwrite(v.data(), 4, n, fo);
            std::fwrite(z.data(), 4, n, fo);
            std::fwrite(it.data(), 4, n, fo);
            std::fwrite(fl.data(), 1, n, fo);
            std::fwrite(p6.data(), 4, 6 * n, fo);
            std::fwrite(cur.data(), 4, 2 * n, fo);
            std::fwrite(au.data(), 4, n, fo);
            std::fwrite(av.data(), 4, n, fo);
        }
        std::fclose(fo);
        twdic_free(fc, d_accU);
        twdic_free(fc, d_accV);
        twdic_free(fc, d_curPOI);
        fftcc.cuDestroyFFTCC(d_U, d_V, d_ZNCC);
        std::printf("frame_loop ok: %d POIs, %d frames\n", n, argc - 14);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "frame_loop error: %s\n", e.what());
        return 1;
    }
    return 0;
}
