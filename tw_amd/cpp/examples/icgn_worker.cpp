// icgn_worker.cpp -- the per-frame driver of TW_Core's IC-GN worker
// (ICGNWorkerThread ctor + processFrame, TW/TW_Core/icgnworkerthread.cpp:4-83)
// with the GPU class in place of ICGN2D_CPU: the same constructor arguments
// (20 iterations, delta_p 0.001, bicubic), and per frame the same two calls on
// host buffers, ResetRefImg(reference) then ICGN2D_Algorithm(fU, fV,
// iterations, POIs, target).  The integer guesses fU / fV and the POIs come in
// host arrays, as TW_Core's FFT-CC thread hands them over its queues.
//
// Reads: POIs (int32 {y, x} x n), per frame the initial u, v (float x n each),
// and raw 8-bit (reference, target) frame pairs.  Writes per frame: u, v (float),
// iterations (int32), flags (uint8).  Usage:
//   tw_icgn_worker W H roiX roiY roiW roiH S nx ny pois.bin uv0.bin out.bin ref1.raw tar1.raw [ref2.raw tar2.raw ...]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "tw/TW_paDIC_cuICGN2D.h"

using namespace TW;
using namespace TW::paDIC;

static std::vector<unsigned char> read_file(const char* path, size_t n)
{
    std::vector<unsigned char> v(n);
    FILE* f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), 1, n, f) != n) {
        if (f) std::fclose(f);
        throw std::runtime_error(std::string("cannot read ") + path);
    }
    std::fclose(f);
    return v;
}

int main(int argc, char** argv)
{
    if (argc < 15 || (argc - 13) % 2 != 0) {
        std::fprintf(stderr, "usage: %s W H roiX roiY roiW roiH S nx ny pois uv0 out ref1 tar1 [ref2 tar2 ...]\n",
                     argv[0]);
        return 2;
    }
    try {
        const int W = std::atoi(argv[1]), H = std::atoi(argv[2]);
        const int rx = std::atoi(argv[3]), ry = std::atoi(argv[4]), rw = std::atoi(argv[5]), rh = std::atoi(argv[6]);
        const int S = std::atoi(argv[7]), nx = std::atoi(argv[8]), ny = std::atoi(argv[9]);
        const size_t n = (size_t)nx * ny;
        const int frames = (argc - 13) / 2;
        const auto poi_raw = read_file(argv[10], 8 * n);
        const auto uv_raw = read_file(argv[11], 8 * n * frames);
        std::vector<int> poi(2 * n);
        std::memcpy(poi.data(), poi_raw.data(), 8 * n);

        // icgnworkerthread.cpp:27-36 (the thread flag has no GPU meaning)
        cuICGN2D icgn(W, H, rx, ry, rw, rh, S, S, nx, ny, 20, 0.001f, ICGN2DInterpolationFLag::Bicubic);

        FILE* fo = std::fopen(argv[12], "wb");
        if (!fo) throw std::runtime_error("cannot open output");
        std::vector<real_t> fU(n), fV(n);
        std::vector<int> iters(n);
        std::vector<ICGN2DFlag> flags(n);
        std::vector<unsigned char> fl(n);
        for (int k = 0; k < frames; ++k) {
            const auto ref = read_file(argv[13 + 2 * k], (size_t)W * H);
            const auto tar = read_file(argv[14 + 2 * k], (size_t)W * H);
            // processFrame (icgnworkerthread.cpp:45-73): new reference, then
            // IC-GN from the FFT-CC guesses on the host buffers
            icgn.ResetRefImg(MatView(ref.data(), H, W, W));
            std::memcpy(fU.data(), uv_raw.data() + 8 * n * k, 4 * n);
            std::memcpy(fV.data(), uv_raw.data() + 8 * n * k + 4 * n, 4 * n);
            icgn.ICGN2D_Algorithm(fU.data(), fV.data(), iters.data(), poi.data(), MatView(tar.data(), H, W, W),
                                  flags.data());
            for (size_t i = 0; i < n; ++i) fl[i] = (unsigned char)flags[i];
            // the [head, tail) overload runs every POI, as the reference's
            // does: the same inputs must give the same outputs
            std::vector<real_t> hU(n), hV(n);
            std::vector<int> hIt(iters);
            std::memcpy(hU.data(), uv_raw.data() + 8 * n * k, 4 * n);
            std::memcpy(hV.data(), uv_raw.data() + 8 * n * k + 4 * n, 4 * n);
            icgn.ICGN2D_Algorithm(hU.data(), hV.data(), hIt.data(), poi.data(), MatView(tar.data(), H, W, W),
                                  (int_t)0, (int_t)(n / 2));
            if (std::memcmp(hU.data(), fU.data(), 4 * n) || std::memcmp(hV.data(), fV.data(), 4 * n) ||
                hIt != iters)
                throw std::runtime_error("the head/tail overload differs from the full one");
            std::fwrite(fU.data(), sizeof(real_t), n, fo);
            std::fwrite(fV.data(), sizeof(real_t), n, fo);
            std::fwrite(iters.data(), sizeof(int), n, fo);
            std::fwrite(fl.data(), 1, n, fo);
        }
        std::fclose(fo);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
