// twdic_internal.h -- shared declarations of the HIP implementation of the
// paDIC hot path (context, kernel parameter blocks, launch wrappers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "twdic.h"

namespace twdic {

constexpr int kWave = 64;             // CDNA wavefront width
constexpr int kMaxFftN = 64;          // largest FFT-CC window edge held in LDS
constexpr int kRefStateStride = 48;   // doubles of per-reference state per POI (384 B)

// Per-reference, per-POI state (doubles, kRefStateStride per POI):
//   [0..20]  FP64 Cholesky factor L of the 6x6 Hessian, packed lower, row-major
//   [21..26] A_k = sum SD_k               (steepest-descent sums)
//   [27..32] B_k = sum SD_k (R - mean R)  (BR_k - mean R A_k with BR_k = sum SD_k R, one FMA)
//   [33]     |R - mean R|_2               (R_s)
//   [34]     reference flag: 0 ok, TWDIC_ICGN_DARK_SUBSET, TWDIC_ICGN_SINGULAR_HESSIAN,
//            TWDIC_ICGN_OUT_OF_ROI (subset leaves the ROI)
//   [35]     mean R
//   [36..41] 1 / L_ii (reciprocal pivots of the factor)
enum { RS_L = 0, RS_A = 21, RS_B = 27, RS_NORM = 33, RS_FLAG = 34, RS_MEAN = 35, RS_RDIAG = 36 };
constexpr int kRefStateUsed = 42;     // the doubles above that the iteration kernels read

// IC-GN runs one POI per group of L lanes of a wavefront, 256 / L POIs per
// 256-thread workgroup: L = 32 (a half-wavefront) for sets of at least
// kIcgnHalfWaveMin POIs, where the two POIs of a wavefront share the per-pass
// work between the pixel loops; L = 64 for smaller sets, which leave the GPU
// partly idle and run faster with half the pixel rounds per pass.  The
// choice follows the whole set's POI count (also for a chunk of it, and per
// pair in a batch), so a POI's result does not depend on the path that ran it.
constexpr int kIcgnHalfWaveMin = 32768;
inline int icgn_lanes_for(int n) { return n >= kIcgnHalfWaveMin ? 32 : 64; }

// Mean and L2 spread of n samples from their exact-or-FP64 sums s1 = sum x,
// s2 = sum x^2.  Shared by the reference-state and iteration kernels so that
// identical subsets give bit-identical statistics.
__device__ __forceinline__ void mean_spread(double s1, double s2, double n, double& mean, double& spread)
{
    mean = s1 / n;
    double var = __fma_rn(-s1, mean, s2);
    spread = sqrt(var > 0.0 ? var : 0.0);
}

// sum SD_k (x - mean) from sum SD_k x and sum SD_k.
__device__ __forceinline__ double centred(double sdx, double mean, double a)
{
    return __fma_rn(-mean, a, sdx);
}

struct FftccParams {
    const uint8_t* ref;
    const uint8_t* tar;
    int stride;
    int img_w, img_h;
    const int32_t* poi;
    int n;
    int N;
    const float2* tw;     // N twiddles exp(-2 pi i e / N)
    float tol;            // near-tie tolerance in ZNCC units
    float* u;
    float* v;
    float* zncc;
    uint8_t* status;
};

struct IcgnParams {
    const uint8_t* ref;
    const uint8_t* tar;
    int stride;
    int sx0, sy0, rw, rh;     // ROI
    int Sx, Sy, Wsub, Hsub, s;
    int max_iter;
    int lanes;                // lanes per POI: 32 or 64 (icgn_lanes_for of the set's POI count)
    double delta_p;           // the caller's float, widened: the loop test compares in FP64
    const uint32_t* g2;       // reference gradients (2Rx, 2Ry) as int16 pairs, ROI-relative
    const float* lut;         // target LUT [rh][rw][16]
    const uint4* rp;          // target row polynomials [rh + 1][rw] (fused mode; null: from the image bytes)
    double* refstate;         // kRefStateStride doubles per POI
    const int32_t* poi;
    int n;
    float* u;
    float* v;
    float* p6;
    int32_t* iters;
    uint8_t* flags;
    float* dpnorm;
};

// One image pair of a batch: the same parameter blocks the single-context
// kernels take, so the persistent kernels reuse the per-POI device code.
struct PairDesc {
    IcgnParams icgn;
    FftccParams fftcc;
};

// Descriptors are never written while a batch kernel runs: reading them
// through the constant address space lets the compiler use scalar loads and
// keep each wave's descriptor in SGPRs.
#if defined(__HIP_DEVICE_COMPILE__)
#define TW_CONST_AS __attribute__((address_space(4)))
#else
#define TW_CONST_AS
#endif
typedef const TW_CONST_AS PairDesc* ConstPairDescPtr;
__device__ __forceinline__ ConstPairDescPtr const_pairs(const PairDesc* p)
{
    return (ConstPairDescPtr)p;
}

// Kernel launch wrappers (return hipError_t of the launch).
hipError_t launch_gradients(const uint8_t* img, int stride, int sx0, int sy0, int rw, int rh,
                            float* gx, float* gy, float* gxy, uint32_t* g2, hipStream_t st);
hipError_t launch_bspline_lut(const uint8_t* tar, int stride, int sx0, int sy0, int rw, int rh, float* lut,
                              hipStream_t st);
// fp16 Catmull-Rom row polynomials of the target, (rh + 1) x rw entries of 16 bytes
hipError_t launch_row_poly(const uint8_t* tar, int stride, int sx0, int sy0, int rw, int rh, uint4* rp,
                           hipStream_t st);
// the same for every pair of a batch with a plane (descriptor icgn.rp set);
// max_rh: the largest ROI height among them
hipError_t launch_batch_row_poly(const PairDesc* pairs, int npairs, int max_rh, hipStream_t st);
hipError_t launch_bicubic_lut(const uint8_t* tar, int stride, int sx0, int sy0, int rw, int rh,
                              float* lut, hipStream_t st);
hipError_t launch_poi_grid(int32_t* poi, int start_x, int start_y, int nx, int ny, int mx, int my,
                           int sx, int sy, int gx, int gy, hipStream_t st);
// variant kFftccMma: exact integer correlation on the matrix cores (any even
// N <= 64); 2, 3, 4: the FFT register engine's occupancy target (N = 32, 64,
// else the direct-DFT engine)
constexpr int kFftccMma = 8;
// kFftccDftMfma: N = 64 on fp16 matrix-core DFTs with exact re-evaluation
// of the near-maximal shifts (other N: the integer engine); needs
// dft_tables_init on the device first
constexpr int kFftccDftMfma = 9;
hipError_t dft_tables_init(int device);
hipError_t launch_fftcc(const FftccParams& p, int variant, hipStream_t st);
// batched form: the matrix-core kernel over items (pair, POI) of pairs that
// share a kernel form (mma_form of their window edge N: 0 for N = 64,
// 1 for 48 < N < 64, 2 for 32 < N <= 48, 3 for N <= 32)
int mma_form(int N);
hipError_t launch_batch_fftcc_mma(const PairDesc* pairs, const int2* items, int count, int form, hipStream_t st);
// batched form: IC-GN (fused interpolation) over workgroup items (pair, first
// POI), 256 / lanes POIs of one pair per workgroup, every listed pair with
// that many lanes per POI; `table` when every listed pair's
// subset fits the step table (step_table_fits).
bool step_table_fits(int wsub, int hsub);
// The row-polynomial plane of an ROI is addressed with 32-bit byte offsets
// 16 (iy rw + ix) from entry -(sy0 rw + sx0); true when they all fit.
inline bool row_poly_fits(int sx0, int sy0, int rw, int rh)
{
    return 16ull * ((unsigned long long)(sy0 + rh + 2) * (unsigned long long)rw + (unsigned long long)sx0) <
           (1ull << 32);
}
inline size_t row_poly_bytes(int rw, int rh)
{
    return 16ull * (size_t)rw * (size_t)(rh + 1);
}
// waves: occupancy target 5 or 6; xcd_ranges: contiguous per-XCD item ranges;
// all_rp: every pair's descriptor has its row-polynomial plane
hipError_t launch_batch_icgn_wg(const PairDesc* pairs, const int2* wg, int count, bool table, int lanes, int waves,
                                bool xcd_ranges, bool all_rp, hipStream_t st);
hipError_t launch_ref_state(const IcgnParams& p, hipStream_t st);
hipError_t launch_icgn(const IcgnParams& p, int interp, int variant, hipStream_t st);
// persistent batch kernels: `blocks` workgroups drain a device work counter
hipError_t launch_batch_icgn(const PairDesc* pairs, const int2* items, int total, int* counter, int chunk,
                             int blocks, hipStream_t st);
hipError_t launch_batch_ref_state(const PairDesc* pairs, const int2* items, int total, int* counter, int blocks,
                                  hipStream_t st);
hipError_t launch_batch_fftcc(const PairDesc* pairs, const int2* items, int total, int* counter, int N,
                              int variant, int blocks, hipStream_t st);
hipError_t launch_fill_zero(int* p, int n, hipStream_t st);
hipError_t launch_strain_ls(const float* f, int nx, int ny, int grid, int M, double* dfdx, double* dfdy,
                            hipStream_t st);
// POI tracking across reference resets
hipError_t launch_track_update_poi(int32_t* poi, int n, const float* u, const float* v, hipStream_t st);
hipError_t launch_track_current_poi(const int32_t* poi, int n, const float* u, const float* v, int32_t* out,
                                    hipStream_t st);
hipError_t launch_track_current_poi_acc(const int32_t* poi, int n, const float* u, const float* v, const float* acc_u,
                                        const float* acc_v, int32_t* out, hipStream_t st);
hipError_t launch_track_accumulate_uv(int n, const float* u, const float* v, float* acc_u, float* acc_v,
                                      hipStream_t st);
// occupancy-derived persistent grid sizes
int persistent_blocks_icgn(int device);
int persistent_blocks_state(int device);
int persistent_blocks_fftcc(int device, int N, int variant);

}  // namespace twdic

struct twdic_ctx {
    twdic_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    int nx = 0, ny = 0;
    int dstride = 0;                  // device image row stride (bytes)
    uint8_t* d_ref = nullptr;
    uint8_t* d_tar = nullptr;
    float* d_rx = nullptr;
    float* d_ry = nullptr;
    float* d_tx = nullptr;
    float* d_ty = nullptr;
    float* d_txy = nullptr;
    float* d_lut = nullptr;
    uint4* d_rp = nullptr;            // target row polynomials (fused interpolation)
    uint32_t* d_g2 = nullptr;         // packed (2Rx, 2Ry) int16 pairs
    float2* d_fft_tw = nullptr;
    bool ref_loaded = false, tar_loaded = false;
    bool ref_grads_valid = false;     // Rx/Ry match d_ref
    bool tar_pre_valid = false;       // Tx/Ty/Txy/LUT match d_tar
    bool tar_rp_valid = false;        // d_rp matches d_tar
    bool tar_exposed = false;         // d_tar was handed out: it may change without twdic_set_target
    uint64_t ref_version = 0;
    // per-reference POI state cache key
    double* d_refstate = nullptr;
    size_t refstate_cap = 0;          // POIs
    uint64_t refstate_version = ~0ull;
    const int32_t* refstate_poi = nullptr;
    int refstate_n = -1;
    // FFT-CC engine (TWDIC_FFTCC_VARIANT): 9 fp16 matrix-core DFTs for N = 64 (else
    // 8), 8 exact integer correlation on the matrix cores, 2-4 FFT registers
    int fftcc_variant = twdic::kFftccDftMfma;
    int icgn_variant = 5;             // occupancy target of the IC-GN kernel (0: compiler's choice)
    uint8_t* h_stage = nullptr;       // pinned staging for host images
    size_t h_stage_bytes = 0;
    hipEvent_t ev[16] = {};
    // frame pipeline: FFT-CC of POI chunk i + 1 (on `stream`) overlaps IC-GN of
    // chunk i (on `stream2`)
    static constexpr int kMaxFrameChunks = 16;
    hipStream_t stream2 = nullptr;
    hipEvent_t pev[kMaxFrameChunks + 1] = {};
    int frame_chunks = 1;             // TWDIC_FRAME_CHUNKS; 1 runs the two stages back to back
    // captured frame step (twdic_frame_capture): its graph, the IC-GN
    // parameters it was recorded with, and the per-POI state buffer they name
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    twdic::IcgnParams graph_ip{};
    std::string err;
};
