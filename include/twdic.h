/*
 * twdic.h -- C ABI of the MI355X-native paDIC hot path (FFT-CC integer-pixel
 * search + first-order IC-GN with bicubic interpolation).
 *
 * Plain pointers and sizes only.  Every call returns a twdic_status and never
 * exits the process (the reference calls exit() on any CUDA error,
 * helper_cuda.h:964-997).  One context owns one HIP stream and all device
 * state; calls on one context are asynchronous on that stream unless stated.
 * A context is not thread-safe; use one per host thread.
 *
 * Pointer convention: arguments named d_* are device pointers (allocate them
 * with twdic_malloc or any HIP allocator on the context's device); arguments
 * named h_* are host pointers.  Images are 8-bit single channel, row-major,
 * with a row stride in bytes; the context's own device copies use a row
 * stride equal to the image width (as the reference's device images do).
 *
 * The entries below replace the reference's class methods; each comment
 * names the reference interface it replaces (paths relative to the reference
 * checkout).
 */
#ifndef TWDIC_H
#define TWDIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct twdic_ctx twdic_ctx;

typedef enum {
    TWDIC_OK = 0,
    TWDIC_EINVAL = 1,   /* bad argument or geometry */
    TWDIC_ENOMEM = 2,   /* device or pinned allocation failed */
    TWDIC_EBOUNDS = 3,  /* a device buffer is smaller than the call needs (n entries,
                           an image of stride x height); nothing was launched */
    TWDIC_EHIP = 4,     /* HIP runtime error (message in twdic_last_error) */
    TWDIC_EBUSY = 5,    /* serving queue full (admission refused) */
    TWDIC_ENODEV = 6    /* no usable GPU */
} twdic_status;

/* Per-POI IC-GN result flag: ICGN2DFlag (TW/TW_Engine/TW_paDIC.h:74-81). */
enum { TWDIC_ICGN_SUCCESS = 0, TWDIC_ICGN_DARK_SUBSET = 1,
       TWDIC_ICGN_SINGULAR_HESSIAN = 2, TWDIC_ICGN_SINGULAR_WARP = 3,
       TWDIC_ICGN_OUT_OF_ROI = 4 };

/* Per-POI FFT-CC status.  FLAT: a window is uniform; the reference returns
 * early and leaves u/v/zncc untouched (TW_paDIC_FFTCC2D_CPU.cpp:271-275), and
 * so does this library.  OUT: the N x N window leaves the image (outputs
 * untouched; the reference reads out of bounds there). */
enum { TWDIC_FFTCC_OK = 0, TWDIC_FFTCC_FLAT = 1, TWDIC_FFTCC_OUT = 2 };

/* IC-GN interpolation source. */
enum {
    TWDIC_INTERP_LUT = 0,   /* materialise the 16-coefficient bicubic LUT of the
                               target each frame and gather from it, as the
                               reference does (TW_cuUtils.cu:314-435) */
    TWDIC_INTERP_FUSED = 1, /* evaluate the same bicubic interpolant (tensor
                               Catmull-Rom, which equals the LUT polynomial built
                               from central-difference derivatives) from the
                               target rows' cubic coefficients: integers stored
                               exactly in fp16, 16 bytes per ROI pixel, built once
                               per target (every call once TWDIC_BUF_TAR has been
                               handed out), instead of the 64-byte LUT */
    TWDIC_INTERP_BSPLINE = 2 /* ICGN2DInterpolationFLag::BicubicSpline: the cubic
                               B-spline LUT of BicubicSplineCoefficients_s
                               (TW_utils.cpp:521-602), CPU-only in the reference
                               (its GPU version is unfinished, TW_cuUtils.cu:437-489).
                               Needs a 2-pixel frame right and below the ROI. */
};

/* Configuration: the union of the reference constructors' arguments
 * (cuFFTCC2D, TW_paDIC_cuFFTCC2D.h:52-57; cuICGN2D, TW_paDIC_cuICGN2D.h:14-22).
 * fft_n: FFT-CC window edge N (even, 2..64); N = 2*sx reproduces the
 * reference (TW_paDIC_FFTCC2D_CPU.cpp:24-25). */
typedef struct {
    int img_w, img_h;
    int roi_x, roi_y, roi_w, roi_h;
    int sx, sy;          /* subset half sizes: subset (2sy+1) x (2sx+1) */
    int gx, gy;          /* POI grid spacing */
    int mx, my;          /* ROI margins */
    int fft_n;           /* FFT-CC window edge */
    int max_iter;        /* IC-GN maximum iterations (reference: 20) */
    float delta_p;       /* IC-GN convergence threshold (reference: 0.001) */
    int interp;          /* TWDIC_INTERP_* */
} twdic_config;

/* ---------------------------------------------------------------- context */

/* Replaces the cuFFTCC2D / cuICGN2D constructors + cuICGN2D::prepare
 * (TW_paDIC_cuFFTCC2D.cu:192-210, TW_paDIC_cuICGN2D.cu:694-768):
 * validates the geometry (std::logic_error "Number of POIs is below 0!" in
 * the reference becomes TWDIC_EINVAL; images are limited to 65535 px per
 * side and 2^31 bytes, since the kernels use 32-bit image offsets) and
 * allocates ROI-sized state. */
twdic_status twdic_create(const twdic_config* cfg, int device, twdic_ctx** out);
/* Replaces ~cuFFTCC2D / ~cuICGN2D / cuDestroyFFTCC / cuFinalize
 * (TW_paDIC_cuFFTCC2D.cu:212-233,649-677; TW_paDIC_cuICGN2D.cu:719-729,970-989).
 * Frees only what the context allocated (borrowed pointers are never freed). */
twdic_status twdic_destroy(twdic_ctx* ctx);
const char* twdic_last_error(const twdic_ctx* ctx);
/* Blocks until all work queued on the context's stream has finished. */
twdic_status twdic_sync(twdic_ctx* ctx);
/* The context's hipStream_t, as void*. */
void* twdic_stream(twdic_ctx* ctx);

/* POI counts, Fftcc2D::GetNumPOIsX/Y (TW_paDIC_FFTCC2D.h:99-101) after
 * recomputeNumPOI (TW_paDIC_FFTCC2D.cpp:88-94). */
twdic_status twdic_num_poi(const twdic_config* cfg, int32_t* nx, int32_t* ny);

/* ------------------------------------------------------------------ images */

/* Replaces cuFFTCC2D::ResetRefImg / cuICGN2D::cuInitialize / ResetRefImg
 * (TW_paDIC_cuFFTCC2D.cu:679-685; TW_paDIC_cuICGN2D.cu:770-793).  Copies the
 * image into the context (h_img host, or d_img device when on_device != 0;
 * stride in bytes) and marks the per-reference state stale; it is rebuilt
 * lazily by the next IC-GN call for the POI set used there.  Both copies are
 * ordered on this context's stream only: a device source written on another
 * stream must be complete before the call, and the source may be reused
 * once twdic_sync(ctx) has returned. */
twdic_status twdic_set_reference(twdic_ctx* ctx, const uint8_t* img, int stride, int on_device);
/* Replaces cuFFTCC2D::SetTarImg / the target upload in cuComputeFFTCC
 * (TW_paDIC_cuFFTCC2D.cu:560-563,687-693).  Host images are staged through
 * pinned memory and copied asynchronously on the context's stream. */
twdic_status twdic_set_target(twdic_ctx* ctx, const uint8_t* img, int stride, int on_device);

/* -------------------------------------------------------------- POI grid */

/* Replaces cuComputePOIPositions (TW_cuUtils.cu:496-605) and
 * ComputePOIPositions_s/_m (TW_utils.cpp:8-94): writes int32[n][2] = {y, x}
 * (y first, whole-image coordinates) to d_poi_yx.  n = nx*ny. */
twdic_status twdic_poi_grid(twdic_ctx* ctx, int32_t* d_poi_yx);

/* ---------------------------------------------------------------- FFT-CC */

/* Replaces cuFFTCC2D::cuComputeFFTCC's device pipeline (cufft_prepare_kernel,
 * cufftExecR2C x2, complexMulandScale_kernel, cufftExecC2R, findMax;
 * TW_paDIC_cuFFTCC2D.cu:575-638) with CPU-path semantics
 * (Fftcc2D_CPU::ComputeFFTCC, TW_paDIC_FFTCC2D_CPU.cpp:222-326):
 * integer peak = exact argmax of the circular cross-correlation of the two
 * N x N windows, lowest row-major index on ties, wrapped to (-N/2, N/2];
 * ZNCC = peak / (N^2 sqrt(sum R~^2 sum T~^2)).  d_status may be NULL. */
twdic_status twdic_fftcc(twdic_ctx* ctx, const int32_t* d_poi_yx, int n,
                         float* d_u, float* d_v, float* d_zncc, uint8_t* d_status);

/* ----------------------------------------------------------------- IC-GN */

/* Replaces cuICGN2D::cuCompute (TW_paDIC_cuICGN2D.cu:811-968) with CPU-path
 * semantics (ICGN2D_CPU::ICGN2D_Compute, TW_paDIC_ICGN2D_CPU.cpp:481-925):
 * per-frame target precompute (gradients + bicubic LUT when interp = LUT),
 * per-reference state (reference subsets, steepest-descent sums, FP64
 * Cholesky factor of the Hessian) rebuilt when the reference or the POI set
 * changed, then the IC-GN iterations.
 * d_u/d_v: in = initial guess (e.g. FFT-CC result), out = sub-pixel result.
 * d_p6 (may be NULL): final (u, ux, uy, v, vx, vy) per POI.
 * d_iters (may be NULL): 1-based iteration count; untouched when a flag is
 * raised before the first update (as in the reference).
 * d_flags (may be NULL): ICGN2DFlag per POI.
 * d_dpnorm (may be NULL): weighted |dp| of the last update (audit aid).
 * The per-POI reference state is cached under (reference upload, d_poi_yx,
 * n); after rewriting a POI array in place call twdic_reset_poi_state. */
twdic_status twdic_icgn(twdic_ctx* ctx, const int32_t* d_poi_yx, int n,
                        float* d_u, float* d_v, float* d_p6, int32_t* d_iters,
                        uint8_t* d_flags, float* d_dpnorm);

/* twdic_icgn without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same arguments and
 * results (bit for bit) as twdic_icgn without d_dpnorm; it calls no
 * hipMalloc/hipFree.  The per-POI reference state lives in one fixed device
 * array per GPU (24 MiB, held until the process ends) instead of a context
 * buffer, so it is rebuilt on every call and nothing is cached between calls;
 * at most 65536 POIs per call (TWDIC_EINVAL beyond).  Every call on a GPU
 * shares that array: calls that overlap on one GPU (other contexts, other
 * streams) race and must be ordered by the caller; calls on one context are
 * ordered by its stream. */
twdic_status twdic_icgn_noalloc(twdic_ctx* ctx, const int32_t* d_poi_yx, int n, float* d_u, float* d_v,
                                float* d_p6, int32_t* d_iters, uint8_t* d_flags);

/* Marks the cached per-POI reference state stale (rebuilt by the next
 * twdic_icgn); the reference rebuilds it only on a reference update
 * (TW_paDIC_cuICGN2D.cu:883-903), so it has no counterpart call. */
twdic_status twdic_reset_poi_state(twdic_ctx* ctx);

/* The plain per-frame step of FFTCCTWorkerThread::processFrameFFTCC_ICGN
 * (TW/TW_Core/fftcctworkerthread.cpp:445-461): FFT-CC then IC-GN on the
 * current target, results left on the device. */
twdic_status twdic_frame(twdic_ctx* ctx, const int32_t* d_poi_yx, int n,
                         float* d_u, float* d_v, float* d_zncc, float* d_p6,
                         int32_t* d_iters, uint8_t* d_flags);

/* Graph-captured frame step (extension; the reference launches each frame's
 * kernels one by one).  twdic_frame_capture records the frame over these
 * buffers into a HIP graph the context holds: the target LUT in the LUT and
 * B-spline modes, then FFT-CC and IC-GN exactly as twdic_frame.  The
 * reference gradients and the per-POI reference state are built before
 * capture, not recorded.  twdic_frame_replay launches the graph on the
 * context stream after rebuilding that state if the reference changed; its
 * results are twdic_frame's on the current target.  New frames go in through
 * twdic_set_target, which copies into the context's own buffer.  The POI and
 * output arrays must stay allocated while the graph is held; replay returns
 * TWDIC_EINVAL when a later call has reallocated the context's per-POI state
 * (capture again then).  Capturing again replaces the graph. */
twdic_status twdic_frame_capture(twdic_ctx* ctx, const int32_t* d_poi_yx, int n,
                                 float* d_u, float* d_v, float* d_zncc, float* d_p6,
                                 int32_t* d_iters, uint8_t* d_flags);
twdic_status twdic_frame_replay(twdic_ctx* ctx);

/* ------------------------------------------------------------ POI tracking */

/* The tracking step TW_Core runs around a reference reset
 * (FFTCCTWorkerThread::processFrameFFTCC_ICGN,
 * TW/TW_Core/fftcctworkerthread.cpp:408-438,476-481).  All three are
 * asynchronous on ctx's stream; with ctx == NULL they run on the default
 * stream and complete before returning (the reference's wrappers launch on
 * the legacy default stream).  int() truncates toward zero, as the reference. */

/* Replaces cuUpdatePOIpos / updatePOIpos_Kernel (TW/TW_Core/cuda_utils.cu:
 * 174-186,271-284): d_poi_yx[i] += (int(v[i]), int(u[i])) in place.  When
 * d_poi_yx is the POI set ctx's per-POI reference state was built for, that
 * state is marked stale. */
twdic_status twdic_track_update_poi(twdic_ctx* ctx, int32_t* d_poi_yx, int n, const float* d_u, const float* d_v);
/* Replaces cuAccumulatePOI / accumulatePOI_Kernel (cuda_utils.cu:188-201,
 * 286-301): d_out_yx[i] = d_poi_yx[i] + (int(v[i]), int(u[i])), the POIs'
 * integer positions in the current target. */
twdic_status twdic_track_current_poi(twdic_ctx* ctx, const int32_t* d_poi_yx, int n, const float* d_u,
                                     const float* d_v, int32_t* d_out_yx);
/* Replaces the 7-argument cuAccumulatePOI / accumulatePOINew_Kernel
 * (cuda_utils.cu:203-223, 305-325): d_out_yx[i] = d_poi_yx[i] +
 * (int(v[i]) - int(acc_u[i]), int(u[i]) - int(acc_v[i])), the accumulated U
 * taken from y and the accumulated V from x, as the reference pairs them. */
twdic_status twdic_track_current_poi_acc(twdic_ctx* ctx, const int32_t* d_poi_yx, int n, const float* d_u,
                                         const float* d_v, const float* d_acc_u, const float* d_acc_v,
                                         int32_t* d_out_yx);
/* Replaces cuAccumulateUV / accumulateUV_Kernel (cuda_utils.cu:225-239,
 * 323-338): d_acc_u[i] += d_u[i], d_acc_v[i] += d_v[i]. */
twdic_status twdic_track_accumulate_uv(twdic_ctx* ctx, int n, const float* d_u, const float* d_v, float* d_acc_u,
                                       float* d_acc_v);

/* --------------------------------------------------------------- results */

/* Writes the per-POI result table in the 14-column CSV layout of the
 * reference's result files (TW/MATLAB/Sample 14 L1 Amp0_8.3_Data.csv, read by
 * calcustrain4_x_0_L1_5_40.m:5-6): x, y, integer u, u, du/dx, du/dy,
 * integer v, v, dv/dx, dv/dy, iterations, ZNCC, FFT-CC status, IC-GN flag;
 * one row per POI in POI order (x fastest), reals with 6 significant digits.
 * All pointers are HOST arrays of n entries (p6: n x 6 = u, ux, uy, v, vx,
 * vy; poi_yx: n x 2 = y, x).  int_u/int_v are the FFT-CC shifts.  zncc,
 * fftcc_status and flags may be NULL (written as 0). */
twdic_status twdic_write_csv(const char* path, int n, const int32_t* poi_yx, const float* int_u,
                             const float* int_v, const float* p6, const int32_t* iters, const float* zncc,
                             const uint8_t* fftcc_status, const uint8_t* flags);

/* Least-squares strain of a displacement field on the POI grid, as
 * calcustrain4_x_0_L1_5_40.m:23-33 computes it: for every grid point with a
 * full (2M+1) x (2M+1) window, d/dx = 3 / (G M (M+1) (2M+1)^2) sum ii f and
 * d/dy the same with jj (grid offsets ii along x, jj along y; G = grid step
 * in pixels), in FP64.  d_field: device float[ny][nx]; outputs device
 * double[ny-2M][nx-2M] (either may be NULL), per pixel (the script reports
 * x 1e6, microstrain).  Asynchronous on ctx's stream. */
twdic_status twdic_strain_ls(twdic_ctx* ctx, const float* d_field, int nx, int ny, int grid, int half_window,
                             double* d_dfdx, double* d_dfdy);

/* ------------------------------------------------ precompute (inspection) */

/* Device buffers the reference exposes as public members of
 * GPUHandle_ICGN (TW/TW_Engine/TW_paDIC.h:41-70): gradients of the reference
 * (Rx, Ry) and target (Tx, Ty, Txy), ROI-relative row-major fp32; the target
 * bicubic LUT, [roi_h][roi_w][16] fp32 with element 4k+l multiplying y^k x^l
 * (same order as the CPU LUT, TW_utils.cpp:673-688; the last ROI row and
 * column are zero).  Target buffers are (re)computed by twdic_precompute_target. */
enum { TWDIC_BUF_RX = 0, TWDIC_BUF_RY = 1, TWDIC_BUF_TX = 2, TWDIC_BUF_TY = 3,
       TWDIC_BUF_TXY = 4, TWDIC_BUF_LUT = 5, TWDIC_BUF_REF = 6, TWDIC_BUF_TAR = 7 };
twdic_status twdic_get_buffer(twdic_ctx* ctx, int which, void** d_ptr, size_t* bytes);
/* cuGradientXY_2Images + cuBicubicCoefficients (TW_cuUtils.cu:679-776):
 * gradients of both images and the target LUT, written to the buffers above. */
twdic_status twdic_precompute_target(twdic_ctx* ctx);

/* ---------------------------------------------------------- device memory */

/* Bytes of memory from d_ptr to the end of the HIP allocation that holds it
 * (device or pinned host memory); TWDIC_EINVAL when the runtime knows no
 * allocation there.  No counterpart in the reference.  Every entry that takes
 * device buffers checks each one this way before launching anything and
 * returns TWDIC_EBOUNDS when one is shorter than the call needs (n entries, or
 * stride x (h - 1) + w bytes of an image), TWDIC_EINVAL when one is not in a
 * HIP allocation.  The check sees the allocation, not the caller's view of
 * it: a buffer carved out of a larger allocation is held to that allocation's
 * end. */
twdic_status twdic_device_span(const void* d_ptr, size_t* bytes);

twdic_status twdic_malloc(twdic_ctx* ctx, void** d_ptr, size_t bytes);
twdic_status twdic_free(twdic_ctx* ctx, void* d_ptr);
/* kind: 0 = host to device, 1 = device to host, 2 = device to device.
 * Host transfers complete before the call returns; device-to-device copies
 * are asynchronous on the context stream. */
twdic_status twdic_memcpy(twdic_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);
twdic_status twdic_memset(twdic_ctx* ctx, void* d_ptr, int value, size_t bytes);
/* Pinned (page-locked) host memory, for frames and results that move to and
 * from the device by DMA without a staging copy. */
twdic_status twdic_host_alloc(void** h_ptr, size_t bytes);
twdic_status twdic_host_free(void* h_ptr);

/* ------------------------------------------------------------- timing */

/* Events recorded on the context stream, for device-side timing. */
twdic_status twdic_event_record(twdic_ctx* ctx, int slot);
/* Milliseconds between two recorded slots (synchronises on the later one). */
twdic_status twdic_event_elapsed(twdic_ctx* ctx, int slot_a, int slot_b, float* ms);

#ifdef __cplusplus
}
#endif
#endif
