/*
 * tw_oracle.h -- CPU restatement of TW_Engine's paDIC CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP implementation in tw_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline.  The product path never calls it.
 *
 * Every routine follows the reference's arithmetic order line by line; each
 * one cites the reference file:line it restates (paths relative to the
 * reference checkout, TW/TW_Engine/...).  Two instantiations exist:
 *   *_f32  -- real_t = float, the reference's own build (TW.h:50-58);
 *   *_f64  -- real_t = double, the full-FP64 accuracy oracle.
 *
 * Documented, intentional extensions over the reference:
 *   - the FFT-CC window edge N is a parameter (reference: N = 2*S, which this
 *     API reproduces when called with N = 2*S); the window is anchored at
 *     POI - N/2 and the peak wraps when its index exceeds N/2;
 *   - the FFT is this file's own mixed-radix complex FFT (MKL/FFTW absent);
 *   - the 6x6 solve is a textbook Cholesky (MKL LAPACKE absent);
 *   - IC-GN also returns the full p[6], the ICGN2DFlag and the last
 *     convergence norm (the reference keeps them local);
 *   - the gradient bounds check is the corrected one (the reference's check
 *     at TW_utils.cpp:246-257 admits a read one column/row past the image).
 *
 * Parity pins (which test pins which function) are listed in the header of
 * tests/test_oracle.py.
 */
#ifndef TW_ORACLE_H
#define TW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ICGN2DFlag (TW_paDIC.h:74-81) */
enum { ORC_SUCCESS = 0, ORC_DARK_SUBSET = 1, ORC_SINGULAR_HESSIAN = 2,
       ORC_SINGULAR_WARP = 3, ORC_OUT_OF_ROI = 4 };
/* FFT-CC per-POI status: the reference returns early on a flat window and
 * leaves the outputs untouched (TW_paDIC_FFTCC2D_CPU.cpp:271-275). */
enum { ORC_FFTCC_OK = 0, ORC_FFTCC_FLAT = 1 };

/* A1: Fftcc2D::recomputeNumPOI (TW_paDIC_FFTCC2D.cpp:88-94).
 * Returns 1 when both counts are > 0, else 0. */
int orc_num_poi(int roi_w, int roi_h, int sx, int sy, int gx, int gy,
                int mx, int my, int* nx, int* ny);

/* A2: ComputePOIPositions_s (TW_utils.cpp:29-49); int32[ny*nx][2] = {y, x}. */
void orc_poi_grid(int start_x, int start_y, int nx, int ny, int mx, int my,
                  int sx, int sy, int gx, int gy, int32_t* poi_yx);

/* A9: GradientXY_2Images_s (TW_utils.cpp:232-305) for one image.
 * Outputs are ROI-relative row-major [roi_h][roi_w]; gxy may be NULL.
 * Returns 0, or -1 when the ROI has no 1-pixel frame inside the image. */
int orc_gradients_f32(const uint8_t* img, int stride, int img_w, int img_h,
                      int start_x, int start_y, int roi_w, int roi_h,
                      float* gx, float* gy, float* gxy);

/* A10: BicubicCoefficients_s (TW_utils.cpp:604-700).
 * lut is [roi_h][roi_w][4][4]; lut[..][k][l] multiplies y^k x^l.  Cells in
 * the last ROI row/column other than the corner are not written (the
 * reference's buffer is calloc-zeroed there; callers pass a zeroed buffer);
 * the bottom-right corner cell is zeroed explicitly.  Returns 0 or -1. */
int orc_bicubic_lut_f32(const uint8_t* img, int stride, int img_w, int img_h,
                        const float* tx, const float* ty, const float* txy,
                        int start_x, int start_y, int roi_w, int roi_h,
                        float* lut);

/* A4-A7: Fftcc2D_CPU::ComputeFFTCC over n POIs (TW_paDIC_FFTCC2D_CPU.cpp:
 * 174-326) with window edge N.  u/v/zncc are written only for status OK.
 * status[i] gets ORC_FFTCC_OK / ORC_FFTCC_FLAT.  If gap != NULL, gap[i] =
 * (max - second max) of the unnormalised surface divided by
 * sqrt(norR*norT)*N^2, i.e. in ZNCC units (used to audit near ties).
 * Returns 0, or -1 on a bad N. */
int orc_fftcc_f32(const uint8_t* ref, const uint8_t* tar, int stride,
                  const int32_t* poi_yx, int n, int N,
                  float* u, float* v, float* zncc, uint8_t* status,
                  float* gap, int nthreads);
int orc_fftcc_f64(const uint8_t* ref, const uint8_t* tar, int stride,
                  const int32_t* poi_yx, int n, int N,
                  double* u, double* v, double* zncc, uint8_t* status,
                  double* gap, int nthreads);

/* Exact integer-pixel peak by brute force: argmax over k of the integer
 * circular correlation I[k] = sum_x R(x) T((x+k) mod N) of the two N x N
 * windows, lowest row-major index on ties, then the same wrap as the
 * reference (TW_paDIC_FFTCC2D_CPU.cpp:316-322).  The zero-mean correlation
 * the reference maximises differs from I[k] by a constant, so this is the
 * exact answer the FP32 FFT approximates.  zncc is evaluated in FP64 from
 * the integer sums.  status as above. */
int orc_fftcc_exact(const uint8_t* ref, const uint8_t* tar, int stride,
                    const int32_t* poi_yx, int n, int N,
                    int32_t* u, int32_t* v, double* zncc, uint8_t* status,
                    int nthreads);

/* A11-A15: ICGN2D_CPU::ICGN2D_Algorithm + ICGN2D_Compute, Bicubic,
 * Multicore (TW_paDIC_ICGN2D_CPU.cpp:134-195, 287-441, 481-925).
 * Computes the gradients and target LUT over the ROI, then runs every POI.
 * u, v: in = initial guess, out = result (untouched when a flag is raised
 * before the first update, as in the reference).  iters: 1-based count,
 * untouched when a flag is raised before the first update.  p6 (may be NULL):
 * final (u, ux, uy, v, vx, vy).  flags: ICGN2DFlag.  dpnorm (may be NULL):
 * last weighted |dp| evaluated by the loop test (NaN if none).
 * times_ms (may be NULL): {precompute, iterations} wall times.
 * Returns 0, or -1 on bad geometry. */
/* F2: BicubicSplineCoefficients_s (TW_utils.cpp:521-602) into a [roi_h][roi_w]
 * [16] LUT laid out as the bicubic one.  Needs a 1-pixel frame left/top and
 * a 2-pixel frame right/bottom; returns -1 otherwise. */
int orc_bspline_lut_f32(const uint8_t* img, int stride, int img_w, int img_h,
                        int start_x, int start_y, int roi_w, int roi_h, float* lut);
/* IC-GN as orc_icgn_* with the target interpolation chosen by interp:
 * 0 = Bicubic, 1 = BicubicSpline (TW_paDIC_ICGN2D_CPU.cpp:322-341). */
int orc_icgn_interp_f32(const uint8_t* ref, const uint8_t* tar, int stride,
                        int img_w, int img_h, int start_x, int start_y,
                        int roi_w, int roi_h, int sx, int sy, int max_iter,
                        float delta_p, const int32_t* poi_yx, int n,
                        float* u, float* v, float* p6, int32_t* iters,
                        uint8_t* flags, double* dpnorm, double* times_ms,
                        int nthreads, int interp);
int orc_icgn_interp_f64(const uint8_t* ref, const uint8_t* tar, int stride,
                        int img_w, int img_h, int start_x, int start_y,
                        int roi_w, int roi_h, int sx, int sy, int max_iter,
                        double delta_p, const int32_t* poi_yx, int n,
                        double* u, double* v, double* p6, int32_t* iters,
                        uint8_t* flags, double* dpnorm, double* times_ms,
                        int nthreads, int interp);
int orc_icgn_f32(const uint8_t* ref, const uint8_t* tar, int stride,
                 int img_w, int img_h, int start_x, int start_y,
                 int roi_w, int roi_h, int sx, int sy, int max_iter,
                 float delta_p, const int32_t* poi_yx, int n,
                 float* u, float* v, float* p6, int32_t* iters,
                 uint8_t* flags, double* dpnorm, double* times_ms,
                 int nthreads);
int orc_icgn_f64(const uint8_t* ref, const uint8_t* tar, int stride,
                 int img_w, int img_h, int start_x, int start_y,
                 int roi_w, int roi_h, int sx, int sy, int max_iter,
                 double delta_p, const int32_t* poi_yx, int n,
                 double* u, double* v, double* p6, int32_t* iters,
                 uint8_t* flags, double* dpnorm, double* times_ms,
                 int nthreads);

/* Building blocks exported so the tests can pin them in isolation. */
/* In-place complex FFT of length n (interleaved re/im), sign -1 forward,
 * +1 inverse (unnormalised). */
void orc_fft1d_f32(float* data, int n, int sign);
void orc_fft1d_f64(double* data, int n, int sign);
/* Textbook Cholesky of the lower triangle of a row-major 6x6 (LAPACKE
 * spotrf 'L' semantics: info > 0 when pivot info is not positive) and the
 * matching two-triangle solve (spotrs 'L'). */
int orc_potrf6_f32(float* a);
void orc_potrs6_f32(const float* l, float* b);
int orc_potrf6_f64(double* a);
void orc_potrs6_f64(const double* l, double* b);
/* One inverse-compositional update W(p) <- W(p) o W(dp)^-1 with the closed
 * form of TW_paDIC_ICGN2D_CPU.cpp:713-745; p is updated in place.  Returns 1
 * when |det W(dp)| <= epsilon (SingularWarp; p untouched), else 0. */
int orc_warp_update_f32(float* p, const float* dp);
int orc_warp_update_f64(double* p, const double* dp);
/* Bicubic evaluation of one LUT cell the way ICGN2D_Compute does
 * (TW_paDIC_ICGN2D_CPU.cpp:775-784). */
float orc_lut_eval_f32(const float* cell16, float tx, float ty);

/* F1: POI tracking around a reference reset (TW/TW_Core/cuda_utils.cu), in
 * the caller's order of use at fftcctworkerthread.cpp:419-437,476-481.
 *   update_poi   (updatePOIpos_Kernel, :174-186): poi[i] += (int(v), int(u))
 *   current_poi  (accumulatePOI_Kernel, :188-201): out[i] = poi[i] + (int(v), int(u))
 *   accumulate_uv(accumulateUV_Kernel, :225-239): acc_u += u, acc_v += v */
void orc_track_update_poi(int32_t* poi_yx, int n, const float* u, const float* v);
void orc_track_current_poi(const int32_t* poi_yx, int n, const float* u, const float* v, int32_t* out_yx);
void orc_track_accumulate_uv(int n, const float* u, const float* v, float* acc_u, float* acc_v);

/* F3: least-squares strain on the POI grid, calcustrain4_x_0_L1_5_40.m:23-33
 * (MATLAB, double): for each point with a full (2M+1)^2 window,
 * d/dx = 3/(G M (M+1) (2M+1)^2) * sum_{ii} sum_{jj} ii * f(i+ii, j+jj), ii outer;
 * d/dy the same weighted by jj.  f is [ny][nx]; outputs [ny-2M][nx-2M]. */
void orc_strain_ls(const float* f, int nx, int ny, int grid, int M, double* dfdx, double* dfdy);

#ifdef __cplusplus
}
#endif
#endif
