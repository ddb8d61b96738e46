/*
 * tw_oracle.c -- CPU restatement of TW_Engine's paDIC CPU path.
 * TEST INFRASTRUCTURE ONLY: see tw_oracle.h for the contract and for what
 * may load this library.  Built by oracle/Makefile with -ffp-contract=off
 * so that every a*b+c is rounded twice, as the reference's MSVC x64 build
 * (SSE2, no FMA contraction) does.
 */
#define _GNU_SOURCE
#include "tw_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- FP32 instantiation (real_t = float, TW.h:50-58) ------- */
#define REAL float
#define SFX(name) name##_f32
#define REAL_EPS FLT_EPSILON
#define RSQRT sqrtf
#define RABS fabsf
#include "tw_oracle_body.h"
#undef REAL
#undef SFX
#undef REAL_EPS
#undef RSQRT
#undef RABS

/* ---------------- FP64 instantiation (full-FP64 accuracy oracle) -------- */
#define REAL double
#define SFX(name) name##_f64
#define REAL_EPS DBL_EPSILON
#define RSQRT sqrt
#define RABS fabs
#include "tw_oracle_body.h"
#undef REAL
#undef SFX
#undef REAL_EPS
#undef RSQRT
#undef RABS

/* A1: TW_paDIC_FFTCC2D.cpp:88-94 (float division, floor, truncation). */
int orc_num_poi(int roi_w, int roi_h, int sx, int sy, int gx, int gy,
                int mx, int my, int* nx, int* ny)
{
    *nx = (int)floorf((float)(roi_w - sx * 2 - mx * 2) / (float)gx) + 1;
    *ny = (int)floorf((float)(roi_h - sy * 2 - my * 2) / (float)gy) + 1;
    return (*nx > 0 && *ny > 0) ? 1 : 0;
}

/* A2: TW_utils.cpp:29-49 (y first, then x). */
void orc_poi_grid(int start_x, int start_y, int nx, int ny, int mx, int my,
                  int sx, int sy, int gx, int gy, int32_t* poi_yx)
{
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            poi_yx[2 * (i * nx + j) + 0] = start_y + my + sy + i * gy;
            poi_yx[2 * (i * nx + j) + 1] = start_x + mx + sx + j * gx;
        }
}

int orc_gradients_f32(const uint8_t* img, int stride, int img_w, int img_h,
                      int start_x, int start_y, int roi_w, int roi_h,
                      float* gx, float* gy, float* gxy)
{
    return gradients_f32(img, stride, img_w, img_h, start_x, start_y, roi_w, roi_h, gx, gy, gxy);
}

int orc_bicubic_lut_f32(const uint8_t* img, int stride, int img_w, int img_h,
                        const float* tx, const float* ty, const float* txy,
                        int start_x, int start_y, int roi_w, int roi_h,
                        float* lut)
{
    return bicubic_lut_f32(img, stride, img_w, img_h, tx, ty, txy, start_x, start_y,
                           roi_w, roi_h, lut);
}

/* TW_paDIC_ICGN2D_CPU.cpp:775-784 for one cell. */
float orc_lut_eval_f32(const float* cell, float tx, float ty)
{
    float acc = 0;
    for (int k = 0; k < 4; ++k)
        for (int q = 0; q < 4; ++q)
            acc = (float)((double)acc + (double)cell[k * 4 + q] * pow((double)ty, (double)k) * pow((double)tx, (double)q));
    return acc;
}

/* Exact integer-pixel peak by brute-force integer correlation. */
int orc_fftcc_exact(const uint8_t* ref, const uint8_t* tar, int stride,
                    const int32_t* poi_yx, int n, int N,
                    int32_t* u, int32_t* v, double* zncc, uint8_t* status,
                    int nthreads)
{
    if (N < 2 || (N & 1) || N > 1024) return -1;
    const int NN = N * N, half = N / 2;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        int* r = (int*)malloc(sizeof(int) * NN);
        int* t2 = (int*)malloc(sizeof(int) * 4 * NN);   /* T tiled 2x2 for wrap-free access */
#pragma omp for schedule(dynamic, 4)
        for (int id = 0; id < n; ++id) {
            const int py = poi_yx[2 * id], px = poi_yx[2 * id + 1];
            long long sr = 0, st = 0, srr = 0, stt = 0;
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) {
                    size_t o = (size_t)(py - half + i) * stride + (px - half + j);
                    int a = ref[o], b = tar[o];
                    r[i * N + j] = a;
                    t2[(i) * 2 * N + j] = b; t2[(i) * 2 * N + j + N] = b;
                    t2[(i + N) * 2 * N + j] = b; t2[(i + N) * 2 * N + j + N] = b;
                    sr += a; st += b; srr += (long long)a * a; stt += (long long)b * b;
                }
            if (srr * NN == sr * sr || stt * NN == st * st) { status[id] = ORC_FFTCC_FLAT; continue; }
            long long best = -1; int peak = 0;
            for (int kv = 0; kv < N; ++kv)
                for (int ku = 0; ku < N; ++ku) {
                    long long acc = 0;
                    for (int i = 0; i < N; ++i) {
                        const int* rr = r + i * N;
                        const int* tt = t2 + (i + kv) * 2 * N + ku;
                        for (int j = 0; j < N; ++j) acc += (long long)rr[j] * tt[j];
                    }
                    if (acc > best) { best = acc; peak = kv * N + ku; }
                }
            int iu = peak % N, iv = peak / N;
            if (iu > half) iu -= N;
            if (iv > half) iv -= N;
            u[id] = iu; v[id] = iv;
            double dn = (double)NN;
            double num = (double)best - (double)sr * (double)st / dn;
            double den = sqrt(((double)srr - (double)sr * (double)sr / dn) *
                              ((double)stt - (double)st * (double)st / dn));
            zncc[id] = num / den;
            status[id] = ORC_FFTCC_OK;
        }
        free(r); free(t2);
    }
    return 0;
}

/* F1: TW/TW_Core/cuda_utils.cu:174-186 (updatePOIpos_Kernel): y += int(v), x += int(u). */
void orc_track_update_poi(int32_t* poi_yx, int n, const float* u, const float* v)
{
    for (int i = 0; i < n; ++i) {
        poi_yx[i * 2 + 0] += (int)v[i];
        poi_yx[i * 2 + 1] += (int)u[i];
    }
}

/* cuda_utils.cu:188-201 (accumulatePOI_Kernel). */
void orc_track_current_poi(const int32_t* poi_yx, int n, const float* u, const float* v, int32_t* out_yx)
{
    for (int i = 0; i < n; ++i) {
        out_yx[i * 2 + 0] = poi_yx[i * 2 + 0] + (int)v[i];
        out_yx[i * 2 + 1] = poi_yx[i * 2 + 1] + (int)u[i];
    }
}

/* cuda_utils.cu:203-223 (accumulatePOINew_Kernel): y takes int(v) - int(accumulated u),
 * x takes int(u) - int(accumulated v), the pairing the reference writes. */
void orc_track_current_poi_acc(const int32_t* poi_yx, int n, const float* u, const float* v, const float* acc_u,
                               const float* acc_v, int32_t* out_yx)
{
    for (int i = 0; i < n; ++i) {
        out_yx[i * 2 + 0] = poi_yx[i * 2 + 0] + (int)v[i] - (int)acc_u[i];
        out_yx[i * 2 + 1] = poi_yx[i * 2 + 1] + (int)u[i] - (int)acc_v[i];
    }
}

/* cuda_utils.cu:225-239 (accumulateUV_Kernel): the first pair accumulates. */
void orc_track_accumulate_uv(int n, const float* u, const float* v, float* acc_u, float* acc_v)
{
    for (int i = 0; i < n; ++i) {
        acc_u[i] += u[i];
        acc_v[i] += v[i];
    }
}

/* F3: TW/MATLAB/calcustrain4_x_0_L1_5_40.m:23-33. */
void orc_strain_ls(const float* f, int nx, int ny, int grid, int M, double* dfdx, double* dfdy)
{
    const int ox = nx - 2 * M, oy = ny - 2 * M;
    const double w = (double)(2 * M + 1);
    const double k = 3.0 / ((double)grid * (double)M * (double)(M + 1) * (w * w));
    for (int j = M; j < ny - M; ++j)
        for (int i = M; i < nx - M; ++i) {
            double sx = 0.0, sy = 0.0;
            for (int ii = -M; ii <= M; ++ii)
                for (int jj = -M; jj <= M; ++jj) {
                    double v = (double)f[(size_t)(j + jj) * nx + (i + ii)];
                    double a = (double)ii * v, b = (double)jj * v;
                    sx = sx + a;
                    sy = sy + b;
                }
            dfdx[(size_t)(j - M) * ox + (i - M)] = k * sx;
            dfdy[(size_t)(j - M) * ox + (i - M)] = k * sy;
        }
    (void)oy;
}
