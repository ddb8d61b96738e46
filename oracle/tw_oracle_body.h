/*
 * tw_oracle_body.h -- precision-generic body of the oracle, included twice
 * by tw_oracle.c with REAL = float and REAL = double.  TEST INFRASTRUCTURE
 * ONLY (see tw_oracle.h).
 *
 * Required macros: REAL, SFX(name), REAL_EPS, RSQRT (sqrt for REAL),
 * RABS (fabs for REAL).
 */

typedef struct { REAL re, im; } SFX(cplx);

/* ---------------------------------------------------------------------- */
/* Mixed-radix complex FFT (own implementation; MKL/FFTW are not present). */
/* Recursive decimation in time.  tw[e] = exp(sign*2*pi*i*e/N) rounded to  */
/* REAL from double.  Radix 4, then 2, 3, 5, then any prime (plain DFT).   */
/* ---------------------------------------------------------------------- */
static void SFX(fft_rec)(const SFX(cplx)* in, int is, SFX(cplx)* out, int n,
                         const SFX(cplx)* tw, int N, int tws)
{
    if (n == 1) { out[0] = in[0]; return; }
    int r;
    if (n % 4 == 0) r = 4;
    else if (n % 2 == 0) r = 2;
    else if (n % 3 == 0) r = 3;
    else if (n % 5 == 0) r = 5;
    else { r = n; for (int p = 7; p * p <= n; p += 2) if (n % p == 0) { r = p; break; } }
    int m = n / r;
    for (int j = 0; j < r; ++j)
        SFX(fft_rec)(in + (size_t)j * is, is * r, out + (size_t)j * m, m, tw, N, tws * r);
    SFX(cplx) t[64], o[64];
    for (int k = 0; k < m; ++k) {
        for (int j = 0; j < r; ++j) {
            SFX(cplx) a = out[j * m + k];
            SFX(cplx) w = tw[((long)j * k * tws) % N];
            t[j].re = a.re * w.re - a.im * w.im;
            t[j].im = a.re * w.im + a.im * w.re;
        }
        for (int q = 0; q < r; ++q) {
            REAL sr = 0, si = 0;
            for (int j = 0; j < r; ++j) {
                SFX(cplx) w = tw[((long)j * q * m * tws) % N];
                sr += t[j].re * w.re - t[j].im * w.im;
                si += t[j].re * w.im + t[j].im * w.re;
            }
            o[q].re = sr; o[q].im = si;
        }
        for (int q = 0; q < r; ++q) out[k + q * m] = o[q];
    }
}

static void SFX(make_twiddles)(SFX(cplx)* tw, int N, int sign)
{
    for (int e = 0; e < N; ++e) {
        double a = sign * 2.0 * M_PI * (double)e / (double)N;
        tw[e].re = (REAL)cos(a);
        tw[e].im = (REAL)sin(a);
    }
}

void SFX(orc_fft1d)(REAL* data, int n, int sign)
{
    SFX(cplx)* tw = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* in = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(make_twiddles)(tw, n, sign);
    memcpy(in, data, sizeof(SFX(cplx)) * n);
    SFX(fft_rec)(in, 1, (SFX(cplx)*)data, n, tw, n, 1);
    free(tw); free(in);
}

/* 2-D transform of an N x N row-major complex array, rows then columns. */
static void SFX(fft2d)(SFX(cplx)* a, int N, const SFX(cplx)* tw,
                       SFX(cplx)* line_in, SFX(cplx)* line_out)
{
    for (int y = 0; y < N; ++y) {
        memcpy(line_in, a + (size_t)y * N, sizeof(SFX(cplx)) * N);
        SFX(fft_rec)(line_in, 1, a + (size_t)y * N, N, tw, N, 1);
    }
    for (int x = 0; x < N; ++x) {
        for (int y = 0; y < N; ++y) line_in[y] = a[(size_t)y * N + x];
        SFX(fft_rec)(line_in, 1, line_out, N, tw, N, 1);
        for (int y = 0; y < N; ++y) a[(size_t)y * N + x] = line_out[y];
    }
}

/* ---------------------------------------------------------------------- */
/* A4-A7: FFT-CC per POI, Fftcc2D_CPU::ComputeFFTCC                       */
/* (TW_paDIC_FFTCC2D_CPU.cpp:222-326), window edge N, anchor POI - N/2.   */
/* ---------------------------------------------------------------------- */
int SFX(orc_fftcc)(const uint8_t* ref, const uint8_t* tar, int stride,
                   const int32_t* poi_yx, int n, int N,
                   REAL* u, REAL* v, REAL* zncc, uint8_t* status,
                   REAL* gap, int nthreads)
{
    if (N < 2 || (N & 1) || N > 1024) return -1;
    const int NN = N * N, half = N / 2;
    SFX(cplx)* twf = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * N);
    SFX(cplx)* twi = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * N);
    SFX(make_twiddles)(twf, N, -1);
    SFX(make_twiddles)(twi, N, +1);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        SFX(cplx)* f1 = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * NN);
        SFX(cplx)* f2 = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * NN);
        REAL* s1 = (REAL*)malloc(sizeof(REAL) * NN);
        REAL* s2 = (REAL*)malloc(sizeof(REAL) * NN);
        SFX(cplx)* li = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * N);
        SFX(cplx)* lo = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * N);
#pragma omp for schedule(dynamic, 16)
        for (int id = 0; id < n; ++id) {
            const int py = poi_yx[2 * id], px = poi_yx[2 * id + 1];
            REAL aveR = 0, aveT = 0;
            /* :239-252 fill + sums (sums of integers, exact in float). */
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) {
                    size_t o = (size_t)(py - half + i) * stride + (px - half + j);
                    REAL t;
                    t = s1[i * N + j] = (REAL)ref[o]; aveR += t;
                    t = s2[i * N + j] = (REAL)tar[o]; aveT += t;
                }
            aveR = aveR / (REAL)NN;                                 /* :253 */
            aveT = aveT / (REAL)NN;                                 /* :254 */
            REAL norR = 0, norT = 0;
            for (int k = 0; k < NN; ++k) {                          /* :259-270 */
                REAL t;
                t = s1[k] -= aveR; norR += t * t;
                t = s2[k] -= aveT; norT += t * t;
            }
            if ((double)norR < 0.0000001 || (double)norT < 0.0000001) { /* :272-275 */
                status[id] = ORC_FFTCC_FLAT;
                continue;
            }
            for (int k = 0; k < NN; ++k) {
                f1[k].re = s1[k]; f1[k].im = 0;
                f2[k].re = s2[k]; f2[k].im = 0;
            }
            SFX(fft2d)(f1, N, twf, li, lo);                         /* :282 */
            SFX(fft2d)(f2, N, twf, li, lo);                         /* :283 */
            for (int p = 0; p < NN; ++p) {                          /* :285-293 */
                SFX(cplx) a = f1[p], b = f2[p], c;
                c.re = (a.re * b.re) + (a.im * b.im);
                c.im = (a.re * b.im) - (a.im * b.re);
                f1[p] = c;
            }
            SFX(fft2d)(f1, N, twi, li, lo);                         /* :297 */
            REAL best = -2, second = -2;                            /* :300 */
            int peak = 0;
            for (int k = 0; k < NN; ++k) {                          /* :304-311 */
                REAL c = f1[k].re;
                if (best < c) { second = best; best = c; peak = k; }
                else if (second < c) second = c;
            }
            REAL scale = RSQRT(norR * norT) * (REAL)NN;             /* :313 */
            REAL zn = best / scale;
            int iu = peak % N, iv = peak / N;                       /* :316-322 */
            if (iu > half) iu -= N;
            if (iv > half) iv -= N;
            u[id] = (REAL)iu; v[id] = (REAL)iv; zncc[id] = zn;
            status[id] = ORC_FFTCC_OK;
            if (gap) gap[id] = (best - second) / scale;
        }
        free(f1); free(f2); free(s1); free(s2); free(li); free(lo);
    }
    free(twf); free(twi);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Cholesky (LAPACKE_spotrf/spotrs 'L', row major, n = 6;                 */
/* TW_paDIC_ICGN2D_CPU.cpp:682-683, 853).  Textbook left-looking form.    */
/* ---------------------------------------------------------------------- */
int SFX(orc_potrf6)(REAL* a)
{
    for (int j = 0; j < 6; ++j) {
        REAL d = a[j * 6 + j];
        for (int k = 0; k < j; ++k) d -= a[j * 6 + k] * a[j * 6 + k];
        if (!(d > 0)) return j + 1;
        d = RSQRT(d);
        a[j * 6 + j] = d;
        for (int i = j + 1; i < 6; ++i) {
            REAL s = a[i * 6 + j];
            for (int k = 0; k < j; ++k) s -= a[i * 6 + k] * a[j * 6 + k];
            a[i * 6 + j] = s / d;
        }
    }
    return 0;
}

void SFX(orc_potrs6)(const REAL* l, REAL* b)
{
    for (int i = 0; i < 6; ++i) {            /* L y = b */
        REAL s = b[i];
        for (int k = 0; k < i; ++k) s -= l[i * 6 + k] * b[k];
        b[i] = s / l[i * 6 + i];
    }
    for (int i = 5; i >= 0; --i) {           /* L^T x = y */
        REAL s = b[i];
        for (int k = i + 1; k < 6; ++k) s -= l[k * 6 + i] * b[k];
        b[i] = s / l[i * 6 + i];
    }
}

/* ---------------------------------------------------------------------- */
/* A9/A10 precompute in REAL.                                             */
/* ---------------------------------------------------------------------- */
static int SFX(gradients)(const uint8_t* img, int stride, int img_w, int img_h,
                          int sx0, int sy0, int rw, int rh,
                          REAL* gx, REAL* gy, REAL* gxy)
{
    /* TW_utils.cpp:246-257 with the off-by-one corrected: a 1-pixel frame
     * around the ROI must lie inside the image. */
    if (sx0 < 1 || sy0 < 1 || sx0 + rw + 1 > img_w || sy0 + rh + 1 > img_h) return -1;
#define PIX(yy, xx) ((int)img[(size_t)(yy) * stride + (xx)])
    /* rows in parallel, as GradientXY_2Images_m (TW_utils.cpp:445-519);
     * every element is computed on its own, so the output is the same */
#pragma omp parallel for schedule(static)
    for (int i = sy0; i < sy0 + rh; ++i)                    /* :260-273 */
        for (int j = sx0; j < sx0 + rw; ++j) {
            size_t o = (size_t)(i - sy0) * rw + (j - sx0);
            gx[o] = (REAL)(0.5 * (REAL)(PIX(i, j + 1) - PIX(i, j - 1)));
            gy[o] = (REAL)(0.5 * (REAL)(PIX(i + 1, j) - PIX(i - 1, j)));
            if (gxy)
                gxy[o] = (REAL)(0.25 * (REAL)(PIX(i + 1, j + 1) - PIX(i - 1, j + 1)
                                               - PIX(i + 1, j - 1) + PIX(i - 1, j - 1)));
        }
#undef PIX
    return 0;
}

static const signed char SFX(bicubic_matrix)[16][16] = {   /* TW_utils.cpp:622-638 */
    {  1,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0 },
    {  0,  0,  0,  0,  1,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0 },
    { -3,  3,  0,  0, -2, -1,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0 },
    {  2, -2,  0,  0,  1,  1,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0 },
    {  0,  0,  0,  0,  0,  0,  0,  0,  1,  0,  0,  0,  0,  0,  0,  0 },
    {  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  1,  0,  0,  0 },
    {  0,  0,  0,  0,  0,  0,  0,  0, -3,  3,  0,  0, -2, -1,  0,  0 },
    {  0,  0,  0,  0,  0,  0,  0,  0,  2, -2,  0,  0,  1,  1,  0,  0 },
    { -3,  0,  3,  0,  0,  0,  0,  0, -2,  0, -1,  0,  0,  0,  0,  0 },
    {  0,  0,  0,  0, -3,  0,  3,  0,  0,  0,  0,  0, -2,  0, -1,  0 },
    {  9, -9, -9,  9,  6,  3, -6, -3,  6, -6,  3, -3,  4,  2,  2,  1 },
    { -6,  6,  6, -6, -3, -3,  3,  3, -4,  4, -2,  2, -2, -2, -1, -1 },
    {  2,  0, -2,  0,  0,  0,  0,  0,  1,  0,  1,  0,  0,  0,  0,  0 },
    {  0,  0,  0,  0,  2,  0, -2,  0,  0,  0,  0,  0,  1,  0,  1,  0 },
    { -6,  6,  6, -6, -4, -2,  4,  2, -3,  3, -3,  3, -2, -1, -2, -1 },
    {  4, -4, -4,  4,  2,  2, -2, -2,  2, -2,  2, -2,  1,  1,  1,  1 }};

static int SFX(bicubic_lut)(const uint8_t* img, int stride, int img_w, int img_h,
                            const REAL* tx, const REAL* ty, const REAL* txy,
                            int sx0, int sy0, int rw, int rh, REAL* lut)
{
    /* TW_utils.cpp:615-619 (the reference passes (cols, cols) as (W, H),
     * TW_paDIC_ICGN2D_CPU.cpp:317; the check below uses the true height). */
    if (img_h - (rh + sy0 + 1) < 0 || img_w - (rw + sx0 + 1) < 0) return -1;
    /* rows in parallel, as BicubicCoefficients_m (TW_utils.cpp:789-888) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rh - 1; ++i)                        /* :643-690 */
        for (int j = 0; j < rw - 1; ++j) {
            REAL tao[16], alpha[16];
            tao[0] = (REAL)img[(size_t)(i + sy0) * stride + j + sx0];
            tao[1] = (REAL)img[(size_t)(i + sy0) * stride + j + sx0 + 1];
            tao[2] = (REAL)img[(size_t)(i + sy0 + 1) * stride + j + sx0];
            tao[3] = (REAL)img[(size_t)(i + sy0 + 1) * stride + j + sx0 + 1];
            tao[4] = tx[(size_t)i * rw + j];       tao[5] = tx[(size_t)i * rw + j + 1];
            tao[6] = tx[(size_t)(i + 1) * rw + j]; tao[7] = tx[(size_t)(i + 1) * rw + j + 1];
            tao[8] = ty[(size_t)i * rw + j];       tao[9] = ty[(size_t)i * rw + j + 1];
            tao[10] = ty[(size_t)(i + 1) * rw + j]; tao[11] = ty[(size_t)(i + 1) * rw + j + 1];
            tao[12] = txy[(size_t)i * rw + j];      tao[13] = txy[(size_t)i * rw + j + 1];
            tao[14] = txy[(size_t)(i + 1) * rw + j]; tao[15] = txy[(size_t)(i + 1) * rw + j + 1];
            for (int k = 0; k < 16; ++k) {
                alpha[k] = 0;
                for (int l = 0; l < 16; ++l)
                    alpha[k] += (REAL)SFX(bicubic_matrix)[k][l] * tao[l];
            }
            memcpy(lut + ((size_t)i * rw + j) * 16, alpha, sizeof(alpha));
        }
    memset(lut + ((size_t)(rh - 1) * rw + (rw - 1)) * 16, 0, 16 * sizeof(REAL)); /* :692-699 */
    return 0;
}

/* F2: BicubicSplineCoefficients_s (TW_utils.cpp:521-602): per ROI pixel,
 * beta = CP . Omega^T . CP^T from the 4x4 neighbourhood Omega (rows y-1..y+2,
 * columns x-1..x+2), then B . beta^T . B^T, stored rotated by 180 degrees
 * (element 4k+l multiplies y^k x^l, as the bicubic LUT).  Every ROI pixel,
 * the last row and column included, gets coefficients.  Returns -1 when the
 * neighbourhood of the last row/column would leave the image (the
 * reference's check at :529-533 admits one row/column past it). */
static const double SFX(bspline_cp)[4][4] = {
    {  71 / 56.0, -19 / 56.0,   5 / 56.0,  -1 / 56.0 },
    { -19 / 56.0,  95 / 56.0, -25 / 56.0,   5 / 56.0 },
    {   5 / 56.0, -25 / 56.0,  95 / 56.0, -19 / 56.0 },
    {  -1 / 56.0,   5 / 56.0, -19 / 56.0,  71 / 56.0 }};
static const double SFX(bspline_base)[4][4] = {
    { -1 / 6.0,  3 / 6.0, -3 / 6.0, 1 / 6.0 },
    {  3 / 6.0, -6 / 6.0,  3 / 6.0,       0 },
    { -3 / 6.0,        0,  3 / 6.0,       0 },
    {  1 / 6.0,  4 / 6.0,  1 / 6.0,       0 }};

static int SFX(bspline_lut)(const uint8_t* img, int stride, int img_w, int img_h,
                            int sx0, int sy0, int rw, int rh, REAL* lut)
{
    if (sx0 < 1 || sy0 < 1 || img_h - (rh + sy0 + 2) < 0 || img_w - (rw + sx0 + 2) < 0) return -1;
    REAL cp[4][4], bb[4][4];
    for (int k = 0; k < 4; ++k)
        for (int l = 0; l < 4; ++l) {
            cp[k][l] = (REAL)SFX(bspline_cp)[k][l];
            bb[k][l] = (REAL)SFX(bspline_base)[k][l];
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rh; ++i)
        for (int j = 0; j < rw; ++j) {
            REAL om[4][4], beta[4][4], c[4][4];
            for (int k = 0; k < 4; ++k)
                for (int l = 0; l < 4; ++l)
                    om[k][l] = (REAL)img[(size_t)(i + sy0 - 1 + k) * stride + (j + sx0 - 1 + l)];
            for (int k = 0; k < 4; ++k)
                for (int l = 0; l < 4; ++l) {
                    beta[k][l] = 0;
                    for (int m = 0; m < 4; ++m)
                        for (int n = 0; n < 4; ++n) beta[k][l] += cp[k][m] * cp[l][n] * om[n][m];
                }
            for (int k = 0; k < 4; ++k)
                for (int l = 0; l < 4; ++l) {
                    c[k][l] = 0;
                    for (int m = 0; m < 4; ++m)
                        for (int n = 0; n < 4; ++n) c[k][l] += bb[k][m] * bb[l][n] * beta[n][m];
                }
            for (int k = 0; k < 2; ++k)                     /* :591-599 */
                for (int l = 0; l < 4; ++l) {
                    REAL t = c[k][l];
                    c[k][l] = c[3 - k][3 - l];
                    c[3 - k][3 - l] = t;
                }
            REAL* o = lut + ((size_t)i * rw + j) * 16;
            for (int k = 0; k < 4; ++k)
                for (int l = 0; l < 4; ++l) o[4 * k + l] = c[k][l];
        }
    return 0;
}

int SFX(orc_bspline_lut)(const uint8_t* img, int stride, int img_w, int img_h,
                         int start_x, int start_y, int roi_w, int roi_h, REAL* lut)
{
    return SFX(bspline_lut)(img, stride, img_w, img_h, start_x, start_y, roi_w, roi_h, lut);
}

/* ---------------------------------------------------------------------- */
/* A14: ICGN2D_CPU::ICGN2D_Compute (TW_paDIC_ICGN2D_CPU.cpp:481-925).     */
/* ---------------------------------------------------------------------- */
typedef struct {
    const uint8_t* ref; const uint8_t* tar; int stride;
    int sx0, sy0, rw, rh, Sx, Sy, W, H, s, max_iter;
    REAL delta_p;
    const REAL* rx; const REAL* ry; const REAL* lut;
} SFX(icgn_ctx);

static void SFX(warp_update)(REAL* P, const REAL* dp, REAL W[2][3], REAL den)
{
    /* :718-723 and :880-885, then :740-745 */
    W[0][0] = ((1 + P[1]) * (1 + dp[5]) - P[2] * dp[4]) / den;
    W[0][1] = (P[2] * (1 + dp[1]) - (1 + P[1]) * dp[2]) / den;
    W[0][2] = P[0] + (P[2] * (dp[0] * dp[4] - dp[3] - dp[3] * dp[1]) - (1 + P[1]) * (dp[0] * dp[5] + dp[0] - dp[2] * dp[3])) / den;
    W[1][0] = (P[4] * (1 + dp[5]) - (1 + P[5]) * dp[4]) / den;
    W[1][1] = ((1 + P[5]) * (1 + dp[1]) - P[4] * dp[2]) / den;
    W[1][2] = P[3] + ((1 + P[5]) * (dp[0] * dp[4] - dp[3] - dp[3] * dp[1]) - P[4] * (dp[0] * dp[5] + dp[0] - dp[2] * dp[3])) / den;
    P[0] = W[0][2];
    P[1] = W[0][0] - 1;
    P[2] = W[0][1];
    P[3] = W[1][2];
    P[4] = W[1][0];
    P[5] = W[1][1] - 1;
}

int SFX(orc_warp_update)(REAL* p, const REAL* dp)
{
    REAL W[2][3];
    REAL den = (1 + dp[1]) * (1 + dp[5]) - dp[2] * dp[4];
    if (RABS(den) <= REAL_EPS) return 1;
    SFX(warp_update)(p, dp, W, den);
    return 0;
}

static double SFX(conv_norm)(const REAL* dp, int Sx, int Sy)
{
    /* :751 -- pow(real_t, int) promotes to double under ISO C++11. */
    return sqrt(pow((double)dp[0], 2) + pow((double)(REAL)(dp[1] * (REAL)Sx), 2)
                + pow((double)(REAL)(dp[2] * (REAL)Sy), 2) + pow((double)dp[3], 2)
                + pow((double)(REAL)(dp[4] * (REAL)Sx), 2) + pow((double)(REAL)(dp[5] * (REAL)Sy), 2));
}

static int SFX(icgn_poi)(const SFX(icgn_ctx)* c, REAL* fU, REAL* fV, int32_t* iters,
                         int px, int py, REAL* subR, REAL* subT, REAL* sd,
                         REAL Pout[6], double* dpn)
{
    const int W = c->W, H = c->H, Sx = c->Sx, Sy = c->Sy, s = c->s;
    REAL P[6] = {0, 0, 0, 0, 0, 0}, RHS[6] = {0, 0, 0, 0, 0, 0}, Hs[36];
    REAL Wm[2][3];
    REAL meanR = 0, normR = 0, meanT = 0, normT = 0;
    memset(Hs, 0, sizeof(Hs));
    *dpn = NAN;
    P[0] = *fU; P[3] = *fV;
    memcpy(Pout, P, sizeof(P));
    /* Extension: a subset that leaves the ROI would make the reference read
     * outside its gradient arrays (:543-544); report OutofROI instead. */
    if (px - Sx < c->sx0 || py - Sy < c->sy0 || px + Sx >= c->sx0 + c->rw || py + Sy >= c->sy0 + c->rh)
        return ORC_OUT_OF_ROI;
    for (int l = 0; l < H; ++l)                                  /* :516-564 */
        for (int m = 0; m < W; ++m) {
            int idY = py - Sy + l, idX = px - Sx + m;
            REAL* sdp = sd + (size_t)(l * W + m) * 6;
            subR[l * W + m] = (REAL)c->ref[(size_t)idY * c->stride + idX];
            meanR += subR[l * W + m];
            REAL J0[6] = {1, (REAL)(m - Sx), (REAL)(l - Sy), 0, 0, 0};
            REAL J1[6] = {0, 0, 0, 1, (REAL)(m - Sx), (REAL)(l - Sy)};
            REAL gxv = c->rx[(size_t)(idY - c->sy0) * c->rw + (idX - c->sx0)];
            REAL gyv = c->ry[(size_t)(idY - c->sy0) * c->rw + (idX - c->sx0)];
            for (int k = 0; k < 6; ++k) sdp[k] = gxv * J0[k] + gyv * J1[k];
            for (int k = 0; k < 6; ++k)
                for (int q = 0; q <= k; ++q) Hs[k * 6 + q] += sdp[k] * sdp[q];
        }
    meanR /= (REAL)s;                                            /* :580-582 */
    if (RABS(meanR) <= REAL_EPS) return ORC_DARK_SUBSET;
    for (int k = 0; k < s; ++k) {                                /* :585-592 */
        subR[k] = subR[k] - meanR;
        normR += subR[k] * subR[k];
    }
    normR = RSQRT(normR);                                        /* :593-595 */
    if (RABS(normR) <= REAL_EPS) return ORC_DARK_SUBSET;

    Wm[0][0] = 1 + P[1]; Wm[0][1] = P[2];     Wm[0][2] = P[0];   /* :605-607 */
    Wm[1][0] = P[4];     Wm[1][1] = 1 + P[5]; Wm[1][2] = P[3];
    for (int l = 0; l < H; ++l)                                  /* :610-637 */
        for (int m = 0; m < W; ++m) {
            REAL wx = (REAL)px + Wm[0][0] * (REAL)(m - Sx) + Wm[0][1] * (REAL)(l - Sy) + Wm[0][2];
            REAL wy = (REAL)py + Wm[1][0] * (REAL)(m - Sx) + Wm[1][1] * (REAL)(l - Sy) + Wm[1][2];
            int ix = (int)wx, iy = (int)wy;
            if (ix >= c->sx0 && iy >= c->sy0 && ix < c->sx0 + c->rw && iy < c->sy0 + c->rh) {
                subT[l * W + m] = (REAL)c->tar[(size_t)iy * c->stride + ix];
                meanT += subT[l * W + m];
            } else return ORC_OUT_OF_ROI;
        }
    meanT /= (REAL)s;                                            /* :638-640 */
    if (RABS(meanT) <= REAL_EPS) return ORC_DARK_SUBSET;
    for (int k = 0; k < s; ++k) {                                /* :643-650 */
        subT[k] = subT[k] - meanT;
        normT += subT[k] * subT[k];
    }
    normT = RSQRT(normT);                                        /* :651-653 */
    if (RABS(normT) <= REAL_EPS) return ORC_DARK_SUBSET;
    for (int k = 0; k < s; ++k) {                                /* :657-667 */
        REAL err = (normR / normT) * subT[k] - subR[k];
        for (int q = 0; q < 6; ++q) RHS[q] += (sd[(size_t)k * 6 + q] * err);
    }
    int info = SFX(orc_potrf6)(Hs);                              /* :682 */
    SFX(orc_potrs6)(Hs, RHS);                                    /* :683 */
    if (info > 0) return ORC_SINGULAR_HESSIAN;                   /* :688-693 */
    REAL den = (1 + RHS[1]) * (1 + RHS[5]) - RHS[2] * RHS[4];    /* :713-715 */
    if (RABS(den) <= REAL_EPS) return ORC_SINGULAR_WARP;
    SFX(warp_update)(P, RHS, Wm, den);
    *fU = P[0]; *fV = P[3];
    memcpy(Pout, P, sizeof(P));
    *dpn = SFX(conv_norm)(RHS, Sx, Sy);
    *iters = 1;                                                  /* :749 */
    /* :750-751; the weighted norm of each dp is kept in *dpn for audits. */
    while (*iters < c->max_iter && *dpn >= (double)c->delta_p) {
        ++*iters;                                                /* :753 */
        meanT = 0; normT = 0;
        for (int l = 0; l < H; ++l)                              /* :757-792 */
            for (int m = 0; m < W; ++m) {
                REAL wx = (REAL)px + Wm[0][0] * (REAL)(m - Sx) + Wm[0][1] * (REAL)(l - Sy) + Wm[0][2];
                REAL wy = (REAL)py + Wm[1][0] * (REAL)(m - Sx) + Wm[1][1] * (REAL)(l - Sy) + Wm[1][2];
                int ix = (int)wx, iy = (int)wy;
                if (ix >= c->sx0 && iy >= c->sy0 && ix < c->sx0 + c->rw && iy < c->sy0 + c->rh) {
                    REAL ftx = wx - (REAL)ix, fty = wy - (REAL)iy;
                    const REAL* cell = c->lut + ((size_t)(iy - c->sy0) * c->rw + (ix - c->sx0)) * 16;
                    REAL acc = 0;
                    for (int k = 0; k < 4; ++k)
                        for (int q = 0; q < 4; ++q)
                            acc = (REAL)((double)acc + (double)cell[k * 4 + q] * pow((double)fty, (double)k) * pow((double)ftx, (double)q));
                    subT[l * W + m] = acc;
                    meanT += acc;
                } else return ORC_OUT_OF_ROI;
            }
        meanT /= (REAL)s;                                        /* :793-795 */
        if (RABS(meanT) <= REAL_EPS) return ORC_DARK_SUBSET;
        for (int k = 0; k < s; ++k) {                            /* :798-805 */
            subT[k] = subT[k] - meanT;
            normT += subT[k] * subT[k];
        }
        normT = RSQRT(normT);                                    /* :806-808 */
        if (RABS(normT) <= REAL_EPS) return ORC_DARK_SUBSET;
        for (int q = 0; q < 6; ++q) RHS[q] = 0;                  /* :811-826 */
        for (int k = 0; k < s; ++k) {
            REAL err = (normR / normT) * subT[k];
            err -= subR[k];
            for (int q = 0; q < 6; ++q) RHS[q] += (sd[(size_t)k * 6 + q] * err);
        }
        SFX(orc_potrs6)(Hs, RHS);                                /* :853 */
        den = (1 + RHS[1]) * (1 + RHS[5]) - RHS[2] * RHS[4];     /* :875-877 */
        if (RABS(den) <= REAL_EPS) return ORC_SINGULAR_WARP;
        SFX(warp_update)(P, RHS, Wm, den);                       /* :880-907 */
        *fU = P[0]; *fV = P[3];
        memcpy(Pout, P, sizeof(P));
        *dpn = SFX(conv_norm)(RHS, Sx, Sy);
    }
    return ORC_SUCCESS;
}

int SFX(orc_icgn_interp)(const uint8_t* ref, const uint8_t* tar, int stride,
                         int img_w, int img_h, int start_x, int start_y,
                         int roi_w, int roi_h, int sx, int sy, int max_iter,
                         REAL delta_p, const int32_t* poi_yx, int n,
                         REAL* u, REAL* v, REAL* p6, int32_t* iters,
                         uint8_t* flags, double* dpnorm, double* times_ms,
                         int nthreads, int interp);

int SFX(orc_icgn)(const uint8_t* ref, const uint8_t* tar, int stride,
                  int img_w, int img_h, int start_x, int start_y,
                  int roi_w, int roi_h, int sx, int sy, int max_iter,
                  REAL delta_p, const int32_t* poi_yx, int n,
                  REAL* u, REAL* v, REAL* p6, int32_t* iters,
                  uint8_t* flags, double* dpnorm, double* times_ms,
                  int nthreads)
{
    return SFX(orc_icgn_interp)(ref, tar, stride, img_w, img_h, start_x, start_y, roi_w, roi_h, sx, sy,
                                max_iter, delta_p, poi_yx, n, u, v, p6, iters, flags, dpnorm, times_ms,
                                nthreads, 0);
}

/* interp: 0 = Bicubic (TW_paDIC_ICGN2D_CPU.cpp:287-318), 1 = BicubicSpline
 * (:322-341): same reference gradients, the target LUT from
 * BicubicSplineCoefficients_s. */
int SFX(orc_icgn_interp)(const uint8_t* ref, const uint8_t* tar, int stride,
                         int img_w, int img_h, int start_x, int start_y,
                         int roi_w, int roi_h, int sx, int sy, int max_iter,
                         REAL delta_p, const int32_t* poi_yx, int n,
                         REAL* u, REAL* v, REAL* p6, int32_t* iters,
                         uint8_t* flags, double* dpnorm, double* times_ms,
                         int nthreads, int interp)
{
    if (roi_w < 2 || roi_h < 2 || sx < 0 || sy < 0) return -1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    size_t P = (size_t)roi_w * roi_h;
    REAL* rx = (REAL*)calloc(P, sizeof(REAL));
    REAL* ry = (REAL*)calloc(P, sizeof(REAL));
    REAL* tx = (REAL*)calloc(P, sizeof(REAL));
    REAL* ty = (REAL*)calloc(P, sizeof(REAL));
    REAL* txy = (REAL*)calloc(P, sizeof(REAL));
    REAL* lut = (REAL*)calloc(P * 16, sizeof(REAL));
    int rc = 0;
    double t0 = omp_get_wtime();
    if (SFX(gradients)(ref, stride, img_w, img_h, start_x, start_y, roi_w, roi_h, rx, ry, NULL))
        rc = -1;
    else if (interp == 1)
        rc = SFX(bspline_lut)(tar, stride, img_w, img_h, start_x, start_y, roi_w, roi_h, lut) ? -1 : 0;
    else if (SFX(gradients)(tar, stride, img_w, img_h, start_x, start_y, roi_w, roi_h, tx, ty, txy) ||
             SFX(bicubic_lut)(tar, stride, img_w, img_h, tx, ty, txy, start_x, start_y, roi_w, roi_h, lut))
        rc = -1;
    double t1 = omp_get_wtime();
    if (rc == 0) {
        SFX(icgn_ctx) c;
        c.ref = ref; c.tar = tar; c.stride = stride;
        c.sx0 = start_x; c.sy0 = start_y; c.rw = roi_w; c.rh = roi_h;
        c.Sx = sx; c.Sy = sy; c.W = 2 * sx + 1; c.H = 2 * sy + 1; c.s = c.W * c.H;
        c.max_iter = max_iter; c.delta_p = delta_p;
        c.rx = rx; c.ry = ry; c.lut = lut;
#pragma omp parallel
        {
            REAL* subR = (REAL*)malloc(sizeof(REAL) * c.s);
            REAL* subT = (REAL*)malloc(sizeof(REAL) * c.s);
            REAL* sd = (REAL*)malloc(sizeof(REAL) * c.s * 6);
#pragma omp for schedule(dynamic, 16)
            for (int i = 0; i < n; ++i) {
                REAL pout[6];
                double dn;
                flags[i] = (uint8_t)SFX(icgn_poi)(&c, &u[i], &v[i], &iters[i],
                                                  poi_yx[2 * i + 1], poi_yx[2 * i],
                                                  subR, subT, sd, pout, &dn);
                if (p6) memcpy(p6 + (size_t)i * 6, pout, sizeof(pout));
                if (dpnorm) dpnorm[i] = dn;
            }
            free(subR); free(subT); free(sd);
        }
    }
    double t2 = omp_get_wtime();
    if (times_ms) { times_ms[0] = 1e3 * (t1 - t0); times_ms[1] = 1e3 * (t2 - t1); }
    free(rx); free(ry); free(tx); free(ty); free(txy); free(lut);
    return rc;
}
