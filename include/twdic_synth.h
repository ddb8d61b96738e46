/*
 * twdic_synth.h -- deterministic synthetic speckle pairs for benchmarks and
 * tests.  No counterpart in the reference's engine; its analogue is the
 * offline pattern script TW/MATLAB/pattern_generation.m.  Host code only.
 *
 * Speckle: I(x, y) = clip(sum_s A_s exp(-((x - cx_s)^2 + (y - cy_s)^2) /
 * (2 sigma_s^2)), 0, 255), pixel centres at integer (col, row); spots drawn
 * from SplitMix64(seed) in order cx ~ U[0, w), cy ~ U[0, h),
 * sigma ~ U[sigma_min, sigma_max], A ~ U[amp_min, amp_max].
 * Target: the same continuous pattern seen through the forward map
 * x' = x + u(x), i.e. T(x') = I(phi^-1(x')), plus Gaussian noise, rounded to
 * the nearest integer and clipped to [0, 255].
 */
#ifndef TWDIC_SYNTH_H
#define TWDIC_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TWDIC_DEFORM_AFFINE = 0,       /* u = u0 + exx (x-xc) + exy (y-yc); v = v0 + exy (x-xc) + eyy (y-yc) */
    TWDIC_DEFORM_SINUSOID = 1,     /* u = amp_u sin(2 pi x / per_u); v = amp_v cos(2 pi y / per_v) */
    TWDIC_DEFORM_RESAMPLE = 2      /* target = bicubic (Catmull-Rom) resample of the rounded
                                      reference at (x - u0, y - v0): uniform translation */
};

typedef struct {
    int kind;
    double u0, v0, exx, eyy, exy, xc, yc;
    double amp_u, per_u, amp_v, per_v;
} twdic_deform;

typedef struct {
    int w, h;
    int nspots;
    uint64_t seed;
    double sigma_min, sigma_max, amp_min, amp_max;
    double noise_sigma;
    uint64_t noise_seed;
    /* spots whose centre lies in any box [x0, x1) x [y0, y1) are removed
     * (low-texture patches); boxes: int32[nboxes][4] = {x0, y0, x1, y1} */
    const int32_t* boxes;
    int nboxes;
} twdic_speckle;

/* Writes the reference (no noise) and the deformed target (with noise);
 * either output may be NULL.  Row stride of both outputs is `stride`.
 * Returns 0, or -1 on bad arguments. */
int twdic_synth_pair(const twdic_speckle* sp, const twdic_deform* d, uint8_t* ref, uint8_t* tar, int stride);

/* Exact displacement of the forward map at (x, y): writes u, v and the four
 * gradients (ux, uy, vx, vy) of the displacement as a function of the
 * reference position. */
void twdic_synth_displacement(const twdic_deform* d, double x, double y, double out6[6]);

#ifdef __cplusplus
}
#endif
#endif
