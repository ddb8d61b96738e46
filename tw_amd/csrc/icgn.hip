// icgn.hip -- first-order IC-GN sub-pixel registration for gfx950: one POI
// per 32-lane half of a wavefront, eight per 256-thread workgroup, for large
// POI sets; one POI per wavefront, four per workgroup, for small ones
// (icgn_lanes_for).
//
// Semantics follow ICGN2D_CPU::ICGN2D_Compute (TW_paDIC_ICGN2D_CPU.cpp:
// 481-925): first pass at integer target pixels, later passes bicubic; flags
// DarkSubset / OutofROI / SingularHessian / SingularWarp in the same order;
// 1-based iteration count; W <- W(p) o W(dp)^-1 by the same closed form
// (:713-745) in FP32 without contraction; the loop test (:750-751) on the
// FP32 dp with the norm evaluated in FP64 as the reference's pow() does.
//
// Per-iteration work stays in registers:
//   * each lane of a POI's group of L lanes walks pixels l, l+L, ...
//     (l = lane % L) of the (2Sy+1)x(2Sx+1) subset;
//   * one pass forms sum T, sum T^2 and the six sums SD_k T in FP64 (every
//     FP32 product SD_k T is exact in FP64), so the residual vector is
//       RHS_k = (R_s/T_s) (sum SD_k T - mean(T) A_k) - B_k
//     with A_k = sum SD_k and B_k = sum SD_k (R - mean R) precomputed per
//     reference (k_ref_state).  This equals the reference's three passes
//     (:793-826) algebraically and is more accurate;
//   * butterflies within the group reduce the 8 sums (its lanes end
//     bit-identical);
//   * the 6x6 solve uses the per-POI FP64 Cholesky factor.
// With half-wavefronts the per-POI work between passes (reductions, solve,
// warp update, loop test) is one instruction stream for the two POIs of a
// wavefront, issued once for both; a half whose POI has finished idles until
// the other's has.  That halves the cost of that work when the POIs fill the
// GPU many times over; with few POIs a lone wavefront's latency per pass
// binds, and a whole wavefront per POI takes half the pixel rounds.
#include "twdic_internal.h"

#include <cfloat>

namespace twdic {

namespace {

template <int O>
__device__ __forceinline__ double add_lane_xor(double v)
{
    // the value of lane ^ O: ds_swizzle's bit mode (and-mask 0x1f, or-mask 0,
    // xor-mask O) within 32-lane groups, one dword at a time
    constexpr int pattern = 0x1f | (O << 10);
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), pattern);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), pattern);
    return v + __hiloint2double(hi, lo);
}

// Sum over the L lanes of a POI's group (a wavefront or a half); every
// lane of the group ends with the same bits.
template <int L>
__device__ __forceinline__ double group_sum(double v)
{
    static_assert(L == 64 || L == 32, "lane groups of 64 or 32");
    if constexpr (L == 64) v += __shfl_xor(v, 32);
    v = add_lane_xor<16>(v);
    v = add_lane_xor<8>(v);
    v = add_lane_xor<4>(v);
    v = add_lane_xor<2>(v);
    return add_lane_xor<1>(v);
}

// True when pred holds in any active lane of the caller's group.
template <int L>
__device__ __forceinline__ bool group_any(bool pred, int lane)
{
    const unsigned long long b = __ballot(pred);
    if constexpr (L == 64) return b != 0ull;
    return (uint32_t)((lane & 32) ? (b >> 32) : b) != 0u;
}

// Warp coordinate exactly as the CPU path evaluates it (:618-619, :761-762):
// ((POI + W0*dx) + W1*dy) + W2, every product and sum rounded to FP32 (the
// pragma keeps the compiler from fusing a product into the following sum).
__device__ __forceinline__ float warp_coord(float poi, float w0, float w1, float w2, float dx, float dy)
{
#pragma clang fp contract(off)
    return ((poi + w0 * dx) + w1 * dy) + w2;
}

// Bicubic value from the materialised LUT cell (16 floats, 4k+l ~ y^k x^l).
__device__ __forceinline__ float interp_lut(const float* __restrict__ lut, uint32_t cell, float tx, float ty)
{
    const float4* c = reinterpret_cast<const float4*>(lut) + (size_t)cell * 4;
    float4 r0 = c[0], r1 = c[1], r2 = c[2], r3 = c[3];
    float a0 = fmaf(fmaf(fmaf(r0.w, tx, r0.z), tx, r0.y), tx, r0.x);
    float a1 = fmaf(fmaf(fmaf(r1.w, tx, r1.z), tx, r1.y), tx, r1.x);
    float a2 = fmaf(fmaf(fmaf(r2.w, tx, r2.z), tx, r2.y), tx, r2.x);
    float a3 = fmaf(fmaf(fmaf(r3.w, tx, r3.z), tx, r3.y), tx, r3.x);
    return fmaf(fmaf(fmaf(a3, ty, a2), ty, a1), ty, a0);
}

// The same bicubic interpolant evaluated from the 4x4 pixel neighbourhood:
// the LUT polynomial built from central-difference corner derivatives is the
// tensor-product Catmull-Rom spline.  Weights in Horner form.
__device__ __forceinline__ void catmull_rom_w(float t, float w[4])
{
    const float t2 = t * t;
    w[0] = t * fmaf(t, fmaf(-0.5f, t, 1.f), -0.5f);
    w[1] = fmaf(t2, fmaf(1.5f, t, -2.5f), 1.f);
    w[2] = t * fmaf(t, fmaf(-1.5f, t, 2.f), 0.5f);
    w[3] = t2 * fmaf(0.5f, t, -0.5f);
}

// One neighbourhood row (its four pixels are the bytes of r) against the x
// weights.  The bytes are widened into 16-bit halves, where each is read as
// the fp16 subnormal byte * 2^-24, exactly; the mixed-precision FMA takes such
// a half as an exact f32 operand, so no conversion instruction is spent per
// pixel.  The result is 2^-24 times the weighted sum taken in the same order
// with the same roundings (a power-of-two scale commutes with rounding while
// nothing reaches the f32 subnormal range, and every product here is above
// 2^-80).  The first term adds +0 where a plain product would give -0 for a
// negative weight on a zero pixel; the row sums are equal otherwise.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ half2_t bytes02(uint32_t r)
{
    return __builtin_bit_cast(half2_t, r & 0x00ff00ffu);
}

__device__ __forceinline__ half2_t bytes13(uint32_t r)
{
    return __builtin_bit_cast(half2_t, __builtin_amdgcn_perm(0u, r, 0x0c030c01u));
}

__device__ __forceinline__ float row_dot_s(uint32_t r, const float wx[4])
{
    const half2_t e = bytes02(r), o = bytes13(r);
    float s = fmaf(wx[0], (float)e.x, 0.f);
    s = fmaf(wx[1], (float)o.x, s);
    s = fmaf(wx[2], (float)e.y, s);
    return fmaf(wx[3], (float)o.y, s);
}

// Bicubic value of a 4x4 neighbourhood given as four row words, scaled by
// 2^-24 (see row_dot_s).  A uniform neighbourhood returns its value exactly
// (as the LUT does: all higher coefficients vanish), so uniform target regions
// keep the reference's DarkSubset behaviour: rows 0, 2, 3 equal row 1, and
// row 1 equals itself turned by a byte, i.e. a chain of byte
// absolute-difference sums is zero.
__device__ __forceinline__ float interp_rows_s(uint32_t r0, uint32_t r1, uint32_t r2, uint32_t r3, float tx,
                                               float ty)
{
    float wx[4], wy[4];
    catmull_rom_w(tx, wx);
    catmull_rom_w(ty, wy);
    float acc = wy[0] * row_dot_s(r0, wx);
    acc = fmaf(wy[1], row_dot_s(r1, wx), acc);
    acc = fmaf(wy[2], row_dot_s(r2, wx), acc);
    acc = fmaf(wy[3], row_dot_s(r3, wx), acc);
    uint32_t diff = __builtin_amdgcn_sad_u8(r0, r1, 0u);
    diff = __builtin_amdgcn_sad_u8(r2, r1, diff);
    diff = __builtin_amdgcn_sad_u8(r3, r1, diff);
    diff = __builtin_amdgcn_sad_u8(r1, __builtin_amdgcn_alignbit(r1, r1, 8), diff);
    return diff == 0u ? (float)bytes13(r1).x : acc;
}

// The target image as a buffer resource whose base is the pixel (-1, -1):
// the 4x4 neighbourhood of (ix, iy) is then four unaligned dwords at byte
// offset iy stride + ix, plus 0, stride, 2 stride and 3 stride in the scalar
// offset, and the pixel itself is at that offset plus stride + 1.  ix, iy >= 1
// (the ROI keeps a 1-pixel frame inside the image wherever IC-GN runs), so
// every address lies in the image; offsets are 32-bit (images below 2 GiB).
// The range check (vector plus scalar offset below the ROI's last row + 4
// rows from the base) turns a stray offset into zeros instead of a fault.
struct TargetBuf {
    __amdgpu_buffer_rsrc_t r;
    int stride;
};

__device__ __forceinline__ TargetBuf target_buf(const IcgnParams& P)
{
    const int stride = P.stride;
    const uint8_t* base = P.tar - stride - 1;
    const int records = (P.sy0 + P.rh + 3) * stride;
    return {__builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, records, 0x00020000), stride};
}

__device__ __forceinline__ int target_offset(const TargetBuf& B, int ix, int iy)
{
    return (int)__umul24((uint32_t)iy, (uint32_t)B.stride) + ix;
}

// One unaligned dword per neighbourhood row; the value is scaled by 2^-24.
__device__ __forceinline__ float interp_fused_s(const TargetBuf& B, int ix, int iy, float tx, float ty)
{
    const int o = target_offset(B, ix, iy);
    const uint32_t r0 = __builtin_amdgcn_raw_buffer_load_b32(B.r, o, 0, 0);
    const uint32_t r1 = __builtin_amdgcn_raw_buffer_load_b32(B.r, o, B.stride, 0);
    const uint32_t r2 = __builtin_amdgcn_raw_buffer_load_b32(B.r, o, 2 * B.stride, 0);
    const uint32_t r3 = __builtin_amdgcn_raw_buffer_load_b32(B.r, o, 3 * B.stride, 0);
    return interp_rows_s(r0, r1, r2, r3, tx, ty);
}

// The target's row polynomials (launch_row_poly) as a buffer resource whose
// base is entry -(sy0 rw + sx0): cell (ix, iy) reads the 16-byte entries at
// byte offsets 16 (iy rw + ix) (rows iy - 1, iy) and that plus 32 rw (rows
// iy + 1, iy + 2).  The host passes the plane only when every such offset is
// below 2^32.
struct RowPolyBuf {
    __amdgpu_buffer_rsrc_t r;
    int rw16;      // 16 rw: bytes per entry row
};

__device__ __forceinline__ RowPolyBuf row_poly_buf(const IcgnParams& P)
{
    const uint32_t lead = 16u * ((uint32_t)P.sy0 * (uint32_t)P.rw + (uint32_t)P.sx0);
    const uint8_t* base = reinterpret_cast<const uint8_t*>(P.rp) - lead;
    const int records = (int)(lead + 16u * (uint32_t)(P.rh + 1) * (uint32_t)P.rw);
    return {__builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, records, 0x00020000), 16 * P.rw};
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// One row's cubic at t: c0 + t (c1/2 + t (c2/2 + t c3/2)), the fp16
// coefficients (stored halved, exactly) taken as exact f32 operands by the
// mixed-precision FMA.  Every intermediate is exactly half the one of
// c0 + (t/2) (c1 + t (c2 + t c3)) evaluated in the same order, so the two
// forms round identically.
__device__ __forceinline__ float row_poly(uint32_t c01, uint32_t c23, float t)
{
    const half2_t a = __builtin_bit_cast(half2_t, c01), b = __builtin_bit_cast(half2_t, c23);
    return fmaf(t, fmaf(t, fmaf(t, (float)b.y, (float)b.x), (float)a.y), (float)a.x);
}

// Bicubic value at (ix + tx, iy + ty): the four rows' cubics at tx, then the
// Catmull-Rom cubic through them at ty in the same form.  Its coefficients
// are formed from the row values with exact small-integer weights, so a
// uniform neighbourhood gives back its value exactly (every higher
// coefficient is an exact zero), as the LUT does: uniform target regions keep
// the reference's DarkSubset behaviour without a test.
__device__ __forceinline__ float interp_row_poly(const RowPolyBuf& B, int ix, int iy, float tx, float ty)
{
    const int o = (int)__umul24((uint32_t)iy, (uint32_t)B.rw16) + 16 * ix;
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(B.r, o, 0, 0);
    const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(B.r, o, 2 * B.rw16, 0);
    const float R0 = row_poly(a.x, a.y, tx), R1 = row_poly(a.z, a.w, tx);
    const float R2 = row_poly(b.x, b.y, tx), R3 = row_poly(b.z, b.w, tx);
    const float d1 = R2 - R0;
    const float d2 = fmaf(4.f, R2, fmaf(-5.f, R1, fmaf(2.f, R0, -R3)));
    const float d3 = fmaf(3.f, R1 - R2, R3 - R0);
    return fmaf(0.5f * ty, fmaf(fmaf(d3, ty, d2), ty, d1), R1);
}

// p from W, as the update sets it (TW_paDIC_ICGN2D_CPU.cpp:902-907): p0 = W02,
// p1 = W00 - 1, p2 = W01, p3 = W12, p4 = W10, p5 = W11 - 1.  For the initial
// W = [[1 + 0, 0, u], [0, 1 + 0, v]] (:605-607) this gives back (u, 0, 0, v,
// 0, 0), so p need not be kept beside W.
__device__ __forceinline__ void params_of(const float W[6], float P[6])
{
#pragma clang fp contract(off)
    P[0] = W[2];
    P[1] = W[0] - 1;
    P[2] = W[1];
    P[3] = W[5];
    P[4] = W[3];
    P[5] = W[4] - 1;
}

// a / b correctly rounded from y = RN(1 / b) by Markstein's correction
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990: when y is within half an ulp
// of 1/b and q0 within an ulp of a/b, the remainder r = a - b q0 is exact and
// RN(q0 + r y) = RN(a / b)), two FMAs instead of a full division per
// quotient.  A zero numerator returns q0 = a y, whose zero has the sign of
// a / b (the correction would add a +0 to it).  Used only away from overflow
// and subnormal quotients: |b| > FLT_EPSILON and the numerators are warp
// terms of order 1 or exact zeros.
__device__ __forceinline__ float div_rcp(float a, float b, float y)
{
    const float q0 = a * y;
    const float r = fmaf(-q0, b, a);
    return a == 0.f ? q0 : fmaf(r, y, q0);
}

// W <- W(p) o W(dp)^-1 with p = params_of(W): TW_paDIC_ICGN2D_CPU.cpp:713-745,
// in FP32 with every operation rounded separately (no contraction); the six
// quotients by den are the correctly rounded ones, through one reciprocal.
__device__ __forceinline__ void warp_update(const float d[6], float den, float W[6])
{
#pragma clang fp contract(off)
    float P[6];
    params_of(W, P);
    const float y = 1.f / den;
    float w00 = div_rcp((1 + P[1]) * (1 + d[5]) - P[2] * d[4], den, y);
    float w01 = div_rcp(P[2] * (1 + d[1]) - (1 + P[1]) * d[2], den, y);
    float w02 = P[0] + div_rcp(P[2] * (d[0] * d[4] - d[3] - d[3] * d[1]) - (1 + P[1]) * (d[0] * d[5] + d[0] - d[2] * d[3]), den, y);
    float w10 = div_rcp(P[4] * (1 + d[5]) - (1 + P[5]) * d[4], den, y);
    float w11 = div_rcp((1 + P[5]) * (1 + d[1]) - P[4] * d[2], den, y);
    float w12 = P[3] + div_rcp((1 + P[5]) * (d[0] * d[4] - d[3] - d[3] * d[1]) - P[4] * (d[0] * d[5] + d[0] - d[2] * d[3]), den, y);
    W[0] = w00; W[1] = w01; W[2] = w02; W[3] = w10; W[4] = w11; W[5] = w12;
}

__device__ __forceinline__ float warp_den(const float d[6])
{
#pragma clang fp contract(off)
    return (1 + d[1]) * (1 + d[5]) - d[2] * d[4];
}

// Loop test quantity (:751): products in FP32, squares and sum in FP64.
__device__ __forceinline__ double conv_norm(const float d[6], int Sx, int Sy)
{
    double t0 = d[0];
    double t1 = __fmul_rn(d[1], (float)Sx);
    double t2 = __fmul_rn(d[2], (float)Sy);
    double t3 = d[3];
    double t4 = __fmul_rn(d[4], (float)Sx);
    double t5 = __fmul_rn(d[5], (float)Sy);
    return sqrt(t0 * t0 + t1 * t1 + t2 * t2 + t3 * t3 + t4 * t4 + t5 * t5);
}

}  // namespace

// FP32 sums with compensated (Kahan) addition of each pixel's T, T^2 and
// SD_k T / 2 terms, every product and sum rounded separately: the compensated
// FP32 accumulation, built as a measured alternative to the FP64 sums
// (TWDIC_ICGN_VARIANT=20), not the default.  The FP32 products SD_k T round
// here, where the FP64 ones are exact.
__device__ __forceinline__ void kahan_pixel(float ks[8], float kc[8], float T, uint32_t g, float dx, float dy)
{
#pragma clang fp contract(off)
    const float gx = (float)(int16_t)(g & 0xffffu), gy = (float)(int16_t)(g >> 16);
    const float a = gx * T, b = gy * T;
    const float x[8] = {T, T * T, a, a * dx, a * dy, b, b * dx, b * dy};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float y = x[i] - kc[i];
        const float t = ks[i] + y;
        kc[i] = (t - ks[i]) - y;
        ks[i] = t;
    }
}

// The subset's pixels in row-major order, p = l * Wsub + m: the exact float
// offsets dx = m - Sx, dy = l - Sy and the byte offset 4 (l * rw + m) into
// the gradient plane from the subset's corner.  Every POI of a launch has the
// same subset and ROI width, so one table in LDS serves a workgroup's eight
// POIs, and a lane's step to its next pixel is an LDS read instead of
// float compares and selects.  The three arrays have the same 4-byte stride,
// so one address register reaches all of them (the array offsets are
// immediates of the LDS reads).
constexpr int kStepTable = 2048;          // pixels: subsets up to 45 x 45
struct StepTable {
    float dx[kStepTable];
    float dy[kStepTable];
    uint32_t go[kStepTable];
};

__device__ __forceinline__ void fill_step_table(StepTable& t, const IcgnParams& P)
{
    for (int p = threadIdx.x; p < P.s; p += blockDim.x) {
        const int l = p / P.Wsub, m = p - l * P.Wsub;
        t.dx[p] = (float)(m - P.Sx);
        t.dy[p] = (float)(l - P.Sy);
        t.go[p] = 4u * (uint32_t)(l * P.rw + m);
    }
}

// One pass over the subset at the current warp, by the L lanes of a POI's
// group (hl = lane % L; px, py and the warp are the group's POI's).  INT:
// first pass at integer target pixels (:610-637); otherwise bicubic
// (:757-792).  Returns false when a warped pixel of the group's POI leaves the
// ROI.  Accumulates sum T, sum T^2 and the six sum SD_k T / 2 (the gradient
// plane stores 2 Rx, 2 Ry) in FP64.  With a step table (tab != null) the lanes
// take pixels hl + L k from it: s / L full rounds, then the s % L remaining pixels;
// without one they step their offsets in registers.  Each lane visits the
// same pixels in the same order either way, so the sums are identical.  RP:
// the fused bicubic passes evaluate the target's row polynomials (P.rp)
// instead of the image bytes.  UNR: pixels per trip of the step-table loop.
template <int L, int INTERP, bool INT, bool CHECK, bool TABLE, bool RP, bool K32 = false, int UNR = 1>
__device__ __forceinline__ bool subset_loop(const IcgnParams& P, int px, int py, const float Wm[6], int lane,
                                            double acc[8], const StepTable* tab)
{
    const int hl = lane & (L - 1);
    // the fused interpolation from the bytes returns T 2^-24 (interp_rows_s);
    // the sums are scaled back at the end, exactly
    constexpr bool SCALED = !INT && INTERP == TWDIC_INTERP_FUSED && !RP;
    const int Sx = P.Sx, Sy = P.Sy, Wsub = P.Wsub, s = P.s;
    const int sx0 = P.sx0, sy0 = P.sy0, rw = P.rw, rh = P.rh;
    const TargetBuf tb = target_buf(P);
    const __amdgpu_buffer_rsrc_t g2r =
        __builtin_amdgcn_make_buffer_rsrc((void*)P.g2, (short)0, 4 * rw * rh, 0x00020000);
    const RowPolyBuf rb = RP ? row_poly_buf(P) : RowPolyBuf{};
    const float fpx = (float)px, fpy = (float)py;
    double st = 0.0, stt = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    float ks[8] = {}, kc[8] = {};
    bool out = false;
    auto pixel = [&](float dx, float dy, uint32_t go, int goff) {
        float wx, wy;
        if constexpr (INT) {
            // the first pass's warp is W = [[1, 0, u], [0, 1, v]] exactly, where
            // ((POI + 1 dx) + 0 dy) + u = (POI + dx) + u bit for bit (POI + dx
            // is a non-zero integer, so adding the zero product changes nothing)
#pragma clang fp contract(off)
            wx = (fpx + dx) + Wm[2];
            wy = (fpy + dy) + Wm[5];
        } else {
            wx = warp_coord(fpx, Wm[0], Wm[1], Wm[2], dx, dy);
            wy = warp_coord(fpy, Wm[3], Wm[4], Wm[5], dx, dy);
        }
        const int ix = (int)wx, iy = (int)wy;          // truncation, as int()
        float T = 0.f;
        bool in = true, interior = true;
        if (CHECK) {
            in = ix >= sx0 && iy >= sy0 && ix < sx0 + rw && iy < sy0 + rh;
            // the CPU bicubic LUT's last row/column is zero; the B-spline LUT covers it
            interior = in && (INTERP == TWDIC_INTERP_BSPLINE || (ix != sx0 + rw - 1 && iy != sy0 + rh - 1));
            out |= !in;
        }
        if (INT) {
            if (in) T = (float)__builtin_amdgcn_raw_buffer_load_b8(tb.r, target_offset(tb, ix, iy), tb.stride + 1, 0); // :629
        } else if (interior) {
            // :771-772; wx, wy >= 1 here, so x - (float)(int)x is fract(x), exactly
            const float tx = __builtin_amdgcn_fractf(wx), ty = __builtin_amdgcn_fractf(wy);
            if constexpr (RP) {
                T = interp_row_poly(rb, ix, iy, tx, ty);
            } else if (INTERP == TWDIC_INTERP_FUSED) {
                T = interp_fused_s(tb, ix, iy, tx, ty);
            } else {
                T = interp_lut(P.lut, (uint32_t)((iy - sy0) * rw + (ix - sx0)), tx, ty);
            }
        }
        const uint32_t g = __builtin_amdgcn_raw_buffer_load_b32(g2r, (int)go + goff, 0, 0);
        if constexpr (K32) {
            kahan_pixel(ks, kc, T, g, dx, dy);
            return;
        }
        // FP64: every product SD_k T is exact, and where T takes integer values
        // (the integer pass, uniform target regions) the sums are exact, so
        // T = R gives an exactly zero update and near-flat subsets, whose
        // residual is scaled by the large R_s / T_s, see no rounding there
        const double gx = (double)(int16_t)(g & 0xffffu), gy = (double)(int16_t)(g >> 16);
        const double Td = T, ddx = dx, ddy = dy;
        st += Td;
        stt = fma(Td, Td, stt);
        const double a = gx * Td, b = gy * Td;
        a0 += a;
        a1 = fma(a, ddx, a1);
        a2 = fma(a, ddy, a2);
        b0 += b;
        b1 = fma(b, ddx, b1);
        b2 = fma(b, ddy, b2);
    };
    // byte offset of the subset's corner pixel in the gradient plane
    const int gcorner = 4 * ((py - Sy - sy0) * rw + (px - Sx - sx0));
    if constexpr (TABLE) {
        const int full = s / L;
        // with UNR = 2 two pixels' loads are in flight together
#pragma unroll UNR
        for (int k = 0; k < full; ++k) {
            const int p = hl + L * k;
            pixel(tab->dx[p], tab->dy[p], tab->go[p], gcorner);
        }
        const int p = hl + L * full;
        if (p < s) {
            pixel(tab->dx[p], tab->dy[p], tab->go[p], gcorner);
        }
    } else {
        // lane walks pixels p = hl, hl + L, ...: column
        // m = p % Wsub, row l = p / Wsub, kept as the exact float offsets
        // dx = m - Sx, dy = l - Sy and the 32-bit byte offset into the
        // gradient plane
        const int qL = L / Wsub, rL = L - qL * Wsub;
        // recomputed per pass: kept live across the iteration loop, this start
        // state is what the register allocator spills
        int lane_v = hl;
        asm volatile("" : "+v"(lane_v));
        const int l0 = lane_v / Wsub, m0 = lane_v - l0 * Wsub;
        float dx = (float)(m0 - Sx), dy = (float)(l0 - Sy);
        uint32_t go = (uint32_t)gcorner + 4u * (uint32_t)(l0 * rw + m0);
        const float fSx = (float)Sx, fWsub = (float)Wsub, frL = (float)rL, fqL = (float)qL;
        const uint32_t gstep = 4u * (uint32_t)(qL * rw + rL), gwrap = 4u * (uint32_t)(rw - Wsub);
#pragma unroll 1
        for (int p = hl; p < s; p += L) {
            pixel(dx, dy, go, 0);
            // next pixel of this lane: L further in row-major subset order
            dx += frL;
            const bool wrap = dx > fSx;
            dx = wrap ? dx - fWsub : dx;
            dy += wrap ? fqL + 1.f : fqL;
            go += wrap ? gstep + gwrap : gstep;
        }
    }
    // 2^24 and 2^48 scale the sums of T 2^-24 back exactly (FP64, far from
    // its range limits)
    const double sc = SCALED ? 0x1p24 : 1.0, sc2 = SCALED ? 0x1p48 : 1.0, hsc = SCALED ? 0x1p23 : 0.5;
    if constexpr (K32) {
        st = (double)ks[0] - (double)kc[0]; stt = (double)ks[1] - (double)kc[1];
        a0 = (double)ks[2] - (double)kc[2]; a1 = (double)ks[3] - (double)kc[3]; a2 = (double)ks[4] - (double)kc[4];
        b0 = (double)ks[5] - (double)kc[5]; b1 = (double)ks[6] - (double)kc[6]; b2 = (double)ks[7] - (double)kc[7];
    }
    acc[0] = sc * st; acc[1] = sc2 * stt;
    acc[2] = hsc * a0; acc[3] = hsc * a1; acc[4] = hsc * a2;
    acc[5] = hsc * b0; acc[6] = hsc * b1; acc[7] = hsc * b2;
    return CHECK ? !group_any<L>(out, lane) : true;
}

// One pass over the subset at the current warp.  INT: first pass at integer
// target pixels (:610-637); otherwise bicubic (:757-792).  Returns false when
// a warped pixel leaves the ROI.  Accumulates sum T, sum T^2 and the six
// sum SD_k T / 2 (the gradient plane stores 2 Rx, 2 Ry) in FP64.
// The warp is affine, so over the subset (offsets |dx| <= Sx, |dy| <= Sy) its
// coordinates lie within POI + W2 -+ (|W0| Sx + |W1| Sy) (and likewise in
// y).  When that range lies at least one pixel inside the region where no
// per-pixel test can fire (the margin covers the FP32 rounding of the
// per-pixel warp and of the range, both below 1e-2 px), the pass runs
// without the per-pixel tests; the results are identical either way.
template <int L, int INTERP, bool INT, bool TABLE, bool RP, bool K32 = false, int UNR = 1>
__device__ __forceinline__ bool subset_pass(const IcgnParams& P, int px, int py, const float Wm[6], int lane,
                                            double acc[8], const StepTable* tab)
{
    const float ex = (float)P.Sx, ey = (float)P.Sy;
    const float cx = (float)px + Wm[2], cy = (float)py + Wm[5];
    const float rx = fabsf(Wm[0] * ex) + fabsf(Wm[1] * ey), ry = fabsf(Wm[3] * ex) + fabsf(Wm[4] * ey);
    const bool safe = cx - rx >= (float)(P.sx0 + 1) && cx + rx < (float)(P.sx0 + P.rw - 2) &&
                      cy - ry >= (float)(P.sy0 + 1) && cy + ry < (float)(P.sy0 + P.rh - 2);
    // both halves take the same loop: the per-pixel tests run unless every
    // active lane's POI is safe
    if (__all(safe)) return subset_loop<L, INTERP, INT, false, TABLE, RP, K32, UNR>(P, px, py, Wm, lane, acc, tab);
    return subset_loop<L, INTERP, INT, true, TABLE, RP, K32, UNR>(P, px, py, Wm, lane, acc, tab);
}

// IC-GN for one POI by a group of L lanes: a whole wavefront (L = 64) or
// one half (L = 32; the other half may run another POI of the same launch or
// none).  rs: the POI's reference state, copied to LDS by the caller (its
// first kRefStateUsed doubles).  Every lane of the group holds the same
// per-POI values, so with halves the two POIs' solves and warp updates issue
// as one instruction stream; with whole wavefronts they are wave-uniform and
// kept in SGPRs.
template <int L, int INTERP, bool TABLE, bool RP, bool K32 = false, int UNR = 1>
__device__ __forceinline__ void icgn_poi_t(const IcgnParams& P, int poi, int lane, const StepTable* tab,
                                           const double* rs)
{
    int py = P.poi[2 * poi];
    int px = P.poi[2 * poi + 1];
    if constexpr (L == 64) {
        py = __builtin_amdgcn_readfirstlane(py);
        px = __builtin_amdgcn_readfirstlane(px);
    }
    const int Sx = P.Sx, Sy = P.Sy;
    const double dn = (double)P.s;

    const float u0 = P.u[poi], v0 = P.v[poi];
    float Wm[6] = {1.f + 0.f, 0.f, u0, 0.f, 1.f + 0.f, v0};  // :605-607 with p = (u, 0, 0, v, 0, 0)
    int iters = 0;                 // 0: leave the caller's value untouched
    int flag = TWDIC_ICGN_SUCCESS;
    float dpn = __builtin_nanf("");

    const int refflag = (int)rs[RS_FLAG];
    if (refflag == TWDIC_ICGN_DARK_SUBSET || refflag == TWDIC_ICGN_OUT_OF_ROI) {
        flag = refflag;
    } else {
        for (int it = 0;; ++it) {
            if (it >= 1) iters = it + 1;                       // ++iNumIterations (:753)
            double acc[8];
            const bool inside = it == 0 ? subset_pass<L, INTERP, true, TABLE, RP, K32, UNR>(P, px, py, Wm, lane, acc, tab)
                                        : subset_pass<L, INTERP, false, TABLE, RP, K32, UNR>(P, px, py, Wm, lane, acc, tab);
            if (!inside) { flag = TWDIC_ICGN_OUT_OF_ROI; break; }         // :632-635, :787-790
#pragma unroll
            for (int i = 0; i < 8; ++i) acc[i] = group_sum<L>(acc[i]);
            double meanT, normT;
            mean_spread(acc[0], acc[1], dn, meanT, normT);
            if (fabsf((float)meanT) <= FLT_EPSILON) { flag = TWDIC_ICGN_DARK_SUBSET; break; }   // :638-640
            if ((float)normT <= FLT_EPSILON) { flag = TWDIC_ICGN_DARK_SUBSET; break; }          // :651-653
            if (it == 0 && refflag == TWDIC_ICGN_SINGULAR_HESSIAN) {                              // :688-693
                flag = TWDIC_ICGN_SINGULAR_HESSIAN;
                break;
            }
            // RHS_k = sum SD_k ((R_s/T_s) (T - mean T) - (R - mean R))   (:657-667),
            // the last sum from the reference state; R_s is read here, not
            // held across the pass, which keeps the pass within the register
            // budget of five waves per SIMD
            int at = 0;
            asm volatile("" : "+v"(at));
            const double normR = rs[RS_NORM + at];
            const double nor = normR / normT;
            double r[6];
#pragma unroll
            for (int k = 0; k < 6; ++k)
                r[k] = __fma_rn(nor, centred(acc[2 + k], meanT, rs[RS_A + k]), -rs[RS_B + k]);
            // L y = r, L^T x = y with the FP64 factor (spotrs 'L', :683, :853)
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                double v = r[i];
#pragma unroll
                for (int k = 0; k < i; ++k) v -= rs[RS_L + i * (i + 1) / 2 + k] * r[k];
                r[i] = v * rs[RS_RDIAG + i];
            }
#pragma unroll
            for (int i = 5; i >= 0; --i) {
                double v = r[i];
#pragma unroll
                for (int k = i + 1; k < 6; ++k) v -= rs[RS_L + k * (k + 1) / 2 + i] * r[k];
                r[i] = v * rs[RS_RDIAG + i];
            }
            float d[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) d[i] = (float)r[i];
            const float den = warp_den(d);
            if (fabsf(den) <= FLT_EPSILON) { flag = TWDIC_ICGN_SINGULAR_WARP; break; }        // :875-877
            warp_update(d, den, Wm);
            if constexpr (L == 64) {
                // the butterflies leave the sums, hence the warp, bit-identical
                // in all lanes: keep it in SGPRs for the next pass
#pragma unroll
                for (int i = 0; i < 6; ++i) Wm[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(Wm[i])));
            }
            const double nrm = conv_norm(d, Sx, Sy);
            dpn = (float)nrm;
            if (it == 0) iters = 1;                                    // :749
            if (!(iters < P.max_iter && nrm >= P.delta_p)) break;
        }
    }
    if ((lane & (L - 1)) == 0) {
        float Pp[6];
        params_of(Wm, Pp);
        P.u[poi] = Pp[0];
        P.v[poi] = Pp[3];
        if (P.p6) {
#pragma unroll
            for (int i = 0; i < 6; ++i) P.p6[(size_t)poi * 6 + i] = Pp[i];
        }
        if (P.iters && iters > 0) P.iters[poi] = iters;
        if (P.flags) P.flags[poi] = (uint8_t)flag;
        if (P.dpnorm) P.dpnorm[poi] = dpn;
    }
}

// The fused mode reads the target's row polynomials when the caller built
// them (P.rp, the same for every wave of a launch), else the image bytes;
// both evaluate the same Catmull-Rom spline.
template <int L, int INTERP, bool TABLE>
__device__ __forceinline__ void icgn_poi(const IcgnParams& P, int poi, int lane, const StepTable* tab,
                                         const double* rs)
{
    if (INTERP == TWDIC_INTERP_FUSED && P.rp != nullptr)
        icgn_poi_t<L, INTERP, TABLE, INTERP == TWDIC_INTERP_FUSED>(P, poi, lane, tab, rs);
    else
        icgn_poi_t<L, INTERP, TABLE, false>(P, poi, lane, tab, rs);
}

// The reference state of the POIs first .. first + count - 1 (those below n)
// into slots 0 .. count - 1 of an LDS array, by the threads t0, t0 + tstride, ...
__device__ __forceinline__ void copy_ref_state(double (*slots)[kRefStateUsed], const double* __restrict__ state,
                                               int first, int count, int n, int t0, int tstride)
{
    for (int i = t0; i < count * kRefStateUsed; i += tstride) {
        const int slot = i / kRefStateUsed, j = i - slot * kRefStateUsed;
        if (first + slot < n) slots[slot][j] = state[(size_t)(first + slot) * kRefStateStride + j];
    }
}

// The single-context kernel takes the choice at launch instead (RP: P.rp is
// set), so each form gets its own register budget.

// MINW: occupancy target (waves per SIMD) the register budget is fitted to;
// 0 leaves the choice to the compiler.
// TABLE: the subset has at most kStepTable pixels; the workgroup builds the
// step table in LDS first.
// L lanes per POI, 256 / L POIs per workgroup (POI (256 / L) b' + t / L for
// thread t); consecutive workgroups go to the 8 XCDs round-robin, so block b
// is remapped to b' = the next block of XCD b % 8's contiguous share and each
// XCD's L2 holds one band of the image.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int x = b & 7, slot = b >> 3, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + slot;
}

template <int INTERP, int MINW, bool TABLE, bool RP, int L, bool K32 = false>
__global__ __launch_bounds__(256, MINW > 0 ? MINW : 1) void k_icgn(IcgnParams P)
{
    constexpr int G = 256 / L;                   // POIs per workgroup
    __shared__ double rs[G][kRefStateUsed];
    const int first = xcd_block(blockIdx.x, gridDim.x) * G;
    const int slot = threadIdx.x / L, poi = first + slot;
    copy_ref_state(rs, P.refstate, first, G, P.n, threadIdx.x, 256);
    if constexpr (TABLE) {
        __shared__ StepTable tab;
        fill_step_table(tab, P);
        __syncthreads();
        if (poi >= P.n) return;
        // two pixels per table-loop trip with halves (0.6 % faster on the
        // 2048^2 frame); the batched kernel keeps one, which was 5 % faster there
        icgn_poi_t<L, INTERP, true, RP, K32, L == 32 ? 2 : 1>(P, poi, threadIdx.x & 63, &tab, rs[slot]);
    } else {
        __syncthreads();
        if (poi >= P.n) return;
        icgn_poi_t<L, INTERP, false, RP, K32>(P, poi, threadIdx.x & 63, nullptr, rs[slot]);
    }
}

// Batched form, one launch over many image pairs: workgroup item w =
// (pair, first POI) holds up to 256 / L POIs of one pair (every listed pair
// takes L lanes per POI, icgn_lanes_for of its POI count), so the
// workgroup's step table is that pair's.  XCD: items in contiguous per-XCD ranges as in
// k_icgn (one band of the work per L2); otherwise dealt out in block order,
// which balances the XCDs when the work per item differs from pair to pair.
// ALL_RP: every pair has its row-polynomial plane (the usual case), so only
// that form is compiled in and gets the register budget; otherwise each pair
// takes its form at run time.
template <int INTERP, int MINW, bool TABLE, bool XCD, bool ALL_RP, int L>
__global__ __launch_bounds__(256, MINW > 0 ? MINW : 1) void k_batch_icgn_wg(const PairDesc* __restrict__ pairs,
                                                                           const int2* __restrict__ wg)
{
    constexpr int G = 256 / L;
    const int2 it = wg[XCD ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x];
    const IcgnParams P = const_pairs(pairs)[__builtin_amdgcn_readfirstlane(it.x)].icgn;
    __shared__ double rs[G][kRefStateUsed];
    const int first = __builtin_amdgcn_readfirstlane(it.y);
    const int slot = threadIdx.x / L, poi = first + slot;
    copy_ref_state(rs, P.refstate, first, G, P.n, threadIdx.x, 256);
    if constexpr (TABLE) {
        __shared__ StepTable tab;
        fill_step_table(tab, P);
        __syncthreads();
        if (poi >= P.n) return;
        if constexpr (ALL_RP)
            icgn_poi_t<L, INTERP, true, true>(P, poi, threadIdx.x & 63, &tab, rs[slot]);
        else
            icgn_poi<L, INTERP, true>(P, poi, threadIdx.x & 63, &tab, rs[slot]);
    } else {
        __syncthreads();
        if (poi >= P.n) return;
        if constexpr (ALL_RP)
            icgn_poi_t<L, INTERP, false, true>(P, poi, threadIdx.x & 63, nullptr, rs[slot]);
        else
            icgn_poi<L, INTERP, false>(P, poi, threadIdx.x & 63, nullptr, rs[slot]);
    }
}

bool step_table_fits(int wsub, int hsub)
{
    return wsub * hsub <= kStepTable;
}

template <int MINW, bool TABLE, bool XCD, bool ALL_RP, int L>
static void launch_wg(const PairDesc* pairs, const int2* wg, int count, hipStream_t st)
{
    hipLaunchKernelGGL((k_batch_icgn_wg<TWDIC_INTERP_FUSED, MINW, TABLE, XCD, ALL_RP, L>), dim3(count), dim3(256), 0,
                       st, pairs, wg);
}

template <int MINW, bool ALL_RP, int L>
static void launch_wg_w(const PairDesc* pairs, const int2* wg, int count, bool table, bool xcd, hipStream_t st)
{
    if (table)
        xcd ? launch_wg<MINW, true, true, ALL_RP, L>(pairs, wg, count, st)
            : launch_wg<MINW, true, false, ALL_RP, L>(pairs, wg, count, st);
    else
        xcd ? launch_wg<MINW, false, true, ALL_RP, L>(pairs, wg, count, st)
            : launch_wg<MINW, false, false, ALL_RP, L>(pairs, wg, count, st);
}

template <int L>
static void launch_wg_l(const PairDesc* pairs, const int2* wg, int count, bool table, int waves, bool xcd_ranges,
                        bool all_rp, hipStream_t st)
{
    if (all_rp)
        waves == 6 ? launch_wg_w<6, true, L>(pairs, wg, count, table, xcd_ranges, st)
                   : launch_wg_w<5, true, L>(pairs, wg, count, table, xcd_ranges, st);
    else
        waves == 6 ? launch_wg_w<6, false, L>(pairs, wg, count, table, xcd_ranges, st)
                   : launch_wg_w<5, false, L>(pairs, wg, count, table, xcd_ranges, st);
}

hipError_t launch_batch_icgn_wg(const PairDesc* pairs, const int2* wg, int count, bool table, int lanes, int waves,
                                bool xcd_ranges, bool all_rp, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    if (lanes == 64)
        launch_wg_l<64>(pairs, wg, count, table, waves, xcd_ranges, all_rp, st);
    else
        launch_wg_l<32>(pairs, wg, count, table, waves, xcd_ranges, all_rp, st);
    return hipGetLastError();
}

// Persistent form: every wavefront takes `chunk` consecutive work items
// (image pair, POI) from a device counter until the list is exhausted; each
// item's geometry and buffers come from its pair's descriptor.  Every wave
// exits once the counter passes `total`.  Consecutive items may belong to
// different pairs, so each item runs alone, on its pair's L lanes (the whole
// wavefront, or its lower half).
__global__ __launch_bounds__(256) void k_batch_icgn(const PairDesc* __restrict__ pairs, const int2* __restrict__ items,
                                                    int total, int* counter, int chunk)
{
    __shared__ double rs[4][1][kRefStateUsed];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(counter, chunk);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
        if (base >= total) break;
        const int end = min(base + chunk, total);
        for (int i = base; i < end; ++i) {
            const int2 it = items[i];
            const int pair = __builtin_amdgcn_readfirstlane(it.x), poi = __builtin_amdgcn_readfirstlane(it.y);
            const IcgnParams P = const_pairs(pairs)[pair].icgn;
            // the wave's LDS slot: written and read by this wavefront only
            __builtin_amdgcn_wave_barrier();
            copy_ref_state(rs[wave], P.refstate, poi, 1, P.n, lane, 64);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (P.lanes == 64)
                icgn_poi<64, TWDIC_INTERP_FUSED, false>(P, poi, lane, nullptr, rs[wave][0]);
            else if (lane < 32)
                icgn_poi<32, TWDIC_INTERP_FUSED, false>(P, poi, lane, nullptr, rs[wave][0]);
        }
    }
}

int persistent_blocks_icgn(int device)
{
    int cus = 0, per = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_icgn, 256, 0) != hipSuccess || per < 1) per = 1;
    return (cus > 0 ? cus : 256) * per;
}

hipError_t launch_batch_icgn(const PairDesc* pairs, const int2* items, int total, int* counter, int chunk,
                             int blocks, hipStream_t st)
{
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_icgn, dim3(blocks), dim3(256), 0, st, pairs, items, total, counter, chunk);
    return hipGetLastError();
}

// variant: the occupancy target of the half-wavefront form (5 or 6 waves per
// SIMD, or 0, the compiler's choice), or 20, the compensated FP32 sums; the
// whole-wavefront form is built for 5 waves per SIMD.  (Seven waves do not
// fit: the workgroup's LDS caps a CU at five.)
template <int INTERP, bool TABLE, bool RP, int L>
static void launch_icgn_v(const IcgnParams& p, int variant, hipStream_t st)
{
    constexpr int G = 256 / L;
    const dim3 grid((p.n + G - 1) / G);
    if (variant == 20) {
        hipLaunchKernelGGL((k_icgn<INTERP, 5, TABLE, RP, L, true>), grid, dim3(256), 0, st, p);
        return;
    }
    if constexpr (L == 64) {
        hipLaunchKernelGGL((k_icgn<INTERP, 5, TABLE, RP, 64>), grid, dim3(256), 0, st, p);
    } else {
        switch (variant) {
        case 5: hipLaunchKernelGGL((k_icgn<INTERP, 5, TABLE, RP, L>), grid, dim3(256), 0, st, p); break;
        case 6: hipLaunchKernelGGL((k_icgn<INTERP, 6, TABLE, RP, L>), grid, dim3(256), 0, st, p); break;
        default: hipLaunchKernelGGL((k_icgn<INTERP, 0, TABLE, RP, L>), grid, dim3(256), 0, st, p); break;
        }
    }
}

template <int INTERP, bool RP>
static void launch_icgn_r(const IcgnParams& p, int variant, hipStream_t st)
{
    const bool table = p.s <= kStepTable;
    if (p.lanes == 64)
        table ? launch_icgn_v<INTERP, true, RP, 64>(p, variant, st) : launch_icgn_v<INTERP, false, RP, 64>(p, variant, st);
    else
        table ? launch_icgn_v<INTERP, true, RP, 32>(p, variant, st) : launch_icgn_v<INTERP, false, RP, 32>(p, variant, st);
}

template <int INTERP>
static void launch_icgn_t(const IcgnParams& p, int variant, hipStream_t st)
{
    if (INTERP == TWDIC_INTERP_FUSED && p.rp != nullptr)
        launch_icgn_r<INTERP, INTERP == TWDIC_INTERP_FUSED>(p, variant, st);
    else
        launch_icgn_r<INTERP, false>(p, variant, st);
}

hipError_t launch_icgn(const IcgnParams& p, int interp, int variant, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    if (p.lanes != 32 && p.lanes != 64) return hipErrorInvalidValue;
    if (interp == TWDIC_INTERP_FUSED)
        launch_icgn_t<TWDIC_INTERP_FUSED>(p, variant, st);
    else if (interp == TWDIC_INTERP_BSPLINE)
        launch_icgn_t<TWDIC_INTERP_BSPLINE>(p, variant, st);
    else
        launch_icgn_t<TWDIC_INTERP_LUT>(p, variant, st);
    return hipGetLastError();
}

}  // namespace twdic
