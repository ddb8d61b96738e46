// fftcc.hip -- FFT-CC integer-pixel search for gfx950.
//
// Semantics follow Fftcc2D_CPU::ComputeFFTCC (TW_paDIC_FFTCC2D_CPU.cpp:
// 222-326) with the window edge N as a parameter (N = 2S is the reference):
//   window rows POIy - N/2 .. POIy + N/2 - 1, same for columns;
//   C[k] = sum_x R~(x) T~(x+k) (circular), peak = argmax C, lowest row-major
//   index on ties; u = k % N, v = k / N, wrapped by -N when > N/2;
//   ZNCC = C[peak] / sqrt(sum R~^2 sum T~^2).
//
// Exactness: the zero-mean correlation differs from the integer correlation
// I[k] = sum_x R(x) T(x+k) by a constant, so the true peak is an integer
// argmax, and every engine here returns that exact peak.  The reported ZNCC
// is evaluated in FP64 from the exact integer sums at the chosen peak, so
// every output is a function of the two windows alone: results do not depend
// on the engine, batching or pairing.
//
// Engines:
//   * k_fftcc_mma<NT, N64, TILE> (default, any even N <= 64): the whole
//     integer correlation surface on the int8 matrix cores, one wavefront per
//     POI (32 x 32 x 32 tiles; 16 x 16 x 64 tiles for 32 < N <= 48).
//   * k_fftcc_fft<P> (N = 4P = 32 or 64): FP32 FFTs.  One workgroup of 4N
//     threads owns a POI pair.  Each window pair is packed into
//     Z = R^ + i T^ (R^, T^ zero-mean, unit norm), so one forward 2-D FFT
//     yields both spectra; the cross-power spectra X_A, X_B are Hermitian, so
//     a single inverse FFT of X_A + i X_B returns both correlation surfaces
//     (real and imaginary parts): 3 complex transforms per 2 POIs.  A 4N-point
//     line transform runs as a P-point radix-4 DFT in registers and a 4-point
//     DFT across a lane quad with DPP quad permutes; transposes go through
//     LDS.  When another shift lies within `tol` (ZNCC units) of the FP32
//     maximum, all such shifts are re-evaluated exactly in integers (lowest
//     index wins ties); with a single candidate the FP32 argmax is provably
//     the exact one.
//   * k_fftcc_dft: FP32 direct DFT along rows/columns for any other even
//     N <= 64, with the same exact re-evaluation.
#include "twdic_internal.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <mutex>

namespace twdic {

namespace {

constexpr float kTwoPi = 6.283185307179586f;

struct Top2 {
    float best;
    int idx;
    float second;
};

__device__ __forceinline__ Top2 top2_merge(Top2 a, Top2 b)
{
    Top2 r;
    if (a.best > b.best || (a.best == b.best && a.idx < b.idx)) {
        r.best = a.best; r.idx = a.idx; r.second = fmaxf(a.second, b.best);
    } else {
        r.best = b.best; r.idx = b.idx; r.second = fmaxf(b.second, a.best);
    }
    return r;
}

__device__ __forceinline__ Top2 wave_top2(Top2 t)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Top2 u;
        u.best = __shfl_xor(t.best, o);
        u.idx = __shfl_xor(t.idx, o);
        u.second = __shfl_xor(t.second, o);
        t = top2_merge(t, u);
    }
    return t;
}

__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// i*SGN*d
template <int SGN>
__device__ __forceinline__ float2 rot90(float2 d)
{
    return SGN > 0 ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
}

// Add one value to a running top-2 whose indices arrive in increasing order
// (equal values keep the earlier, lower index).  Branch-free.
__device__ __forceinline__ void top2_push(Top2& t, float c, int idx)
{
    t.second = fmaxf(t.second, fminf(t.best, c));
    t.idx = c > t.best ? idx : t.idx;
    t.best = fmaxf(t.best, c);
}

// Wrap a row-major peak index to the reference's (u, v) (:316-322).
__device__ __forceinline__ void wrap_peak(int k, int N, int& u, int& v)
{
    u = k % N;
    v = k / N;
    if (u > N / 2) u -= N;
    if (v > N / 2) v -= N;
}

// ----------------------------------------------------------- register DFTs
// 4-point DFT, natural order, X[k] = sum a[n] exp(i SGN 2 pi n k / 4).
template <int SGN>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3)
{
    float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = rot90<SGN>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <int SGN>
__device__ __forceinline__ float2 twc(float2 a, int e, int n)   // a * exp(i SGN 2 pi e / n), e,n constant
{
    const float c = __builtin_cosf(kTwoPi * (float)e / (float)n);
    const float s = (float)SGN * __builtin_sinf(kTwoPi * (float)e / (float)n);
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// P-point DFT in registers, natural order in and out (P = 16: 4 x 4; P = 8: 4 x 2).
template <int P, int SGN>
__device__ __forceinline__ void dft_reg(float2 (&x)[P])
{
    if constexpr (P == 16) {
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4<SGN>(x[n2], x[4 + n2], x[8 + n2], x[12 + n2]);
        // x[4*k1 + n2] now holds Y[n2][k1]; twiddle by W16^(n2 k1)
#pragma unroll
        for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
            for (int k1 = 1; k1 < 4; ++k1) x[4 * k1 + n2] = twc<SGN>(x[4 * k1 + n2], n2 * k1, 16);
        float2 y[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            float2 a0 = x[4 * k1], a1 = x[4 * k1 + 1], a2 = x[4 * k1 + 2], a3 = x[4 * k1 + 3];
            dft4<SGN>(a0, a1, a2, a3);
            y[k1] = a0; y[k1 + 4] = a1; y[k1 + 8] = a2; y[k1 + 12] = a3;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = y[i];
    } else {
        static_assert(P == 8, "P must be 8 or 16");
        // n = 2 n1 + n2: 4-point DFTs over n1 for n2 = 0, 1
        dft4<SGN>(x[0], x[2], x[4], x[6]);
        dft4<SGN>(x[1], x[3], x[5], x[7]);
        float2 y[8];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            float2 a = x[2 * k1], b = twc<SGN>(x[2 * k1 + 1], k1, 8);
            y[k1] = cadd(a, b);
            y[k1 + 4] = csub(a, b);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = y[i];
    }
}

// ------------------------------------------------- 4-point DFT across a quad
// DPP quad permutes: xor 1 = quad_perm [1,0,3,2] (0xB1), xor 2 = [2,3,0,1] (0x4E).
__device__ __forceinline__ float dpp_xor1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_xor2(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
}

// DIT form: lane q holds index q; returns index brev2(q).
template <int SGN>
__device__ __forceinline__ float2 quad_dit(float2 v, int q)
{
    const float s2 = (q < 2) ? 1.f : -1.f, s1 = (q & 1) ? -1.f : 1.f;
    float2 a = make_float2(fmaf(s2, v.x, dpp_xor2(v.x)), fmaf(s2, v.y, dpp_xor2(v.y)));
    if (q == 3) a = rot90<SGN>(a);
    return make_float2(fmaf(s1, a.x, dpp_xor1(a.x)), fmaf(s1, a.y, dpp_xor1(a.y)));
}

// DIF form: lane q holds index brev2(q); returns index q.
template <int SGN>
__device__ __forceinline__ float2 quad_dif(float2 v, int q)
{
    const float s2 = (q < 2) ? 1.f : -1.f, s1 = (q & 1) ? -1.f : 1.f;
    float2 a = make_float2(fmaf(s1, v.x, dpp_xor1(v.x)), fmaf(s1, v.y, dpp_xor1(v.y)));
    if (q == 3) a = rot90<SGN>(a);
    return make_float2(fmaf(s2, a.x, dpp_xor2(a.x)), fmaf(s2, a.y, dpp_xor2(a.y)));
}

// 4P-point line transforms over a lane quad (tw: LDS table exp(-2 pi i e / 4P)).
//   dif: lane q holds x[j + P*brev2(q)]  ->  X[4k + q]
//   dit: lane q holds x[4j + q]          ->  X[k + P*brev2(q)]
template <int P, int SGN>
__device__ __forceinline__ void line_dif(float2 (&x)[P], int q, const float2* tw)
{
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = quad_dif<SGN>(x[j], q);
#pragma unroll
    for (int j = 1; j < P; ++j) {
        float2 w = tw[(q * j) & (4 * P - 1)];
        if (SGN > 0) w.y = -w.y;
        x[j] = cmul(x[j], w);
    }
    dft_reg<P, SGN>(x);
}

template <int P, int SGN>
__device__ __forceinline__ void line_dit(float2 (&x)[P], int q, const float2* tw)
{
    dft_reg<P, SGN>(x);
#pragma unroll
    for (int k = 1; k < P; ++k) {
        float2 w = tw[(q * k) & (4 * P - 1)];
        if (SGN > 0) w.y = -w.y;
        x[k] = cmul(x[k], w);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = quad_dit<SGN>(x[k], q);
}

}  // namespace

// ===========================================================================
// Fast engine: N = 4P in {32, 64}; 4N threads per workgroup, one POI pair.
// Thread t: line = t / 4 (row or column being transformed), q = t % 4.
// ===========================================================================
template <int P>
struct FftSmem {
    static constexpr int N = 4 * P;
    // complex plane, element (r, c) at r*(N+1) + c + r/P (row skew spreads
    // the four P-row blocks of a column read over distinct banks)
    float2 z[N * (N + 1) + 4];
    float2 tw[N];
    int red[N / 16][8];
    Top2 top[N / 16][2];
    int imin[N / 16];
};

template <int P>
__device__ __forceinline__ int zidx(int r, int c)
{
    return r * (4 * P + 1) + c + r / P;
}

// Exact integer correlation of the thread's R pixels with the LDS T window
// (bytes, row-major N x N) at shift (ku, kv), reduced over the workgroup.
template <int P>
__device__ __forceinline__ int exact_corr(const uint32_t (&rw)[P / 4], const uint8_t* tb, int line, int bq, int ku,
                                          int kv, FftSmem<P>& sm, int slot)
{
    constexpr int N = 4 * P;
    const int row = ((line + kv) & (N - 1)) * N;
    int acc = 0;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int r = (rw[j >> 2] >> (8 * (j & 3))) & 255;
        acc += r * (int)tb[row + ((bq * P + j + ku) & (N - 1))];
    }
    acc = wave_sum_i(acc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sm.red[wave][slot] = acc;
    __syncthreads();
    int s = 0;
#pragma unroll
    for (int w = 0; w < N / 16; ++w) s += sm.red[w][slot];
    __syncthreads();
    return s;
}

// One workgroup's work: the POIs (pa, ia) and (pb, ib); ib < 0 leaves the
// second slot empty.  Both descriptors must share the window edge N = 4P.
template <int P>
__device__ __forceinline__ void fftcc_pair(FftSmem<P>& sm, const FftccParams& pa, int ia, const FftccParams& pb,
                                           int ib)
{
    constexpr int N = 4 * P, NN = N * N, NW = N / 16;   // waves per workgroup
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int line = tid >> 2, q = tid & 3, bq = ((q & 1) << 1) | (q >> 1);
    const FftccParams* ps[2] = {&pa, &pb};
    const int idx[2] = {ia, ib};

    if (tid < N) sm.tw[tid] = pa.tw[tid];

    // -- 1. load both window pairs straight into registers (DIF layout: row
    //       `line`, columns bq*P .. bq*P+P-1), exact integer sums ---------------
    uint32_t rw[2][P / 4], tw8[2][P / 4];
    int valid[2], sums[2][4];
    int py[2], px[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const FftccParams& p = *ps[s];
        const int poi = idx[s];
        valid[s] = 0;
        py[s] = px[s] = 0;
        if (poi >= 0) {
            py[s] = p.poi[2 * poi];
            px[s] = p.poi[2 * poi + 1];
            const int y0 = py[s] - N / 2, x0 = px[s] - N / 2;
            valid[s] = (y0 >= 0 && x0 >= 0 && y0 + N <= p.img_h && x0 + N <= p.img_w) ? 1 : 2;
        }
#pragma unroll
        for (int k = 0; k < P / 4; ++k) { rw[s][k] = 0u; tw8[s][k] = 0u; }
        if (valid[s] == 1) {
            const size_t o = (size_t)(py[s] - N / 2 + line) * p.stride + (px[s] - N / 2 + bq * P);
            __builtin_memcpy(rw[s], p.ref + o, P);
            __builtin_memcpy(tw8[s], p.tar + o, P);
        }
        int a = 0, b = 0, aa = 0, bb = 0;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const int r = (rw[s][k >> 2] >> (8 * (k & 3))) & 255, t = (tw8[s][k >> 2] >> (8 * (k & 3))) & 255;
            a += r; b += t; aa += r * r; bb += t * t;
        }
        sums[s][0] = wave_sum_i(a); sums[s][1] = wave_sum_i(b);
        sums[s][2] = wave_sum_i(aa); sums[s][3] = wave_sum_i(bb);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) sm.red[wave][4 * s + k] = sums[s][k];
        }
    }
    __syncthreads();
    double nr[2], nt[2];
    float mr[2], mt[2], gr[2], gt[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int acc = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) acc += sm.red[w][4 * s + k];
            sums[s][k] = acc;
        }
        const long long sr = sums[s][0], st = sums[s][1], srr = sums[s][2], stt = sums[s][3];
        // flat window (TW_paDIC_FFTCC2D_CPU.cpp:271-275): exact integer test
        if (valid[s] == 1 && (srr * NN == sr * sr || stt * NN == st * st)) valid[s] = 3;
        nr[s] = (double)srr - (double)sr * (double)sr / (double)NN;
        nt[s] = (double)stt - (double)st * (double)st / (double)NN;
        mr[s] = (float)sr / (float)NN;          // (:253-254)
        mt[s] = (float)st / (float)NN;
        gr[s] = valid[s] == 1 ? (float)(1.0 / sqrt(nr[s])) : 0.f;
        gt[s] = valid[s] == 1 ? (float)(1.0 / sqrt(nt[s])) : 0.f;
    }
    __syncthreads();

    // -- 2. forward transforms and cross-power spectra ---------------------------
    float2 X[2][P];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        float2 x[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int r = (rw[s][j >> 2] >> (8 * (j & 3))) & 255, t = (tw8[s][j >> 2] >> (8 * (j & 3))) & 255;
            x[j] = make_float2(((float)r - mr[s]) * gr[s], ((float)t - mt[s]) * gt[s]);
        }
        line_dif<P, -1>(x, q, sm.tw);                       // rows: lane holds kx = 4j + q
#pragma unroll
        for (int j = 0; j < P; ++j) sm.z[zidx<P>(line, 4 * j + q)] = x[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < P; ++j) x[j] = sm.z[zidx<P>(bq * P + j, line)];
        __syncthreads();
        line_dif<P, -1>(x, q, sm.tw);                       // columns: lane holds ky = 4j + q, kx = line
#pragma unroll
        for (int j = 0; j < P; ++j) sm.z[zidx<P>(4 * j + q, line)] = x[j];
        __syncthreads();
        // conj(F_R) F_T = Im(Z[k] Z[-k]) / 2 - i (|Z[k]|^2 - |Z[-k]|^2) / 4
        const int mkx = (N - line) & (N - 1);
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int mky = (N - (4 * j + q)) & (N - 1);
            const float2 a = x[j], b = sm.z[zidx<P>(mky, mkx)];
            const float im_ab = a.x * b.y + a.y * b.x;
            const float d = (a.x * a.x + a.y * a.y) - (b.x * b.x + b.y * b.y);
            X[s][j] = make_float2(0.5f * im_ab, -0.25f * d);
        }
        __syncthreads();
    }

    // -- 3. one inverse transform for both: IFFT(X_A + i X_B) = c_A + i c_B -------
    float2 x[P];
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = make_float2(X[0][j].x - X[1][j].y, X[0][j].y + X[1][j].x);
    line_dit<P, 1>(x, q, sm.tw);                            // columns: lane holds y = j + P*bq
#pragma unroll
    for (int j = 0; j < P; ++j) sm.z[zidx<P>(j + P * bq, line)] = x[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = sm.z[zidx<P>(line, 4 * j + q)];
    __syncthreads();
    line_dit<P, 1>(x, q, sm.tw);                            // rows: lane holds x = j + P*bq of row `line`

    // -- 4. peaks: top-2 per POI with lowest-index tie-break (:300-311) ----------
    Top2 t[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        t[s] = Top2{-FLT_MAX, 0x7fffffff, -FLT_MAX};
#pragma unroll
        for (int j = 0; j < P; ++j) top2_push(t[s], s == 0 ? x[j].x : x[j].y, line * N + j + P * bq);
        t[s] = wave_top2(t[s]);
        if (lane == 0) sm.top[wave][s] = t[s];
    }
    // T windows to LDS (bytes, row-major) for the exact integer sums; the
    // complex plane is free once every thread holds its outputs.  The window
    // words are re-read (L2-resident) rather than kept live in registers.
    uint8_t* tb = reinterpret_cast<uint8_t*>(sm.z);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (valid[s] == 1) {
            const FftccParams& p = *ps[s];
            const size_t o = (size_t)(py[s] - N / 2 + line) * p.stride + (px[s] - N / 2 + bq * P);
            __builtin_memcpy(rw[s], p.ref + o, P);
            __builtin_memcpy(tw8[s], p.tar + o, P);
        }
#pragma unroll
        for (int k = 0; k < P / 4; ++k)
            reinterpret_cast<uint32_t*>(tb + s * NN + line * N + bq * P)[k] = tw8[s][k];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        Top2 r = sm.top[0][s];
#pragma unroll
        for (int w = 1; w < NW; ++w) r = top2_merge(r, sm.top[w][s]);
        t[s] = r;
    }

    // -- 5. exact peak and ZNCC from integers -------------------------------------
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (valid[s] != 1) continue;                         // uniform per workgroup
        const float thr = t[s].best - pa.tol * (float)NN;
        int peak = t[s].idx, bestI = 0;
        if (t[s].second < thr) {
            bestI = exact_corr<P>(rw[s], tb + s * NN, line, bq, peak % N, peak / N, sm, 0);
        } else {
            // several shifts within tol of the maximum: evaluate them all exactly,
            // in increasing index order, keeping the first strict maximum
            bestI = -1;
            peak = -1;
            int last = -1;
            for (;;) {
                int cand = 0x7fffffff;
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const int idx = line * N + j + P * bq;
                    const float c = s == 0 ? x[j].x : x[j].y;
                    if (idx > last && c >= thr) cand = min(cand, idx);
                }
                cand = wave_min_i(cand);
                if (lane == 0) sm.imin[wave] = cand;
                __syncthreads();
                int m = sm.imin[0];
#pragma unroll
                for (int w = 1; w < NW; ++w) m = min(m, sm.imin[w]);
                __syncthreads();
                if (m == 0x7fffffff) break;
                const int I = exact_corr<P>(rw[s], tb + s * NN, line, bq, m % N, m / N, sm, 0);
                if (I > bestI) { bestI = I; peak = m; }
                last = m;
            }
        }
        if (tid == 0) {
            const FftccParams& p = *ps[s];
            int iu, iv;
            wrap_peak(peak, N, iu, iv);
            const double sr = sums[s][0], st = sums[s][1];
            p.u[idx[s]] = (float)iu;
            p.v[idx[s]] = (float)iv;
            p.zncc[idx[s]] = (float)(((double)bestI - sr * st / (double)NN) / sqrt(nr[s] * nt[s]));
        }
    }
    if (tid == 0) {
#pragma unroll
        for (int s = 0; s < 2; ++s)
            if (valid[s] && ps[s]->status)
                ps[s]->status[idx[s]] = valid[s] == 1 ? TWDIC_FFTCC_OK : valid[s] == 2 ? TWDIC_FFTCC_OUT
                                                                                       : TWDIC_FFTCC_FLAT;
    }
}

template <int P, int MINW>
__global__ __launch_bounds__(16 * P, MINW) void k_fftcc_fft(FftccParams p)
{
    __shared__ FftSmem<P> sm;
    const int ia = 2 * blockIdx.x, ib = ia + 1 < p.n ? ia + 1 : -1;
    fftcc_pair<P>(sm, p, ia, p, ib);
}

// Persistent form: each workgroup takes the next pair of (image pair, POI)
// items of one window class from a device counter until the list is
// exhausted.  Consecutive items may belong to different images.
template <int P, int MINW>
__global__ __launch_bounds__(16 * P, MINW) void k_batch_fftcc(const PairDesc* __restrict__ pairs,
                                                              const int2* __restrict__ items, int total, int* counter)
{
    __shared__ FftSmem<P> sm;
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int k = next;
        __syncthreads();
        if (2 * k >= total) break;
        const int2 a = items[2 * k];
        const int2 b = 2 * k + 1 < total ? items[2 * k + 1] : make_int2(a.x, -1);
        const FftccParams pa = const_pairs(pairs)[__builtin_amdgcn_readfirstlane(a.x)].fftcc;
        const FftccParams pb = const_pairs(pairs)[__builtin_amdgcn_readfirstlane(b.x)].fftcc;
        fftcc_pair<P>(sm, pa, __builtin_amdgcn_readfirstlane(a.y), pb, __builtin_amdgcn_readfirstlane(b.y));
    }
}

// ===========================================================================
// Direct-DFT engine: any even N <= 64, one POI per 256-thread workgroup.
// ===========================================================================
namespace {

constexpr int kDftThreads = 256;
constexpr int kPitch = kMaxFftN + 1;

struct DftSmem {
    float2 z[kMaxFftN * kPitch];
    uint8_t wr[kMaxFftN * kMaxFftN];
    uint8_t wt[kMaxFftN * kMaxFftN];
    float2 tw[kMaxFftN];
    int red[4][4];
    Top2 top[4];
    int imin[4];
};

// In-place DFT of every row (dir 0) or column (dir 1); each wavefront owns
// whole lines and lane k produces output k.  All LDS reads of a line complete
// before its writes within the wavefront.
__device__ void dft_lines(DftSmem& sm, int N, int dir, bool inverse)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int line = wave; line < N; line += 4) {
        float2 acc = make_float2(0.f, 0.f);
        if (lane < N) {
            int e = 0;
            for (int j = 0; j < N; ++j) {
                float2 z = dir == 0 ? sm.z[line * kPitch + j] : sm.z[j * kPitch + line];
                float2 w = sm.tw[e];
                if (inverse) w.y = -w.y;
                acc.x += z.x * w.x - z.y * w.y;
                acc.y += z.x * w.y + z.y * w.x;
                e += lane;
                if (e >= N) e -= N;   // lane < N: one subtraction suffices
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < N) {
            if (dir == 0) sm.z[line * kPitch + lane] = acc;
            else sm.z[lane * kPitch + line] = acc;
        }
    }
}

__device__ int block_sum_i(DftSmem& sm, int v)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_sum_i(v);
    __syncthreads();
    if (lane == 0) sm.red[wave][0] = v;
    __syncthreads();
    return sm.red[0][0] + sm.red[1][0] + sm.red[2][0] + sm.red[3][0];
}

__device__ int exact_corr_dft(DftSmem& sm, int N, int ku, int kv)
{
    int acc = 0;
    for (int i = threadIdx.x; i < N * N; i += kDftThreads) {
        int y = i / N, x = i - y * N;
        int yy = y + kv; if (yy >= N) yy -= N;
        int xx = x + ku; if (xx >= N) xx -= N;
        acc += (int)sm.wr[i] * (int)sm.wt[yy * N + xx];
    }
    return block_sum_i(sm, acc);
}

}  // namespace

__device__ void fftcc_dft_poi(DftSmem& sm, const FftccParams& p, int poi)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int N = p.N, NN = N * N, half = N >> 1;
    const int py = p.poi[2 * poi], px = p.poi[2 * poi + 1];
    const int y0 = py - half, x0 = px - half;
    if (y0 < 0 || x0 < 0 || y0 + N > p.img_h || x0 + N > p.img_w) {
        if (tid == 0 && p.status) p.status[poi] = TWDIC_FFTCC_OUT;
        return;
    }
    int sr = 0, st = 0, srr = 0, stt = 0;
    for (int i = tid; i < NN; i += kDftThreads) {
        int y = i / N, x = i - y * N;
        size_t o = (size_t)(y0 + y) * p.stride + (x0 + x);
        int a = p.ref[o], b = p.tar[o];
        sm.wr[i] = (uint8_t)a;
        sm.wt[i] = (uint8_t)b;
        sr += a; st += b; srr += a * a; stt += b * b;
    }
    if (tid < N) sm.tw[tid] = p.tw[tid];
    sr = wave_sum_i(sr); st = wave_sum_i(st); srr = wave_sum_i(srr); stt = wave_sum_i(stt);
    if (lane == 0) { sm.red[wave][0] = sr; sm.red[wave][1] = st; sm.red[wave][2] = srr; sm.red[wave][3] = stt; }
    __syncthreads();
    const long long Sr = sm.red[0][0] + sm.red[1][0] + sm.red[2][0] + sm.red[3][0];
    const long long St = sm.red[0][1] + sm.red[1][1] + sm.red[2][1] + sm.red[3][1];
    const long long Srr = sm.red[0][2] + sm.red[1][2] + sm.red[2][2] + sm.red[3][2];
    const long long Stt = sm.red[0][3] + sm.red[1][3] + sm.red[2][3] + sm.red[3][3];
    if (Srr * NN == Sr * Sr || Stt * NN == St * St) {         // (:271-275)
        if (tid == 0 && p.status) p.status[poi] = TWDIC_FFTCC_FLAT;
        return;
    }
    const double nr = (double)Srr - (double)Sr * (double)Sr / (double)NN;
    const double nt = (double)Stt - (double)St * (double)St / (double)NN;
    const float mr = (float)Sr / (float)NN, mt = (float)St / (float)NN;
    const float gr = (float)(1.0 / sqrt(nr)), gt = (float)(1.0 / sqrt(nt));
    for (int i = tid; i < NN; i += kDftThreads) {
        int y = i / N, x = i - y * N;
        sm.z[y * kPitch + x] = make_float2(((float)sm.wr[i] - mr) * gr, ((float)sm.wt[i] - mt) * gt);
    }
    __syncthreads();
    dft_lines(sm, N, 0, false);
    __syncthreads();
    dft_lines(sm, N, 1, false);
    __syncthreads();
    for (int i = tid; i < NN; i += kDftThreads) {
        int ky = i / N, kx = i - ky * N;
        int my = ky ? N - ky : 0, mx = kx ? N - kx : 0;
        int j = my * N + mx;
        if (i <= j) {          // each Hermitian pair is formed by one thread
            float2 a = sm.z[ky * kPitch + kx], b = sm.z[my * kPitch + mx];
            float im_ab = a.x * b.y + a.y * b.x;
            float d = (a.x * a.x + a.y * a.y) - (b.x * b.x + b.y * b.y);
            float2 xk = make_float2(0.5f * im_ab, -0.25f * d);
            sm.z[ky * kPitch + kx] = xk;
            if (j != i) sm.z[my * kPitch + mx] = make_float2(xk.x, -xk.y);
        }
    }
    __syncthreads();
    dft_lines(sm, N, 1, true);
    __syncthreads();
    dft_lines(sm, N, 0, true);
    __syncthreads();

    Top2 t{-FLT_MAX, 0x7fffffff, -FLT_MAX};
    for (int i = tid; i < NN; i += kDftThreads) {
        int y = i / N, x = i - y * N;
        t = top2_merge(t, Top2{sm.z[y * kPitch + x].x, i, -FLT_MAX});
    }
    t = wave_top2(t);
    if (lane == 0) sm.top[wave] = t;
    __syncthreads();
    t = top2_merge(top2_merge(sm.top[0], sm.top[1]), top2_merge(sm.top[2], sm.top[3]));
    const float thr = t.best - p.tol * (float)NN;
    int peak = t.idx, bestI;
    if (t.second < thr) {
        bestI = exact_corr_dft(sm, N, peak % N, peak / N);
    } else {
        bestI = -1;
        peak = -1;
        int last = -1;
        for (;;) {
            int cand = 0x7fffffff;
            for (int i = tid; i < NN; i += kDftThreads) {
                int y = i / N, x = i - y * N;
                if (i > last && sm.z[y * kPitch + x].x >= thr) { cand = i; break; }
            }
            cand = wave_min_i(cand);
            __syncthreads();
            if (lane == 0) sm.imin[wave] = cand;
            __syncthreads();
            cand = min(min(sm.imin[0], sm.imin[1]), min(sm.imin[2], sm.imin[3]));
            if (cand == 0x7fffffff) break;
            const int I = exact_corr_dft(sm, N, cand % N, cand / N);
            if (I > bestI) { bestI = I; peak = cand; }
            last = cand;
        }
    }
    if (tid == 0) {
        int iu, iv;
        wrap_peak(peak, N, iu, iv);
        p.u[poi] = (float)iu;
        p.v[poi] = (float)iv;
        p.zncc[poi] = (float)(((double)bestI - (double)Sr * (double)St / (double)NN) / sqrt(nr * nt));
        if (p.status) p.status[poi] = TWDIC_FFTCC_OK;
    }
}

__global__ __launch_bounds__(kDftThreads) void k_fftcc_dft(FftccParams p)
{
    __shared__ DftSmem sm;
    if ((int)blockIdx.x >= p.n) return;
    fftcc_dft_poi(sm, p, blockIdx.x);
}

__global__ __launch_bounds__(kDftThreads) void k_batch_fftcc_dft(const PairDesc* __restrict__ pairs,
                                                                 const int2* __restrict__ items, int total,
                                                                 int* counter)
{
    __shared__ DftSmem sm;
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int k = next;
        __syncthreads();
        if (k >= total) break;
        const int2 it = items[k];
        const FftccParams p = const_pairs(pairs)[__builtin_amdgcn_readfirstlane(it.x)].fftcc;
        fftcc_dft_poi(sm, p, __builtin_amdgcn_readfirstlane(it.y));
        __syncthreads();
    }
}

// ===========================================================================
// Matrix-core engine: any even N <= 64, one wavefront per POI.
//
// The inverse transform of the cross-power spectrum (:285-297) is the
// circular cross-correlation C[kv][ku] = sum_{y,x} R[y][x] T[(y+kv)%N][(x+ku)%N],
// up to the zero-mean constant and rounding.  This engine computes it
// directly, exactly, in integers on the int8 matrix cores
// (v_mfma_i32_32x32x32_i8, or v_mfma_i32_16x16x64_i8 for 32 < N <= 48).
// Stored as signed bytes R' = R - 128 and
// T' = T - 128, C'[k] = C[k] - 128 (sum R + sum T) + 128^2 N^2 differs from
// C[k] by a constant, so the argmax (lowest row-major index on ties) is the
// exact integer peak, and |C'| <= 2^14 N^2 <= 2^26 fits int32.
//
// GEMM form, K = (y', x) over target rows y' and columns x:
//   C'[kv][ku] = sum_{y'} sum_x R'[(y' - kv) mod N][x] * T'[y'][(x + ku) mod N]
// A[kv][(y', x)] gathers reference rows (the row index turns with y');
// B[(y', x)][ku] is the circulant of target row y', read from a doubled copy
// of the row, T2[c] = T'[c mod N], at byte offset x + ku.  Outputs are
// NT x NT (NT = 32 for N <= 32, else 64) in 32 x 32 tiles, or 48 x 48 in
// 16 x 16 tiles; rows and columns of K or of the output beyond N are zero
// padding or ignored.
//
// Lane layout (one wavefront, lane l, m = l % 32, h = l / 32): an A operand
// holds row m, bytes 16h .. 16h+15 of a 32-byte K chunk; a B operand holds
// column m, the same 16 K bytes; accumulator register r of lane l is output
// (row 8 (r / 4) + 4 h + r % 4, column m) of the tile.
// ===========================================================================
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Row strides are multiples of 16 bytes whose dword counts put 8 consecutive
// rows in distinct 4-bank groups, so the row-per-lane 16-byte stores and the
// reference-row reads are conflict-free.
template <int NT, bool N64>
struct MmaSmem {
    static constexpr int RS = NT == 64 ? 80 : 48;     // reference row stride
    static constexpr int TW = N64 ? 96 : 2 * NT;      // bytes of the doubled target row that are read
    static constexpr int TS = N64 ? 112 : NT == 64 ? 144 : 80;   // target row stride
    alignas(16) uint8_t r[NT * RS];
    alignas(16) uint8_t t[NT * TS];
};

// 32 < N <= 48 on 16 x 16 output tiles: 48 rows; the target row is read up
// to byte 111 (offset 16 g + 16 t' + n <= 95, plus 16 bytes)
struct MmaSmem16 {
    static constexpr int RS = 80;
    static constexpr int TW = 112;
    static constexpr int TS = 112;                    // 28 dwords: 8 consecutive rows in distinct bank groups
    alignas(16) uint8_t r[48 * RS];
    alignas(16) uint8_t t[48 * TS];
};

template <int NT, bool N64, int TILE>
struct MmaSel {
    using type = MmaSmem<NT, N64>;
};
template <>
struct MmaSel<64, false, 16> {
    using type = MmaSmem16;
};

// 16 bytes of a doubled target row starting at byte offset o (any alignment)
__device__ __forceinline__ v4i b_operand(const uint32_t* row, int d, int sh)
{
    const uint32_t w0 = row[d], w1 = row[d + 1], w2 = row[d + 2], w3 = row[d + 3], w4 = row[d + 4];
    v4i b;
    b.x = (int)__builtin_amdgcn_alignbyte(w1, w0, sh);
    b.y = (int)__builtin_amdgcn_alignbyte(w2, w1, sh);
    b.z = (int)__builtin_amdgcn_alignbyte(w3, w2, sh);
    b.w = (int)__builtin_amdgcn_alignbyte(w4, w3, sh);
    return b;
}

// The n bytes (2 <= n <= 4 WORDS) at src, any alignment, into w
// (bytes from n on are unspecified).  Only the 16-byte-aligned chunks that
// hold at least one of the n bytes are read; an aligned chunk never crosses a
// page, so nothing beyond the pages of the row itself is touched.  The
// misalignment sh (0..15) is taken out as whole dwords in two select steps,
// then as bytes with v_alignbyte.
template <int WORDS>
__device__ __forceinline__ void load_row(const uint8_t* src, int n, uint32_t (&w)[WORDS])
{
    constexpr int CH = WORDS / 4 + 1;                  // raw words 0 .. WORDS + 3 are used
    const uintptr_t a = reinterpret_cast<uintptr_t>(src);
    const uint4* c = reinterpret_cast<const uint4*>(a & ~(uintptr_t)15);
    const int sh = (int)(a & 15);
    uint32_t raw[4 * CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (16 * k < sh + n) v = c[k];
        raw[4 * k] = v.x;
        raw[4 * k + 1] = v.y;
        raw[4 * k + 2] = v.z;
        raw[4 * k + 3] = v.w;
    }
    // bitwise selects on opaque masks (v_bfi_b32): a `q ? raw[i + 2] : raw[i]`
    // is folded into a dynamically indexed load, which moves raw to scratch
    uint32_t m2 = (sh & 8) ? ~0u : 0u, m1 = (sh & 4) ? ~0u : 0u;
    asm volatile("" : "+v"(m2), "+v"(m1));
    const int b = sh & 3;
    uint32_t s2[WORDS + 2], s1[WORDS + 1];
#pragma unroll
    for (int i = 0; i < WORDS + 2; ++i) s2[i] = (raw[i + 2] & m2) | (raw[i] & ~m2);
#pragma unroll
    for (int i = 0; i < WORDS + 1; ++i) s1[i] = (s2[i + 1] & m1) | (s2[i] & ~m1);
#pragma unroll
    for (int k = 0; k < WORDS; ++k) w[k] = __builtin_amdgcn_alignbyte(s1[k + 1], s1[k], b);
}

// N64: N == 64 (NT = 64); otherwise N <= NT is a run-time value.  TILE 32:
// v_mfma_i32_32x32x32_i8 on 32 x 32 tiles (N <= 32 in one tile, 32 < N <= 64
// in 2 x 2); TILE 16: v_mfma_i32_16x16x64_i8 on 3 x 3 tiles of 16 x 16 with
// the whole 64-byte row as one K chunk, for 32 < N <= 48, where it issues
// 9 16-cycle MFMAs per row instead of 8 32-cycle ones.
template <int NT, bool N64, int TILE>
__device__ __forceinline__ void fftcc_mma_poi(typename MmaSel<NT, N64, TILE>::type& sm, const FftccParams& p,
                                              int poi)
{
    constexpr int TT = NT / 32;          // output tiles per edge and K chunks per row (TILE 32)
    using Sm = typename MmaSel<NT, N64, TILE>::type;
    const int lane = threadIdx.x, m = lane & 31, h = lane >> 5;
    const int N = N64 ? 64 : p.N, NN = N * N;
    const int py = p.poi[2 * poi], px = p.poi[2 * poi + 1];
    if (py - N / 2 < 0 || px - N / 2 < 0 || py + N / 2 > p.img_h || px + N / 2 > p.img_w) {
        if (lane == 0 && p.status) p.status[poi] = TWDIC_FFTCC_OUT;
        return;
    }

    // -- window rows to LDS (lane = row), exact integer sums -------------------
    int sr = 0, st = 0, srr = 0, stt = 0;
    if (lane < N) {
        const size_t o = (size_t)(py - N / 2 + lane) * p.stride + (px - N / 2);
        uint8_t* rrow = sm.r + lane * Sm::RS;
        uint8_t* trow = sm.t + lane * Sm::TS;
        if (N == NT) {
            uint32_t r[NT / 4], t[NT / 4];
            __builtin_memcpy(r, p.ref + o, NT);
            __builtin_memcpy(t, p.tar + o, NT);
#pragma unroll
            for (int k = 0; k < NT / 4; ++k) {
                sr = (int)__builtin_amdgcn_udot4(r[k], 0x01010101u, (uint32_t)sr, false);
                st = (int)__builtin_amdgcn_udot4(t[k], 0x01010101u, (uint32_t)st, false);
                srr = (int)__builtin_amdgcn_udot4(r[k], r[k], (uint32_t)srr, false);
                stt = (int)__builtin_amdgcn_udot4(t[k], t[k], (uint32_t)stt, false);
            }
            uint4* r4 = reinterpret_cast<uint4*>(rrow);
            uint4* t4 = reinterpret_cast<uint4*>(trow);
#pragma unroll
            for (int k = 0; k < NT / 16; ++k)         // signed bytes
                r4[k] = make_uint4(r[4 * k] ^ 0x80808080u, r[4 * k + 1] ^ 0x80808080u, r[4 * k + 2] ^ 0x80808080u,
                                   r[4 * k + 3] ^ 0x80808080u);
#pragma unroll
            for (int k = 0; k < Sm::TW / 16; ++k) {   // the target row, repeated
                const int q = k % (NT / 16);
                t4[k] = make_uint4(t[4 * q] ^ 0x80808080u, t[4 * q + 1] ^ 0x80808080u, t[4 * q + 2] ^ 0x80808080u,
                                   t[4 * q + 3] ^ 0x80808080u);
            }
        } else {
            // any even N < NT: the rows as words (load_row), signed, reference
            // bytes x >= N zeroed, then the sums from the signed words with
            // explicit dot products.  (A guarded byte sum `if (x < N) s += a * a`
            // was folded by the compiler into a v_dot4 with a wrong accumulator,
            // adding 1 per row.)
            constexpr int WORDS = NT == 32 ? 8 : TILE == 16 ? 12 : 16;   // N <= 30, 48, 62
            uint32_t r[WORDS], t[WORDS];
            load_row<WORDS>(p.ref + o, N, r);
            load_row<WORDS>(p.tar + o, N, t);
            int s1r = 0, s2r = 0, s1t = 0, s2t = 0;
#pragma unroll
            for (int k = 0; k < WORDS; ++k) {
                const int x = 4 * k;
                const uint32_t keep = x + 4 <= N ? ~0u : x >= N ? 0u : 0xffffu;   // N is even
                r[k] = (r[k] ^ 0x80808080u) & keep;
                t[k] ^= 0x80808080u;
                const int wr = (int)r[k], wt = (int)(t[k] & keep);
                s1r = __builtin_amdgcn_sdot4(wr, 0x01010101, s1r, false);
                s2r = __builtin_amdgcn_sdot4(wr, wr, s2r, false);
                s1t = __builtin_amdgcn_sdot4(wt, 0x01010101, s1t, false);
                s2t = __builtin_amdgcn_sdot4(wt, wt, s2t, false);
            }
            uint4* r4 = reinterpret_cast<uint4*>(rrow);
#pragma unroll
            for (int k = 0; k < NT / 16; ++k)
                r4[k] = 4 * k < WORDS ? make_uint4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3])
                                      : make_uint4(0u, 0u, 0u, 0u);
            // the target row at byte 0 and again at byte N.  Products that count
            // read bytes x + ku <= 2N - 2 (x, ku < N); later bytes meet zero
            // reference bytes (x >= N) or output columns ku >= N, which are ignored.
            uint4* t4 = reinterpret_cast<uint4*>(trow);
#pragma unroll
            for (int k = 0; k < WORDS / 4; ++k) t4[k] = make_uint4(t[4 * k], t[4 * k + 1], t[4 * k + 2], t[4 * k + 3]);
            uint32_t* t32 = reinterpret_cast<uint32_t*>(trow) + (N >> 2);
            if ((N & 2) == 0) {
#pragma unroll
                for (int e = 0; e < WORDS; ++e) t32[e] = t[e];
            } else {                                       // N = 4 m + 2: bytes 0, 1 at N, then shifted words
                *reinterpret_cast<uint16_t*>(trow + N) = (uint16_t)t[0];
#pragma unroll
                for (int e = 0; e + 1 < WORDS; ++e) t32[1 + e] = __builtin_amdgcn_alignbyte(t[e + 1], t[e], 2);
            }
            // bytes are v - 128: sum v = s1 + 128 N, sum v^2 = s2 + 256 s1 + 16384 N
            sr = s1r + 128 * N;
            srr = s2r + 256 * s1r + 16384 * N;
            st = s1t + 128 * N;
            stt = s2t + 256 * s1t + 16384 * N;
        }
    }
    sr = wave_sum_i(sr); st = wave_sum_i(st); srr = wave_sum_i(srr); stt = wave_sum_i(stt);
    const long long Sr = sr, St = st, Srr = srr, Stt = stt;
    if (Srr * NN == Sr * Sr || Stt * NN == St * St) {          // flat window (:271-275)
        if (lane == 0 && p.status) p.status[poi] = TWDIC_FFTCC_FLAT;
        return;
    }
    __syncthreads();

    int best = INT_MIN, bk = INT_MAX;
    if constexpr (TILE == 32) {
        // -- C' = A B over K = N rows x TT chunks of 32 bytes ------------------------
        int ra[TT];                                        // A row byte offsets, (y' - kv) mod N
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            const int kv = 32 * t + m;
            ra[t] = (kv < N ? (kv == 0 ? 0 : N - kv) : 0) * Sm::RS + 16 * h;
        }
        // B operand j (j = K chunk + output tile) starts at byte 32 j + 16 h + m
        constexpr int NB = TT == 1 ? 1 : N64 ? 2 : 3;      // distinct B operands per row
        int bd[NB], bs[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int o = 32 * j + 16 * h + m;
            bd[j] = o >> 2;
            bs[j] = o & 3;
        }
        const int rwrap = N * Sm::RS;
        v16i acc[TT][TT];
#pragma unroll
        for (int a = 0; a < TT; ++a)
#pragma unroll
            for (int b = 0; b < TT; ++b) acc[a][b] = v16i{};

        // A operands of row y': lane (m, h) holds reference row (y' - kv) mod N,
        // read from LDS for the next row while the MFMAs of this one run.  (Passing
        // them down one lane per row instead, with a DPP shift and two LDS reads,
        // measured 0.8 % slower: the kernel is power-limited, not LDS-bound.)
        v4i A[TT][TT];                                     // [K chunk][output row tile]
#pragma unroll
        for (int xc = 0; xc < TT; ++xc)
#pragma unroll
            for (int t = 0; t < TT; ++t) A[xc][t] = *reinterpret_cast<const v4i*>(sm.r + ra[t] + 32 * xc);

#pragma unroll 2
        for (int y = 0; y < N; ++y) {
            const uint32_t* trow = reinterpret_cast<const uint32_t*>(sm.t + y * Sm::TS);
            v4i B[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) B[j] = b_operand(trow, bd[j], bs[j]);
#pragma unroll
            for (int xc = 0; xc < TT; ++xc)
#pragma unroll
                for (int t = 0; t < TT; ++t)
#pragma unroll
                    for (int t2 = 0; t2 < TT; ++t2)     // offset 64 is offset 0 when N = 64
                        acc[t][t2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[xc][t], B[(xc + t2) % NB], acc[t][t2], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < TT; ++t) {
                ra[t] += Sm::RS;
                if (ra[t] >= rwrap) ra[t] -= rwrap;
            }
#pragma unroll
            for (int xc = 0; xc < TT; ++xc)
#pragma unroll
                for (int t = 0; t < TT; ++t) A[xc][t] = *reinterpret_cast<const v4i*>(sm.r + ra[t] + 32 * xc);
        }

        // -- exact peak: largest C', lowest row-major index on ties (:300-311) -------
        // a lane's values are visited in increasing index order (kv, then ku), so
        // a strict comparison keeps the lowest index among equal values
#pragma unroll
        for (int t = 0; t < TT; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r)
#pragma unroll
                for (int t2 = 0; t2 < TT; ++t2) {
                    const int kv = 32 * t + 8 * (r >> 2) + 4 * h + (r & 3), ku = 32 * t2 + m;
                    const int v = acc[t][t2][r];
                    const bool take = kv < N && ku < N && v > best;
                    best = take ? v : best;
                    bk = take ? kv * N + ku : bk;
                }
    } else {
        // -- 16 x 16 tiles: lane (n, g), n = lane % 16, g = lane / 16 ----------
        // A tile t: row (y' - kv) mod N, kv = 16 t + n, bytes 16 g .. 16 g + 15
        // of the 64-byte row (zero beyond N); B tile t': the target row at
        // byte offset 16 g + 16 t' + n; accumulator register r of lane l is
        // output (row 4 g + r, column n) of its tile.
        const int n16 = lane & 15, g = lane >> 4;
        int ra[3], bd[3], bs[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int kv = 16 * t + n16;
            ra[t] = (kv < N ? (kv == 0 ? 0 : N - kv) : 0) * Sm::RS + 16 * g;
            const int o = 16 * g + 16 * t + n16;
            bd[t] = o >> 2;
            bs[t] = o & 3;
        }
        const int rwrap = N * Sm::RS;
        v4i acc[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[a][c] = v4i{};
        v4i A[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) A[t] = *reinterpret_cast<const v4i*>(sm.r + ra[t]);
        for (int y = 0; y < N; ++y) {
            const uint32_t* trow = reinterpret_cast<const uint32_t*>(sm.t + y * Sm::TS);
            v4i B[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) B[t] = b_operand(trow, bd[t], bs[t]);
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int t2 = 0; t2 < 3; ++t2)
                    acc[t][t2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[t], B[t2], acc[t][t2], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                ra[t] += Sm::RS;
                if (ra[t] >= rwrap) ra[t] -= rwrap;
                A[t] = *reinterpret_cast<const v4i*>(sm.r + ra[t]);
            }
        }
        // increasing index order within the lane: kv = 16 t + 4 g + r, then ku
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int t2 = 0; t2 < 3; ++t2) {
                    const int kv = 16 * t + 4 * g + r, ku = 16 * t2 + n16;
                    const int v = acc[t][t2][r];
                    const bool take = kv < N && ku < N && v > best;
                    best = take ? v : best;
                    bk = take ? kv * N + ku : bk;
                }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int ov = __shfl_xor(best, o), ok = __shfl_xor(bk, o);
        const bool take = ov > best || (ov == best && ok < bk);
        best = take ? ov : best;
        bk = take ? ok : bk;
    }
    if (lane == 0) {
        const long long I = (long long)best + 128LL * (Sr + St) - 16384LL * NN;   // sum R T at the peak
        const double nr = (double)Srr - (double)Sr * (double)Sr / (double)NN;
        const double nt = (double)Stt - (double)St * (double)St / (double)NN;
        int iu, iv;
        wrap_peak(bk, N, iu, iv);
        p.u[poi] = (float)iu;
        p.v[poi] = (float)iv;
        p.zncc[poi] = (float)(((double)I - (double)Sr * (double)St / (double)NN) / sqrt(nr * nt));
        if (p.status) p.status[poi] = TWDIC_FFTCC_OK;
    }
}

template <int NT, bool N64, int TILE>
__global__ __launch_bounds__(64) void k_fftcc_mma(FftccParams p)
{
    __shared__ typename MmaSel<NT, N64, TILE>::type sm;
    fftcc_mma_poi<NT, N64, TILE>(sm, p, blockIdx.x);
}

// Batched form, one launch over the POIs of many image pairs that share a
// kernel form: block b takes item b, (pair, POI), and that pair's descriptor.
template <int NT, bool N64, int TILE>
__global__ __launch_bounds__(64) void k_batch_fftcc_mma(const PairDesc* __restrict__ pairs,
                                                        const int2* __restrict__ items)
{
    __shared__ typename MmaSel<NT, N64, TILE>::type sm;
    const int2 it = items[blockIdx.x];
    const FftccParams p = const_pairs(pairs)[__builtin_amdgcn_readfirstlane(it.x)].fftcc;
    fftcc_mma_poi<NT, N64, TILE>(sm, p, __builtin_amdgcn_readfirstlane(it.y));
}

// ===========================================================================
// Matrix-core DFT engine, N = 64: the correlation surface from fp16 2-D DFTs
// on v_mfma_f32_32x32x16_f16, then the exact integer correlation at every
// shift the surface leaves within its proven error of the maximum.
//
//   r = R - a, t = T - b with a, b the rounded window means: integers in
//   [-255, 255], exact in fp16; the correlation of r and t differs from I[k]
//   = sum R T(k) by a constant, so it has the same argmax.
//   Forward: R^ = F r F, T^ = F t F (F[j][y] = w^(j y), w = exp(-2 pi i / 64),
//   rounded to fp16), two passes of fp16 MFMAs with fp32 accumulation, the
//   intermediate rounded to fp16.  r and t are real, so only the columns k =
//   0 .. 32 are needed (column 64 - k is the conjugate of column k, rows
//   negated): the first pass computes k = 0 .. 31 with column 0 carrying Y[.][0]
//   + i Y[.][32] (its table column is 1 + i (-1)^x, exact), and after the
//   second pass W[j] = R^[j][0] + i R^[j][32] is unpacked with its partner
//   row -j: R^[j][0] = (W[j] + conj W[-j]) / 2, R^[j][32] = (W[j] - conj
//   W[-j]) / 2i.  P = conj(R^) T^ s in fp32 (s a power of two putting |P|
//   below 2^14), rounded to fp16, column 0 repacked as P[.][0] + i P[.][32].
//   Inverse: Q = P^T F* for k = 0 .. 31 (scaled by 2^-8, rows k >= 1 doubled
//   for their conjugates), row 0 giving Q[0][a] + i Q[32][a]; then C^T =
//   Re(F* Q) + (-1)^b Q[32][a], the last term through the imaginary table
//   part of k = 0 (zero in F, (-1)^b here): 144 MFMAs per POI against the
//   integer engine's 512 of the same length.
//   Layouts: a 32 x 32 x 16 accumulator holds column n in lane n (+32 h)
//   and rows 8 (r / 4) + 4 h + r % 4 in register r, so registers 8 c' .. 8 c'
//   + 7 of a tile are a ready B (or, for a transposed product, A) operand
//   whose 16 K values are those rows in that order; the other operand, the
//   fp16 DFT matrix, is stored with its K index in the same order (table 1),
//   so no pass transposes through LDS.
//   Error: |C^ - C| <= kDftBound ||r|| ||t|| / N^2 everywhere.  The bound
//   follows from ||fp16(F) - F||_2 = 1.29e-3, the largest entry error of
//   fp16(F) (1.73e-4), the fp16 and fp32 unit roundoffs, the worst-case error
//   of an fp32 sum of n terms (gamma_{n-1} sum |terms|) for each pass and
//   Cauchy-Schwarz; completing the half spectrum by conjugation turns the
//   computation into the full-spectrum one with one set of errors.  It is
//   15.1 with fp16 sub-normals kept and 21.7 if they were flushed to zero;
//   kDftBound covers both, about sixteen times the largest error of a
//   matrix-level model of this arithmetic on stress windows.  Every shift
//   whose approximate value is within twice that (x 1.25) of the approximate
//   maximum is evaluated exactly in integers, in increasing index order,
//   keeping the first strict maximum: the exact argmax with the CPU path's
//   lowest-index tie-break.  (Speckle windows leave about two such shifts; a
//   window of pure noise a handful; the count is not capped.)
//   A workgroup of eight wavefronts, one POI each, shares the fp16 tables in
//   LDS; the window bytes stay in LDS for the exact stage.
// ===========================================================================
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef float f16x __attribute__((ext_vector_type(16)));

constexpr double kDftBound = 22.0;        // N^2 |C^ - C| / (||r|| ||t||), see above
// LDS pitch of the R rows (64 B) and the doubled T rows (128 B): 18 and 34
// dwords put the 8-byte operand reads of 16 consecutive rows on distinct banks
constexpr int kExR = 72, kExT = 136;
constexpr int kDftWaves = 8;              // POIs in flight per workgroup (one per wavefront)

// F in fp16 as MFMA operands, [table][part][tile][K chunk][lane], lane = 32 h + m,
// parts re, im, -im:
//   table 0: F[16 c + 8 h + q][32 t + m]                 (B operand, K in natural order);
//            in tile 0 the imaginary part of column 0 holds (-1)^x instead of 0
//   table 1: 2^-8 F[32 t + m][16 c + 8 (q / 4) + 4 h + q % 4]   (A operand, K in accumulator
//            row order; the 2^-8, exact, keeps the first inverse pass's output in fp16 range)
__device__ h8 g_dft_tab[2][3][2][4][64];
// the last pass's imaginary table part for K chunk 0 with the k = 0 entry
// (zero in F) set to 2^-8 (-1)^b: the Nyquist row Q[32][a], carried in the
// imaginary part of row 0, then enters C^T[b][a] with its sign (-1)^b
__device__ h8 g_dft_nyq[2][64];
int g_dft_cus[64] = {};                        // compute units per device, set by dft_tables_init

// One workgroup: the tables (32 KB, read by all its wavefronts) and, per
// wavefront, its POI's R rows and doubled T rows as bytes and the second
// pass's spectrum column 0 (for the j <-> -j unpacking).
struct DftWgSmem {
    h8 ta[2][4][64];
    h8 tb[3][2][4][64];
    h8 tn[2][64];
    uint8_t ex[kDftWaves][64 * kExR + 64 * kExT];
    float col0[kDftWaves][4][64];               // Re R^, Im R^, Re T^, Im T^ of column 0, by row j
    uint32_t xcol[kDftWaves][64];
};

__device__ __forceinline__ f16x mfma16(h8 a, h8 b, f16x c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// accumulator registers 8 half .. 8 half + 7, times scale, rounded to fp16
__device__ __forceinline__ h8 to_h8(const f16x& acc, int half, float scale)
{
    h8 r;
#pragma unroll
    for (int q = 0; q < 8; ++q) r[q] = (_Float16)(acc[8 * half + q] * scale);
    return r;
}

// LDS written by lanes of this wavefront and then read by other lanes of it:
// its LDS accesses complete in issue order; keep the compiler from moving
// them across this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-uniform reductions for the DFT engine: butterflies inside each row of
// 16 lanes by DPP (xor 1, xor 2, half-row mirror, row mirror), then the four
// row results by readlane.  Every lane must be active.
template <class Op>
__device__ __forceinline__ int wave_reduce_dpp(int v, Op op)
{
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));
    const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return op(op(a, b), op(c, d));
}
__device__ __forceinline__ int dpp_sum(int v)
{
    return wave_reduce_dpp(v, [](int x, int y) { return x + y; });
}
__device__ __forceinline__ int dpp_min(int v)
{
    return wave_reduce_dpp(v, [](int x, int y) { return min(x, y); });
}
__device__ __forceinline__ float dpp_max(float v)
{
    return __int_as_float(wave_reduce_dpp(__float_as_int(v), [](int x, int y) {
        return __float_as_int(fmaxf(__int_as_float(x), __int_as_float(y)));
    }));
}

// 2 mask + (x >= thr): the compare's lane mask is the carry-in of an add,
// two instructions per bit
__device__ __forceinline__ uint32_t shift_in_ge(uint32_t mask, float x, float thr)
{
    uint32_t r;
    asm("v_cmp_ge_f32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %3, %3, vcc"
        : "=v"(r)
        : "v"(x), "v"(thr), "v"(mask)
        : "vcc");
    return r;
}

// Eight window bytes as the fp16 values byte - off: the byte v becomes the
// half 1024 + v (bits 0x6400 | v), minus off = 1024 + mean, exactly.  The
// byte shuffle is one v_perm_b32 per two values (selector bytes 0-3 take the
// bytes of the second source, 4-7 those of the first: 0x64).
__device__ __forceinline__ h8 centred_h8(uint2 w, h2v off)
{
    const uint32_t q[4] = {__builtin_amdgcn_perm(0x64646464u, w.x, 0x04010400u),
                           __builtin_amdgcn_perm(0x64646464u, w.x, 0x04030402u),
                           __builtin_amdgcn_perm(0x64646464u, w.y, 0x04010400u),
                           __builtin_amdgcn_perm(0x64646464u, w.y, 0x04030402u)};
    h8 r;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const h2v v = __builtin_bit_cast(h2v, q[k]) - off;
        r[2 * k] = v[0];
        r[2 * k + 1] = v[1];
    }
    return r;
}

__device__ void fftcc_dft_mfma_poi(DftWgSmem& sm, int wave, int lane, const FftccParams& p, int poi)
{
    constexpr int N = 64, NN = 4096;
    const int m = lane & 31, h = lane >> 5;
    const int py = p.poi[2 * poi], px = p.poi[2 * poi + 1];
    if (py - N / 2 < 0 || px - N / 2 < 0 || py + N / 2 > p.img_h || px + N / 2 > p.img_w) {
        if (lane == 0 && p.status) p.status[poi] = TWDIC_FFTCC_OUT;
        return;
    }
    uint8_t* ex = sm.ex[wave];
    const uint8_t* erows = ex;                 // R row y at y kExR
    const uint8_t* etrows = ex + 64 * kExR;    // T row y at y kExT, twice
    float (*col0)[64] = sm.col0[wave];
    uint32_t* xcol = sm.xcol[wave];
    // -- window rows (lane = row) to LDS, exact integer sums -------------------
    int sr = 0, st = 0, srr = 0, stt = 0;
    {
        const size_t o = (size_t)(py - N / 2 + lane) * p.stride + (px - N / 2);
        uint32_t rw[16], tw[16];
        __builtin_memcpy(rw, p.ref + o, 64);
        __builtin_memcpy(tw, p.tar + o, 64);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            sr = (int)__builtin_amdgcn_udot4(rw[k], 0x01010101u, (uint32_t)sr, false);
            st = (int)__builtin_amdgcn_udot4(tw[k], 0x01010101u, (uint32_t)st, false);
            srr = (int)__builtin_amdgcn_udot4(rw[k], rw[k], (uint32_t)srr, false);
            stt = (int)__builtin_amdgcn_udot4(tw[k], tw[k], (uint32_t)stt, false);
        }
        wave_lds_sync();                       // the previous POI's reads of this region are done
        uint2* er = reinterpret_cast<uint2*>(ex + lane * kExR);
        uint2* et = reinterpret_cast<uint2*>(ex + 64 * kExR + lane * kExT);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint2 a = make_uint2(rw[2 * k], rw[2 * k + 1]);
            const uint2 b = make_uint2(tw[2 * k], tw[2 * k + 1]);
            er[k] = a;
            et[k] = b;
            et[8 + k] = b;
        }
    }
    sr = dpp_sum(sr); st = dpp_sum(st); srr = dpp_sum(srr); stt = dpp_sum(stt);
    const long long Sr = sr, St = st, Srr = srr, Stt = stt;
    if (Srr * NN == Sr * Sr || Stt * NN == St * St) {          // flat window (:271-275)
        if (lane == 0 && p.status) p.status[poi] = TWDIC_FFTCC_FLAT;
        return;
    }
    const int a = (sr + NN / 2) >> 12, b = (st + NN / 2) >> 12;    // rounded means
    const h2v oa = {(_Float16)(1024 + a), (_Float16)(1024 + a)};
    const h2v ob = {(_Float16)(1024 + b), (_Float16)(1024 + b)};
    const double nr2 = (double)(Srr - 2LL * a * Sr + (long long)a * a * NN);     // ||r||^2, exact
    const double nt2 = (double)(Stt - 2LL * b * St + (long long)b * b * NN);
    const double rho = sqrt(nr2 * nt2);
    int ex2;
    (void)frexp(4096.0 * rho, &ex2);
    const float s = ldexpf(1.0f, 14 - ex2);                                      // s 4096 rho in [2^13, 2^14)
    const float tol = (float)(2.0 * 1.25 * kDftBound * rho * (double)s / 65536.0); // in C^ units (s 2^-16 N^2 C)
    wave_lds_sync();

    // -- first pass, columns k = 0 .. 31 (column 0 = Y[.][0] + i Y[.][32]) -------
    h8 Yrr[4], Yri[4], Ytr[4], Yti[4];         // B operands, K chunk c = rows 16 c .. 16 c + 15 (accumulator order)
#pragma unroll
    for (int Mt = 0; Mt < 2; ++Mt) {
        f16x yrr = {}, yri = {}, ytr = {}, yti = {};
        const int y = 32 * Mt + m;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const h8 Fr = sm.ta[0][c][lane], Fi = sm.ta[1][c][lane];
            const h8 Ar = centred_h8(*reinterpret_cast<const uint2*>(erows + y * kExR + 16 * c + 8 * h), oa);
            const h8 At = centred_h8(*reinterpret_cast<const uint2*>(etrows + y * kExT + 16 * c + 8 * h), ob);
            yrr = mfma16(Ar, Fr, yrr);
            yri = mfma16(Ar, Fi, yri);
            ytr = mfma16(At, Fr, ytr);
            yti = mfma16(At, Fi, yti);
        }
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            Yrr[2 * Mt + hf] = to_h8(yrr, hf, 1.f);
            Yri[2 * Mt + hf] = to_h8(yri, hf, 1.f);
            Ytr[2 * Mt + hf] = to_h8(ytr, hf, 1.f);
            Yti[2 * Mt + hf] = to_h8(yti, hf, 1.f);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    // -- second pass: R^, T^ for k = 0 .. 31 and P = conj(R^) T^ s ----------------
    h8 Pre[4], Pim[4];                         // A operands of the first inverse pass [j chunk]
#pragma unroll
    for (int Mj = 0; Mj < 2; ++Mj) {
        f16x Rr = {}, Ri = {}, Tr = {}, Ti = {};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const h8 Fr = sm.tb[0][Mj][c][lane], Fi = sm.tb[1][Mj][c][lane], nFi = sm.tb[2][Mj][c][lane];
            Rr = mfma16(Fr, Yrr[c], Rr);
            Ri = mfma16(Fr, Yri[c], Ri);
            Tr = mfma16(Fr, Ytr[c], Tr);
            Ti = mfma16(Fr, Yti[c], Ti);
            Rr = mfma16(nFi, Yri[c], Rr);
            Ri = mfma16(Fi, Yrr[c], Ri);
            Tr = mfma16(nFi, Yti[c], Tr);
            Ti = mfma16(Fi, Ytr[c], Ti);
        }
        if (m == 0) {                          // column 0: W_R = R^[j][0] + i R^[j][32], W_T likewise
#pragma unroll
            for (int g = 0; g < 4; ++g) {      // registers 4 g .. 4 g + 3 hold rows j0 .. j0 + 3
                const int j0 = 32 * Mj + 8 * g + 4 * h;
                *reinterpret_cast<float4*>(&col0[0][j0]) = make_float4(Rr[4 * g], Rr[4 * g + 1], Rr[4 * g + 2], Rr[4 * g + 3]);
                *reinterpret_cast<float4*>(&col0[1][j0]) = make_float4(Ri[4 * g], Ri[4 * g + 1], Ri[4 * g + 2], Ri[4 * g + 3]);
                *reinterpret_cast<float4*>(&col0[2][j0]) = make_float4(Tr[4 * g], Tr[4 * g + 1], Tr[4 * g + 2], Tr[4 * g + 3]);
                *reinterpret_cast<float4*>(&col0[3][j0]) = make_float4(Ti[4 * g], Ti[4 * g + 1], Ti[4 * g + 2], Ti[4 * g + 3]);
            }
        }
        f16x pr, pi;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            pr[i] = fmaf(Rr[i], Tr[i], Ri[i] * Ti[i]);
            pi[i] = fmaf(Rr[i], Ti[i], -(Ri[i] * Tr[i]));
        }
        // columns k >= 1 doubled for their conjugates 64 - k (column 0 is
        // rewritten below); 2^16 undoes the two 2^-8 of the table
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            Pre[2 * Mj + hf] = to_h8(pr, hf, 131072.f * s);
            Pim[2 * Mj + hf] = to_h8(pi, hf, 131072.f * s);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    // column 0 of P becomes P[j][0] + i P[j][32]: unpack R^[j][0] = (W[j] +
    // conj W[-j]) / 2, R^[j][32] = (W[j] - conj W[-j]) / 2i, likewise T^; lane
    // j does row j (the results for j and -j are exact conjugates), lanes m = 0
    // collect theirs
    wave_lds_sync();
    {
        const int jm = (64 - lane) & 63;
        const float4 w = make_float4(col0[0][lane], col0[1][lane], col0[2][lane], col0[3][lane]);
        const float4 wm = make_float4(col0[0][jm], col0[1][jm], col0[2][jm], col0[3][jm]);
        const float ar = 0.5f * (w.x + wm.x), ai = 0.5f * (w.y - wm.y);    // R^[j][0]
        const float br = 0.5f * (w.y + wm.y), bi = 0.5f * (wm.x - w.x);    // R^[j][32]
        const float cr = 0.5f * (w.z + wm.z), ci = 0.5f * (w.w - wm.w);    // T^[j][0]
        const float dr = 0.5f * (w.w + wm.w), di = 0.5f * (wm.z - w.z);    // T^[j][32]
        const float p0r = fmaf(ar, cr, ai * ci), p0i = fmaf(ar, ci, -(ai * cr));   // conj(R^0) T^0
        const float p1r = fmaf(br, dr, bi * di), p1i = fmaf(br, di, -(bi * dr));   // conj(R^32) T^32
        const h2v x = {(_Float16)((p0r - p1i) * (65536.f * s)), (_Float16)((p0i + p1r) * (65536.f * s))};
        xcol[lane] = __builtin_bit_cast(uint32_t, x);
    }
    wave_lds_sync();
    if (m == 0) {
#pragma unroll
        for (int Mj = 0; Mj < 2; ++Mj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const h2v x = __builtin_bit_cast(h2v, xcol[32 * Mj + 8 * (r >> 2) + 4 * h + (r & 3)]);
                Pre[2 * Mj + (r >> 3)][r & 7] = x[0];
                Pim[2 * Mj + (r >> 3)][r & 7] = x[1];
            }
    }
    // -- first inverse pass, rows k = 0 .. 31: Q[k][a] = sum_j P[j][k] conj(F[j][a])
    //    (row 0 comes out as Q[0][a] + i Q[32][a], both real) -----------------------
    h8 Qr[2][2], Qi[2][2];                     // B operands of the last pass [k chunk][a tile]
    {
        f16x qr[2] = {}, qi[2] = {};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int Nc = 0; Nc < 2; ++Nc) {
                const h8 Fr = sm.tb[0][Nc][c][lane], Fi = sm.tb[1][Nc][c][lane], nFi = sm.tb[2][Nc][c][lane];
                qr[Nc] = mfma16(Pre[c], Fr, qr[Nc]);
                qi[Nc] = mfma16(Pim[c], Fr, qi[Nc]);
                qr[Nc] = mfma16(Pim[c], Fi, qr[Nc]);
                qi[Nc] = mfma16(Pre[c], nFi, qi[Nc]);
            }
        }
#pragma unroll
        for (int Nc = 0; Nc < 2; ++Nc) {
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                Qr[hf][Nc] = to_h8(qr[Nc], hf, 1.f);
                Qi[hf][Nc] = to_h8(qi[Nc], hf, 1.f);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    // -- last pass: C^T = Re(F* Q) + (-1)^b Q[32][a], lane -> a, register -> b ------
    f16x Cd[2][2] = {};                        // [b tile][a tile]
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int Md = 0; Md < 2; ++Md) {
            const h8 Fr = sm.tb[0][Md][c][lane], Fi = c == 0 ? sm.tn[Md][lane] : sm.tb[1][Md][c][lane];
#pragma unroll
            for (int Nd = 0; Nd < 2; ++Nd) {
                Cd[Md][Nd] = mfma16(Fr, Qr[c][Nd], Cd[Md][Nd]);
                Cd[Md][Nd] = mfma16(Fi, Qi[c][Nd], Cd[Md][Nd]);
            }
        }
    }

    // -- candidates: every shift within tol of the approximate maximum -----------
    float gmx[2][2];                           // per 16-register group [Nd][Md]
#pragma unroll
    for (int Md = 0; Md < 2; ++Md)
#pragma unroll
        for (int Nd = 0; Nd < 2; ++Nd) {
            float g = Cd[Md][Nd][0];
#pragma unroll
            for (int r = 1; r < 16; ++r) g = fmaxf(g, Cd[Md][Nd][r]);
            gmx[Nd][Md] = g;
        }
    const float mx = dpp_max(fmaxf(fmaxf(gmx[0][0], gmx[0][1]), fmaxf(gmx[1][0], gmx[1][1])));
    const float thr = mx - tol;
    // bit 16 (2 Nd + Md) + r: this lane's shifts in increasing row-major index a
    // N + b (shifted in from the top bit down); a group no lane has a candidate
    // in is skipped
    uint32_t mw[2] = {0u, 0u};
#pragma unroll
    for (int Nd = 0; Nd < 2; ++Nd)
#pragma unroll
        for (int Md = 1; Md >= 0; --Md) {
            uint32_t bits = 0;
            if (gmx[Nd][Md] >= thr) {
#pragma unroll
                for (int r = 15; r >= 0; --r) bits = shift_in_ge(bits, Cd[Md][Nd][r], thr);
            }
            mw[Nd] |= bits << (16 * Md);
        }
    uint64_t mask = ((uint64_t)mw[1] << 32) | mw[0];

    // -- exact integer correlation at the candidates, in increasing index order ---
    const uint32_t* rrow = reinterpret_cast<const uint32_t*>(erows + lane * kExR);
    int bestI = -1, peak = -1;
    for (;;) {
        int mine = INT_MAX;
        if (mask) {
            const int bit = __builtin_ctzll(mask);
            const int Nd = bit >> 5, Md = (bit >> 4) & 1, r = bit & 15;
            mine = (32 * Nd + m) * N + 32 * Md + 8 * (r >> 2) + 4 * h + (r & 3);
        }
        const int k = dpp_min(mine);
        if (k == INT_MAX) break;
        const int kv = k >> 6, ku = k & 63, d = ku >> 2, sh = ku & 3;
        const uint32_t* trow = reinterpret_cast<const uint32_t*>(etrows + ((lane + kv) & 63) * kExT);
        uint32_t acc = 0;
        uint32_t w0 = trow[d];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t w1 = trow[d + j + 1];
            acc = __builtin_amdgcn_udot4(rrow[j], __builtin_amdgcn_alignbyte(w1, w0, sh), acc, false);
            w0 = w1;
        }
        const int I = dpp_sum((int)acc);
        if (I > bestI) {
            bestI = I;
            peak = k;
        }
        if (mine == k) mask &= mask - 1;
    }
    if (lane == 0) {
        const double nr = (double)Srr - (double)Sr * (double)Sr / (double)NN;
        const double nt = (double)Stt - (double)St * (double)St / (double)NN;
        int iu, iv;
        wrap_peak(peak, N, iu, iv);
        p.u[poi] = (float)iu;
        p.v[poi] = (float)iv;
        p.zncc[poi] = (float)(((double)bestI - (double)Sr * (double)St / (double)NN) / sqrt(nr * nt));
        if (p.status) p.status[poi] = TWDIC_FFTCC_OK;
    }
}

// Workgroups of kDftWaves wavefronts share the DFT tables in LDS; workgroup
// b takes the contiguous POI range [g n / G, (g + 1) n / G) (G workgroups),
// its wavefronts every kDftWaves-th POI of it.  Workgroups go to the 8 XCDs
// round-robin, so g = (b mod 8) G / 8 + b / 8 when 8 divides G: each XCD
// (and its L2) then covers one contiguous eighth of the POIs, a band of rows.
__device__ __forceinline__ void dft_tables_to_lds(DftWgSmem& sm)
{
    const h8* gb = &g_dft_tab[1][0][0][0][0];
    h8* la = &sm.ta[0][0][0];
    h8* lb = &sm.tb[0][0][0][0];
    for (int i = threadIdx.x; i < 2 * 4 * 64; i += blockDim.x)        // table 0, tile 0, parts re and im
        la[i] = g_dft_tab[0][i >> 8][0][(i >> 6) & 3][i & 63];
    for (int i = threadIdx.x; i < 3 * 2 * 4 * 64; i += blockDim.x) lb[i] = gb[i];
    for (int i = threadIdx.x; i < 2 * 64; i += blockDim.x) sm.tn[i >> 6][i & 63] = g_dft_nyq[i >> 6][i & 63];
    __syncthreads();
}

__device__ __forceinline__ int dft_range_index()
{
    const int G = gridDim.x, b = blockIdx.x;
    return (G & 7) == 0 ? (b & 7) * (G >> 3) + (b >> 3) : b;
}

__global__ __launch_bounds__(64 * kDftWaves, 1) void k_fftcc_dft_mfma(FftccParams p)
{
    __shared__ DftWgSmem sm;
    dft_tables_to_lds(sm);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = dft_range_index();
    const int lo = (int)((long long)g * p.n / gridDim.x);
    const int hi = (int)((long long)(g + 1) * p.n / gridDim.x);
    for (int poi = lo + wave; poi < hi; poi += kDftWaves) fftcc_dft_mfma_poi(sm, wave, lane, p, poi);
}

// Batched form over items (pair, POI) of pairs with N = 64: the same ranges
// per workgroup, each wavefront reading its item's pair descriptor.
__global__ __launch_bounds__(64 * kDftWaves, 1) void k_batch_fftcc_dft_mfma(const PairDesc* __restrict__ pairs,
                                                                         const int2* __restrict__ items, int count)
{
    __shared__ DftWgSmem sm;
    dft_tables_to_lds(sm);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = dft_range_index();
    const int lo = (int)((long long)g * count / gridDim.x);
    const int hi = (int)((long long)(g + 1) * count / gridDim.x);
    for (int i = lo + wave; i < hi; i += kDftWaves) {
        const int2 it = items[i];
        const FftccParams p = const_pairs(pairs)[__builtin_amdgcn_readfirstlane(it.x)].fftcc;
        fftcc_dft_mfma_poi(sm, wave, lane, p, __builtin_amdgcn_readfirstlane(it.y));
    }
}

// workgroups for n POIs: one per compute unit (each holds one), fewer for
// small n
int dft_blocks(int n)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    const int cus = dev >= 0 && dev < 64 && g_dft_cus[dev] > 0 ? g_dft_cus[dev] : 256;
    return std::max(1, std::min(cus, (n + kDftWaves - 1) / kDftWaves));
}

}  // namespace

int mma_form(int N)
{
    return N == 64 ? 0 : N > 48 ? 1 : N > 32 ? 2 : 3;
}

hipError_t launch_batch_fftcc_mma(const PairDesc* pairs, const int2* items, int count, int form, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    switch (form) {
    case 0:
        hipLaunchKernelGGL(k_batch_fftcc_dft_mfma, dim3(dft_blocks(count)), dim3(64 * kDftWaves), 0, st, pairs, items,
                           count);
        break;
    case 1: hipLaunchKernelGGL((k_batch_fftcc_mma<64, false, 32>), dim3(count), dim3(64), 0, st, pairs, items); break;
    case 2: hipLaunchKernelGGL((k_batch_fftcc_mma<64, false, 16>), dim3(count), dim3(64), 0, st, pairs, items); break;
    default: hipLaunchKernelGGL((k_batch_fftcc_mma<32, false, 32>), dim3(count), dim3(64), 0, st, pairs, items); break;
    }
    return hipGetLastError();
}

// variant: occupancy target of the register engine (waves per SIMD the
// compiler budgets registers for): 2 (no spills), 3 or 4.
hipError_t launch_fftcc(const FftccParams& p, int variant, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    const int pairs = (p.n + 1) / 2;
    if (variant == kFftccDftMfma && p.N == 64) {
        hipLaunchKernelGGL(k_fftcc_dft_mfma, dim3(dft_blocks(p.n)), dim3(64 * kDftWaves), 0, st, p);
    } else if (variant == kFftccMma || variant == kFftccDftMfma) {
        if (p.N == 64)
            hipLaunchKernelGGL((k_fftcc_mma<64, true, 32>), dim3(p.n), dim3(64), 0, st, p);
        else if (p.N > 48)
            hipLaunchKernelGGL((k_fftcc_mma<64, false, 32>), dim3(p.n), dim3(64), 0, st, p);
        else if (p.N > 32)
            hipLaunchKernelGGL((k_fftcc_mma<64, false, 16>), dim3(p.n), dim3(64), 0, st, p);
        else
            hipLaunchKernelGGL((k_fftcc_mma<32, false, 32>), dim3(p.n), dim3(64), 0, st, p);
    } else if (p.N == 64) {
        if (variant == 3)
            hipLaunchKernelGGL((k_fftcc_fft<16, 3>), dim3(pairs), dim3(256), 0, st, p);
        else if (variant == 4)
            hipLaunchKernelGGL((k_fftcc_fft<16, 4>), dim3(pairs), dim3(256), 0, st, p);
        else
            hipLaunchKernelGGL((k_fftcc_fft<16, 2>), dim3(pairs), dim3(256), 0, st, p);
    } else if (p.N == 32) {
        hipLaunchKernelGGL((k_fftcc_fft<8, 4>), dim3(pairs), dim3(128), 0, st, p);
    } else
        hipLaunchKernelGGL(k_fftcc_dft, dim3(p.n), dim3(kDftThreads), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_batch_fftcc(const PairDesc* pairs, const int2* items, int total, int* counter, int N, int variant,
                              int blocks, hipStream_t st)
{
    if (total <= 0) return hipSuccess;
    if (N == 64) {
        if (variant == 2)
            hipLaunchKernelGGL((k_batch_fftcc<16, 2>), dim3(blocks), dim3(256), 0, st, pairs, items, total, counter);
        else if (variant == 4)
            hipLaunchKernelGGL((k_batch_fftcc<16, 4>), dim3(blocks), dim3(256), 0, st, pairs, items, total, counter);
        else
            hipLaunchKernelGGL((k_batch_fftcc<16, 3>), dim3(blocks), dim3(256), 0, st, pairs, items, total, counter);
    } else if (N == 32) {
        hipLaunchKernelGGL((k_batch_fftcc<8, 4>), dim3(blocks), dim3(128), 0, st, pairs, items, total, counter);
    } else {
        hipLaunchKernelGGL(k_batch_fftcc_dft, dim3(blocks), dim3(kDftThreads), 0, st, pairs, items, total, counter);
    }
    return hipGetLastError();
}

int persistent_blocks_fftcc(int device, int N, int variant)
{
    int cus = 0, per = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    hipError_t e;
    if (N == 64)
        e = variant == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_fftcc<16, 2>, 256, 0)
            : variant == 4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_fftcc<16, 4>, 256, 0)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_fftcc<16, 3>, 256, 0);
    else if (N == 32)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_fftcc<8, 4>, 128, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_fftcc_dft, kDftThreads, 0);
    if (e != hipSuccess || per < 1) per = 1;
    return (cus > 0 ? cus : 256) * per;
}

hipError_t dft_tables_init(int device)
{
    static std::mutex mu;
    static bool done[64] = {};
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (done[device]) return hipSuccess;
    static _Float16 host[2][3][2][4][64][8];
    for (int t = 0; t < 2; ++t)
        for (int tile = 0; tile < 2; ++tile)
            for (int c = 0; c < 4; ++c)
                for (int lane = 0; lane < 64; ++lane)
                    for (int q = 0; q < 8; ++q) {
                        const int m = lane & 31, h = lane >> 5;
                        const int row = t == 0 ? 16 * c + 8 * h + q : 32 * tile + m;
                        const int col = t == 0 ? 32 * tile + m : 16 * c + 8 * (q / 4) + 4 * h + q % 4;
                        const double ang = -2.0 * M_PI * (double)((row * col) & 63) / 64.0;
                        // direct double -> fp16, nearest-even; table 1 then times 2^-8 (exact:
                        // its non-zero entries stay normal)
                        const double sc = t == 0 ? 1.0 : 1.0 / 256.0;
                        host[t][0][tile][c][lane][q] = (_Float16)((double)(_Float16)std::cos(ang) * sc);
                        host[t][1][tile][c][lane][q] = (_Float16)((double)(_Float16)std::sin(ang) * sc);
                        if (t == 0 && tile == 0 && m == 0)                        // (-1)^x: Y[.][32]
                            host[t][1][tile][c][lane][q] = (_Float16)(((16 * c + 8 * h + q) & 1) ? -1.0 : 1.0);
                        host[t][2][tile][c][lane][q] = -host[t][1][tile][c][lane][q];
                    }
    static _Float16 nyq[2][64][8];
    for (int tile = 0; tile < 2; ++tile)
        for (int lane = 0; lane < 64; ++lane)
            for (int q = 0; q < 8; ++q) {
                nyq[tile][lane][q] = host[1][1][tile][0][lane][q];
                if ((lane >> 5) == 0 && q == 0)                                   // K index 0: b = 32 tile + m
                    nyq[tile][lane][q] = (_Float16)((lane & 1) ? -1.0 / 256.0 : 1.0 / 256.0);
            }
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&g_dft_cus[device], hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(g_dft_tab), host, sizeof(host));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(g_dft_nyq), nyq, sizeof(nyq));
    if (e == hipSuccess) e = hipSetDevice(cur);
    if (e == hipSuccess) done[device] = true;
    return e;
}

hipError_t launch_fftcc_generic(const FftccParams& p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fftcc_dft, dim3(p.n), dim3(kDftThreads), 0, st, p);
    return hipGetLastError();
}

}  // namespace twdic
