// fftcc.hip -- FFT-CC integer-pixel search for gfx950, one workgroup per POI.
//
// Semantics follow Fftcc2D_CPU::ComputeFFTCC (TW_paDIC_FFTCC2D_CPU.cpp:
// 222-326) with the window edge N as a parameter (N = 2S is the reference):
//   window rows POIy - N/2 .. POIy + N/2 - 1, same for columns;
//   C[k] = sum_x R~(x) T~(x+k) (circular), peak = argmax C, lowest row-major
//   index on ties; u = k % N, v = k / N, wrapped by -N when > N/2;
//   ZNCC = C[peak] / (N^2 sqrt(sum R~^2 sum T~^2)).
//
// MI355X design:
//   * Both windows are packed into one complex array Z = R~ + i T~ held in
//     LDS, so one forward 2-D transform yields both spectra
//     (F_R[k] = (Z[k] + conj Z[-k]) / 2, F_T[k] = (Z[k] - conj Z[-k]) / 2i).
//   * The cross-power spectrum conj(F_R) F_T = Im(Z[k] Z[-k]) / 2
//     - i (|Z[k]|^2 - |Z[-k]|^2) / 4 is formed in place, then inverse
//     transformed; nothing round-trips HBM.
//   * Exact peak: the zero-mean correlation differs from the integer
//     correlation I[k] = sum_x R(x) T(x+k) by a constant, so the true peak is
//     an integer argmax.  The FP32 surface gives the maximum; every shift
//     within `tol` (ZNCC units) of it is re-evaluated exactly in integers
//     from the uint8 windows kept in LDS.  With a single candidate (the
//     usual case) the FP32 argmax is provably exact and no integer work runs.
//
// Two transform engines share the rest of the kernel:
//   * radix-4/2 register FFT for N = 32 and N = 64 (fast path): each of
//     256 threads owns N*N/256 points, 64-point transforms run as 16 x 4
//     (16-point DFT in registers, 4-point DFT across a lane quad via DPP
//     shuffles); transposes go through LDS;
//   * direct DFT by rows/columns for any other even N <= 64 (reference
//     parity mode N = 2S = 20, 30, 40, 60).
#include "twdic_internal.h"

#include <cfloat>

namespace twdic {

namespace {

constexpr int kThreads = 256;
constexpr int kPitch = kMaxFftN + 1;   // LDS row pitch in float2, avoids bank conflicts

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

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Shared-memory layout of one workgroup.
struct FftccSmem {
    float2 z[kMaxFftN * kPitch];         // 33,280 B
    uint8_t wr[kMaxFftN * kMaxFftN];     // reference window (exact refinement)
    uint8_t wt[kMaxFftN * kMaxFftN];     // target window
    float2 tw[kMaxFftN];
    long long red[4][4];
    Top2 top[4];
    int imin[4];
};

// ---------------------------------------------------------------------------
// Direct DFT along rows (dir = 0) or columns (dir = 1), in place, with
// twiddle sign s (conjugated twiddles for the inverse).  Each wavefront owns
// whole lines; lane k computes output k of its line.  All LDS reads of a
// line complete before its writes within the same wavefront.
// ---------------------------------------------------------------------------
__device__ void dft_lines(FftccSmem& sm, int N, int dir, bool inverse)
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
                if (e >= N) e -= N;   // lane < N, so one subtraction suffices
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < N) {
            if (dir == 0) sm.z[line * kPitch + lane] = acc;
            else sm.z[lane * kPitch + line] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// 64-point and 32-point transforms spread over a lane quad (4 lanes x 16 or
// 8 points).  P = points per lane (16 for N = 64, 8 for N = 32).
//
//   dif4(x): input lane q holds x[j + P*brev2(q)], j < P (contiguous block);
//            output lane q holds X[4k + q], k < P.
//   dit4(x): input lane q holds x[4j + q];
//            output lane q holds X[k + P*brev2(q)].
// Both compute X[k] = sum_n x[n] W^{nk}, W = exp(-2 pi i sgn / (4P)).
// ---------------------------------------------------------------------------
// P-point DFT in registers: iterative radix-2 DIT over a bit-reversed copy.
// Twiddles exp(i sgn 2 pi j / m) have compile-time angles after unrolling.
template <int P>
__device__ __forceinline__ void dft_small(float2 (&x)[P], float sgn)
{
    constexpr int LOG = (P == 16) ? 4 : (P == 8) ? 3 : (P == 4) ? 2 : 1;
    float2 y[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        int r = 0;
#pragma unroll
        for (int b = 0; b < LOG; ++b) r |= ((i >> b) & 1) << (LOG - 1 - b);
        y[r] = x[i];
    }
#pragma unroll
    for (int s = 1; s <= LOG; ++s) {
        const int m = 1 << s, h = m >> 1;
#pragma unroll
        for (int k = 0; k < P; k += m) {
#pragma unroll
            for (int j = 0; j < h; ++j) {
                const double th = 2.0 * 3.14159265358979323846 * (double)j / (double)m;
                const float c = (float)__builtin_cos(th);
                const float sn = sgn * (float)__builtin_sin(th);
                float2 t = y[k + j + h];
                float2 tw = make_float2(t.x * c - t.y * sn, t.x * sn + t.y * c);
                float2 a = y[k + j];
                y[k + j] = make_float2(a.x + tw.x, a.y + tw.y);
                y[k + j + h] = make_float2(a.x - tw.x, a.y - tw.y);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < P; ++i) x[i] = y[i];
}

__device__ __forceinline__ float2 shfl_xor2(float2 v, int m)
{
    return make_float2(__shfl_xor(v.x, m), __shfl_xor(v.y, m));
}

// 4-point DFT across a lane quad.  natural_in: input lane q holds index q,
// output lane q holds index brev2(q).  Otherwise input is bit-reversed and
// output natural.  sgn: -1 forward, +1 inverse.
__device__ __forceinline__ float2 quad_dft_dit(float2 v, int q, float sgn)
{
    // stage 1 (xor 2): lanes 0,1 get v + v', lanes 2,3 get v' - v
    float2 o = shfl_xor2(v, 2);
    float2 a = (q < 2) ? make_float2(v.x + o.x, v.y + o.y) : make_float2(o.x - v.x, o.y - v.y);
    // lane 3 twiddle by W4^1 = exp(-i pi/2 sgn'): multiply by (0, sgn)
    if (q == 3) a = make_float2(-sgn * a.y, sgn * a.x);
    float2 b = shfl_xor2(a, 1);
    return (q & 1) ? make_float2(b.x - a.x, b.y - a.y) : make_float2(a.x + b.x, a.y + b.y);
}

__device__ __forceinline__ float2 quad_dft_dif(float2 v, int q, float sgn)
{
    // input lane q holds block brev2(q) = {0,2,1,3}[q]
    float2 o = shfl_xor2(v, 1);
    float2 a = (q & 1) ? make_float2(o.x - v.x, o.y - v.y) : make_float2(v.x + o.x, v.y + o.y);
    if (q == 3) a = make_float2(-sgn * a.y, sgn * a.x);
    float2 b = shfl_xor2(a, 2);
    return (q < 2) ? make_float2(a.x + b.x, a.y + b.y) : make_float2(b.x - a.x, b.y - a.y);
}

// tw: LDS table of exp(-2 pi i e / N); inverse transforms use conj(tw).
template <int P>
__device__ __forceinline__ float2 twiddle(const float2* tw, int e, bool inverse)
{
    float2 w = tw[e & (4 * P - 1)];
    if (inverse) w.y = -w.y;
    return w;
}

template <int P>
__device__ __forceinline__ void fft_dif4(float2 (&x)[P], int q, bool inverse, const float2* tw)
{
    const float sgn = inverse ? 1.f : -1.f;
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = quad_dft_dif(x[j], q, sgn);
#pragma unroll
    for (int j = 1; j < P; ++j) x[j] = cmul(x[j], twiddle<P>(tw, q * j, inverse));
    dft_small<P>(x, sgn);
}

template <int P>
__device__ __forceinline__ void fft_dit4(float2 (&x)[P], int q, bool inverse, const float2* tw)
{
    const float sgn = inverse ? 1.f : -1.f;
    dft_small<P>(x, sgn);
#pragma unroll
    for (int k = 1; k < P; ++k) x[k] = cmul(x[k], twiddle<P>(tw, q * k, inverse));
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = quad_dft_dit(x[k], q, sgn);
}

}  // namespace

// ---------------------------------------------------------------------------
// The FFT-CC kernel.  FAST selects the register radix-4 engine (N = 4*P with
// P = 16 or 8); otherwise the direct DFT handles any even N <= 64.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(kThreads) void k_fftcc(FftccParams p)
{
    __shared__ FftccSmem sm;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int poi = blockIdx.x;
    if (poi >= p.n) return;
    const int N = (P > 0) ? 4 * P : p.N;
    const int NN = N * N, half = N >> 1;
    const int py = p.poi[2 * poi], px = p.poi[2 * poi + 1];
    const int y0 = py - half, x0 = px - half;
    if (y0 < 0 || x0 < 0 || y0 + N > p.img_h || x0 + N > p.img_w) {
        if (tid == 0 && p.status) p.status[poi] = TWDIC_FFTCC_OUT;
        return;
    }

    // -- 1. gather both windows into LDS, exact integer sums ---------------
    long long sr = 0, st = 0, srr = 0, stt = 0;
    for (int i = tid; i < NN; i += kThreads) {
        int y = i / N, x = i - y * N;
        size_t o = (size_t)(y0 + y) * p.stride + (x0 + x);
        int a = p.ref[o], b = p.tar[o];
        sm.wr[i] = (uint8_t)a;
        sm.wt[i] = (uint8_t)b;
        sr += a; st += b; srr += a * a; stt += b * b;
    }
    if (tid < N) sm.tw[tid] = p.tw[tid];
    sr = wave_sum(sr); st = wave_sum(st); srr = wave_sum(srr); stt = wave_sum(stt);
    if (lane == 0) { sm.red[wave][0] = sr; sm.red[wave][1] = st; sm.red[wave][2] = srr; sm.red[wave][3] = stt; }
    __syncthreads();
    sr = sm.red[0][0] + sm.red[1][0] + sm.red[2][0] + sm.red[3][0];
    st = sm.red[0][1] + sm.red[1][1] + sm.red[2][1] + sm.red[3][1];
    srr = sm.red[0][2] + sm.red[1][2] + sm.red[2][2] + sm.red[3][2];
    stt = sm.red[0][3] + sm.red[1][3] + sm.red[2][3] + sm.red[3][3];
    // flat window (TW_paDIC_FFTCC2D_CPU.cpp:271-275): exact integer test
    if (srr * NN == sr * sr || stt * NN == st * st) {
        if (tid == 0 && p.status) p.status[poi] = TWDIC_FFTCC_FLAT;
        return;
    }
    // means as the CPU path forms them: exact sum / N^2 in FP32 (:253-254)
    const float mr = (float)sr / (float)NN, mt = (float)st / (float)NN;
    // exact norms sum (x - mean)^2 = sum x^2 - (sum x)^2 / N^2
    const double nr = (double)srr - (double)sr * (double)sr / (double)NN;
    const double nt = (double)stt - (double)st * (double)st / (double)NN;
    const float scale = (float)((double)NN * sqrt(nr * nt));

    // -- 2. forward 2-D transform of Z = R~ + i T~ ---------------------------
    if constexpr (P > 0) {
        const int q = tid & 3, line = tid >> 2;         // 64 lines x 4 lanes
        const int bq = ((q & 1) << 1) | (q >> 1);       // brev2(q)
        float2 x[P];
        // rows: DIF input = contiguous block bq of row `line`
        if (line < N) {
#pragma unroll
            for (int j = 0; j < P; ++j) {
                int c = bq * P + j;
                x[j] = make_float2((float)sm.wr[line * N + c] - mr, (float)sm.wt[line * N + c] - mt);
            }
            fft_dif4<P>(x, q, false, sm.tw);
            // lane holds row spectrum at kx = 4j + q
#pragma unroll
            for (int j = 0; j < P; ++j) sm.z[line * kPitch + 4 * j + q] = x[j];
        }
        __syncthreads();
        // columns: DIF input = contiguous block bq of column `line`
        if (line < N) {
#pragma unroll
            for (int j = 0; j < P; ++j) x[j] = sm.z[(bq * P + j) * kPitch + line];
            fft_dif4<P>(x, q, false, sm.tw);
        }
        __syncthreads();
        if (line < N) {
#pragma unroll
            for (int j = 0; j < P; ++j) sm.z[(4 * j + q) * kPitch + line] = x[j];
        }
        __syncthreads();
        // -- 3. cross-power spectrum, kept in registers ----------------------
        if (line < N) {
#pragma unroll
            for (int j = 0; j < P; ++j) {
                int ky = 4 * j + q, kx = line;
                int my = (N - ky) & (N - 1), mxx = (N - kx) & (N - 1);
                float2 a = x[j], b = sm.z[my * kPitch + mxx];
                float im_ab = a.x * b.y + a.y * b.x;
                float d = (a.x * a.x + a.y * a.y) - (b.x * b.x + b.y * b.y);
                x[j] = make_float2(0.5f * im_ab, -0.25f * d);
            }
            // -- 4. inverse columns: DIT input lane q holds ky = 4j + q
            fft_dit4<P>(x, q, true, sm.tw);
        }
        __syncthreads();
        if (line < N) {
            // lane holds y = j + P*bq of column `line`
#pragma unroll
            for (int j = 0; j < P; ++j) sm.z[(j + P * bq) * kPitch + line] = x[j];
        }
        __syncthreads();
        // inverse rows: DIT input lane q holds x = 4j + q of row `line`
        if (line < N) {
#pragma unroll
            for (int j = 0; j < P; ++j) x[j] = sm.z[line * kPitch + 4 * j + q];
            fft_dit4<P>(x, q, true, sm.tw);
        }
        __syncthreads();
        if (line < N) {
#pragma unroll
            for (int j = 0; j < P; ++j) sm.z[line * kPitch + j + P * bq] = x[j];
        }
        __syncthreads();
    } else {
        for (int i = tid; i < NN; i += kThreads) {
            int y = i / N, x = i - y * N;
            sm.z[y * kPitch + x] = make_float2((float)sm.wr[i] - mr, (float)sm.wt[i] - mt);
        }
        __syncthreads();
        dft_lines(sm, N, 0, false);
        __syncthreads();
        dft_lines(sm, N, 1, false);
        __syncthreads();
        for (int i = tid; i < NN; i += kThreads) {
            int ky = i / N, kx = i - ky * N;
            int my = ky ? N - ky : 0, mxx = kx ? N - kx : 0;
            int j = my * N + mxx;
            if (i <= j) {
                float2 a = sm.z[ky * kPitch + kx], b = sm.z[my * kPitch + mxx];
                float im_ab = a.x * b.y + a.y * b.x;
                float d = (a.x * a.x + a.y * a.y) - (b.x * b.x + b.y * b.y);
                float2 xk = make_float2(0.5f * im_ab, -0.25f * d);
                sm.z[ky * kPitch + kx] = xk;
                if (j != i) sm.z[my * kPitch + mxx] = make_float2(xk.x, -xk.y);
            }
        }
        __syncthreads();
        dft_lines(sm, N, 1, true);
        __syncthreads();
        dft_lines(sm, N, 0, true);
        __syncthreads();
    }

    // -- 5. peak: top-2 with lowest-index tie-break (:300-311) -----------------
    Top2 t{-FLT_MAX, 0x7fffffff, -FLT_MAX};
    for (int i = tid; i < NN; i += kThreads) {
        int y = i / N, x = i - y * N;
        float c = sm.z[y * kPitch + x].x;
        Top2 u{c, i, -FLT_MAX};
        t = top2_merge(t, u);
    }
    t = wave_top2(t);
    if (lane == 0) sm.top[wave] = t;
    __syncthreads();
    t = top2_merge(top2_merge(sm.top[0], sm.top[1]), top2_merge(sm.top[2], sm.top[3]));
    int peak = t.idx;
    float zn = t.best / scale;
    const float thr = t.best - p.tol * scale;
    if (t.second >= thr) {
        // -- 6. exact refinement over all near-maximal shifts -----------------
        long long bestI = -1;
        int bestk = -1, last = -1;
        for (;;) {
            int cand = 0x7fffffff;
            for (int i = tid; i < NN; i += kThreads) {
                int y = i / N, x = i - y * N;
                if (i > last && sm.z[y * kPitch + x].x >= thr) { cand = i; break; }
            }
            cand = wave_min(cand);
            __syncthreads();
            if (lane == 0) sm.imin[wave] = cand;
            __syncthreads();
            cand = min(min(sm.imin[0], sm.imin[1]), min(sm.imin[2], sm.imin[3]));
            if (cand == 0x7fffffff) break;
            const int kv = cand / N, ku = cand - kv * N;
            long long acc = 0;
            for (int i = tid; i < NN; i += kThreads) {
                int y = i / N, x = i - y * N;
                int yy = y + kv; if (yy >= N) yy -= N;
                int xx = x + ku; if (xx >= N) xx -= N;
                acc += (int)sm.wr[i] * (int)sm.wt[yy * N + xx];
            }
            acc = wave_sum(acc);
            __syncthreads();
            if (lane == 0) sm.red[wave][0] = acc;
            __syncthreads();
            acc = sm.red[0][0] + sm.red[1][0] + sm.red[2][0] + sm.red[3][0];
            if (acc > bestI) { bestI = acc; bestk = cand; }
            last = cand;
        }
        peak = bestk;
        zn = (float)(((double)bestI - (double)sr * (double)st / (double)NN) / sqrt(nr * nt));
    }
    if (tid == 0) {
        int iu = peak % N, iv = peak / N;                   // :316-322
        if (iu > half) iu -= N;
        if (iv > half) iv -= N;
        p.u[poi] = (float)iu;
        p.v[poi] = (float)iv;
        p.zncc[poi] = zn;
        if (p.status) p.status[poi] = TWDIC_FFTCC_OK;
    }
}

hipError_t launch_fftcc(const FftccParams& p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    if (p.N == 64)
        hipLaunchKernelGGL(k_fftcc<16>, dim3(p.n), dim3(kThreads), 0, st, p);
    else if (p.N == 32)
        hipLaunchKernelGGL(k_fftcc<8>, dim3(p.n), dim3(kThreads), 0, st, p);
    else
        hipLaunchKernelGGL(k_fftcc<0>, dim3(p.n), dim3(kThreads), 0, st, p);
    return hipGetLastError();
}

// Generic-engine launch for any even N (used to cross-check the fast path).
hipError_t launch_fftcc_generic(const FftccParams& p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fftcc<0>, dim3(p.n), dim3(kThreads), 0, st, p);
    return hipGetLastError();
}

}  // namespace twdic
