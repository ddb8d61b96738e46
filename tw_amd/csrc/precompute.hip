// precompute.hip -- POI grid, central-difference gradients, bicubic LUT and
// the per-reference IC-GN state (FP64 Hessian + Cholesky), for gfx950.
//
// All gradient and LUT values are exact in FP32 (half/quarter integers and
// sums of at most 16 integer multiples of them, below 2^14), so these kernels
// are bit-identical to the CPU path in any summation order.
#include "twdic_internal.h"

#include <algorithm>
#include <cfloat>

namespace twdic {

// ---------------------------------------------------------------------------
// POI grid: TW_utils.cpp:29-49 / TW_cuUtils.cu:116-134 (y first).
// ---------------------------------------------------------------------------
__global__ void k_poi_grid(int32_t* poi, int start_x, int start_y, int nx, int ny, int mx, int my,
                           int sx, int sy, int gx, int gy)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    int iy = i / nx, ix = i - iy * nx;
    int2 yx = make_int2(start_y + my + sy + iy * gy, start_x + mx + sx + ix * gx);
    reinterpret_cast<int2*>(poi)[i] = yx;
}

hipError_t launch_poi_grid(int32_t* poi, int start_x, int start_y, int nx, int ny, int mx, int my,
                           int sx, int sy, int gx, int gy, hipStream_t st)
{
    int n = nx * ny;
    hipLaunchKernelGGL(k_poi_grid, dim3((n + 255) / 256), dim3(256), 0, st, poi, start_x, start_y,
                       nx, ny, mx, my, sx, sy, gx, gy);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gradients: GradientXY_2Images_s (TW_utils.cpp:260-273), one image.
// Thread per pixel, x fastest; stores are fully coalesced.  g2 (optional)
// receives the doubled gradients (2 Gx, 2 Gy) as a packed int16 pair, the
// 4-byte-per-pixel form the IC-GN kernels read.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gradients(const uint8_t* __restrict__ img, int stride,
                                                   int sx0, int sy0, int rw, int rh,
                                                   float* __restrict__ gx, float* __restrict__ gy,
                                                   float* __restrict__ gxy, uint32_t* __restrict__ g2)
{
    int x = blockIdx.x * 64 + (threadIdx.x & 63);
    int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= rw || y >= rh) return;
    const uint8_t* c = img + (size_t)(y + sy0) * stride + (x + sx0);
    int l = c[-1], r = c[1], u = c[-stride], d = c[stride];
    size_t o = (size_t)y * rw + x;
    if (gx) gx[o] = 0.5f * (float)(r - l);
    if (gy) gy[o] = 0.5f * (float)(d - u);
    if (g2) g2[o] = (uint32_t)(uint16_t)(int16_t)(r - l) | ((uint32_t)(uint16_t)(int16_t)(d - u) << 16);
    if (gxy) {
        int ur = c[-stride + 1], ul = c[-stride - 1], dr = c[stride + 1], dl = c[stride - 1];
        gxy[o] = 0.25f * (float)(dr - ur - dl + ul);
    }
}

hipError_t launch_gradients(const uint8_t* img, int stride, int sx0, int sy0, int rw, int rh,
                            float* gx, float* gy, float* gxy, uint32_t* g2, hipStream_t st)
{
    dim3 grid((rw + 63) / 64, (rh + 3) / 4);
    hipLaunchKernelGGL(k_gradients, grid, dim3(256), 0, st, img, stride, sx0, sy0, rw, rh, gx, gy, gxy, g2);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cubic B-spline LUT: BicubicSplineCoefficients_s (TW_utils.cpp:521-602).
// From the 4x4 neighbourhood Om (rows y-1..y+2, columns x-1..x+2):
// beta = CP Om^T CP^T with the local prefilter CP, then B beta^T B^T with the
// B-spline basis B, rotated by 180 degrees.  FP32, every product and sum
// rounded separately in the reference's order.  Every ROI pixel gets
// coefficients (the caller guarantees a 2-pixel frame right and below).
// ---------------------------------------------------------------------------
namespace {
__device__ __forceinline__ float bsp_cp(int k, int l)
{
    constexpr float t[4][4] = {{(float)(71 / 56.0), (float)(-19 / 56.0), (float)(5 / 56.0), (float)(-1 / 56.0)},
                               {(float)(-19 / 56.0), (float)(95 / 56.0), (float)(-25 / 56.0), (float)(5 / 56.0)},
                               {(float)(5 / 56.0), (float)(-25 / 56.0), (float)(95 / 56.0), (float)(-19 / 56.0)},
                               {(float)(-1 / 56.0), (float)(5 / 56.0), (float)(-19 / 56.0), (float)(71 / 56.0)}};
    return t[k][l];
}
__device__ __forceinline__ float bsp_base(int k, int l)
{
    constexpr float t[4][4] = {{(float)(-1 / 6.0), (float)(3 / 6.0), (float)(-3 / 6.0), (float)(1 / 6.0)},
                               {(float)(3 / 6.0), (float)(-6 / 6.0), (float)(3 / 6.0), 0.f},
                               {(float)(-3 / 6.0), 0.f, (float)(3 / 6.0), 0.f},
                               {(float)(1 / 6.0), (float)(4 / 6.0), (float)(1 / 6.0), 0.f}};
    return t[k][l];
}
}  // namespace

__global__ __launch_bounds__(256) void k_bspline_lut(const uint8_t* __restrict__ img, int stride, int sx0, int sy0,
                                                     int rw, int rh, float* __restrict__ lut)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= rw || y >= rh) return;
    float om[4][4], beta[4][4], c[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int l = 0; l < 4; ++l)
            om[k][l] = (float)img[(size_t)(y + sy0 - 1 + k) * stride + (x + sx0 - 1 + l)];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            float b = 0.f;
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int n = 0; n < 4; ++n) b = b + bsp_cp(k, m) * bsp_cp(l, n) * om[n][m];
            beta[k][l] = b;
        }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            float a = 0.f;
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int n = 0; n < 4; ++n) a = a + bsp_base(k, m) * bsp_base(l, n) * beta[n][m];
            c[k][l] = a;
        }
    float4* out = reinterpret_cast<float4*>(lut + ((size_t)y * rw + x) * 16);
#pragma unroll
    for (int k = 0; k < 4; ++k)      // rotated by 180 degrees (:591-599)
        out[k] = make_float4(c[3 - k][3], c[3 - k][2], c[3 - k][1], c[3 - k][0]);
}

hipError_t launch_bspline_lut(const uint8_t* tar, int stride, int sx0, int sy0, int rw, int rh, float* lut,
                              hipStream_t st)
{
    dim3 grid((rw + 63) / 64, (rh + 3) / 4);
    hipLaunchKernelGGL(k_bspline_lut, grid, dim3(256), 0, st, tar, stride, sx0, sy0, rw, rh, lut);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Bicubic LUT: BicubicCoefficients_s (TW_utils.cpp:604-700), fused with the
// target gradients: tau (corner values, x/y/xy central differences) is built
// straight from the 4x4 pixel neighbourhood, alpha = M tau with the standard
// bicubic matrix (TW_utils.cpp:622-638) expanded by hand.  Layout per pixel:
// 16 contiguous floats, element 4k+l multiplies y^k x^l.  The last ROI row and
// column are written as zeros (the CPU buffer is calloc-zeroed there).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bicubic_lut(const uint8_t* __restrict__ tar, int stride,
                                                     int sx0, int sy0, int rw, int rh,
                                                     float* __restrict__ lut)
{
    int x = blockIdx.x * 64 + (threadIdx.x & 63);
    int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= rw || y >= rh) return;
    float4* out = reinterpret_cast<float4*>(lut + ((size_t)y * rw + x) * 16);
    if (x == rw - 1 || y == rh - 1) {
        float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        out[0] = z; out[1] = z; out[2] = z; out[3] = z;
        return;
    }
    // p[r][c] = T(y0 - 1 + r, x0 - 1 + c), r,c in 0..3
    const uint8_t* b = tar + (size_t)(y + sy0 - 1) * stride + (x + sx0 - 1);
    float p[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) p[r][c] = (float)b[r * stride + c];
    // corner values f, x/y derivatives, cross derivatives (TW_utils.cpp:265-271)
    float f00 = p[1][1], f10 = p[1][2], f01 = p[2][1], f11 = p[2][2];
    float x00 = 0.5f * (p[1][2] - p[1][0]), x10 = 0.5f * (p[1][3] - p[1][1]);
    float x01 = 0.5f * (p[2][2] - p[2][0]), x11 = 0.5f * (p[2][3] - p[2][1]);
    float y00 = 0.5f * (p[2][1] - p[0][1]), y10 = 0.5f * (p[2][2] - p[0][2]);
    float y01 = 0.5f * (p[3][1] - p[1][1]), y11 = 0.5f * (p[3][2] - p[1][2]);
    float c00 = 0.25f * (p[2][2] - p[0][2] - p[2][0] + p[0][0]);
    float c10 = 0.25f * (p[2][3] - p[0][3] - p[2][1] + p[0][1]);
    float c01 = 0.25f * (p[3][2] - p[1][2] - p[3][0] + p[1][0]);
    float c11 = 0.25f * (p[3][3] - p[1][3] - p[3][1] + p[1][1]);
    // tau = [f00 f10 f01 f11 | x00 x10 x01 x11 | y00 y10 y01 y11 | c00 c10 c01 c11]
    float a[16];
    a[0] = f00;
    a[1] = x00;
    a[2] = -3.f * f00 + 3.f * f10 - 2.f * x00 - x10;
    a[3] = 2.f * f00 - 2.f * f10 + x00 + x10;
    a[4] = y00;
    a[5] = c00;
    a[6] = -3.f * y00 + 3.f * y10 - 2.f * c00 - c10;
    a[7] = 2.f * y00 - 2.f * y10 + c00 + c10;
    a[8] = -3.f * f00 + 3.f * f01 - 2.f * y00 - y01;
    a[9] = -3.f * x00 + 3.f * x01 - 2.f * c00 - c01;
    a[10] = 9.f * f00 - 9.f * f10 - 9.f * f01 + 9.f * f11 + 6.f * x00 + 3.f * x10 - 6.f * x01 - 3.f * x11
          + 6.f * y00 - 6.f * y10 + 3.f * y01 - 3.f * y11 + 4.f * c00 + 2.f * c10 + 2.f * c01 + c11;
    a[11] = -6.f * f00 + 6.f * f10 + 6.f * f01 - 6.f * f11 - 3.f * x00 - 3.f * x10 + 3.f * x01 + 3.f * x11
          - 4.f * y00 + 4.f * y10 - 2.f * y01 + 2.f * y11 - 2.f * c00 - 2.f * c10 - c01 - c11;
    a[12] = 2.f * f00 - 2.f * f01 + y00 + y01;
    a[13] = 2.f * x00 - 2.f * x01 + c00 + c01;
    a[14] = -6.f * f00 + 6.f * f10 + 6.f * f01 - 6.f * f11 - 4.f * x00 - 2.f * x10 + 4.f * x01 + 2.f * x11
          - 3.f * y00 + 3.f * y10 - 3.f * y01 + 3.f * y11 - 2.f * c00 - c10 - 2.f * c01 - c11;
    a[15] = 4.f * f00 - 4.f * f10 - 4.f * f01 + 4.f * f11 + 2.f * x00 + 2.f * x10 - 2.f * x01 - 2.f * x11
          + 2.f * y00 - 2.f * y10 + 2.f * y01 - 2.f * y11 + c00 + c10 + c01 + c11;
    out[0] = make_float4(a[0], a[1], a[2], a[3]);
    out[1] = make_float4(a[4], a[5], a[6], a[7]);
    out[2] = make_float4(a[8], a[9], a[10], a[11]);
    out[3] = make_float4(a[12], a[13], a[14], a[15]);
}

hipError_t launch_bicubic_lut(const uint8_t* tar, int stride, int sx0, int sy0, int rw, int rh,
                              float* lut, hipStream_t st)
{
    dim3 grid((rw + 63) / 64, (rh + 3) / 4);
    hipLaunchKernelGGL(k_bicubic_lut, grid, dim3(256), 0, st, tar, stride, sx0, sy0, rw, rh, lut);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Row polynomials of the target for the fused bicubic interpolation.  The
// bicubic LUT above, built from central-difference corner derivatives, is the
// tensor-product Catmull-Rom spline; along one image row y, between pixels x
// and x + 1, its cubic through p0..p3 = T(y, x-1 .. x+2) is
//   R(t) = c0 + t/2 (c1 + t (c2 + t c3)),
//   c0 = p1, c1 = p2 - p0, c2 = 2 p0 - 5 p1 + 4 p2 - p3, c3 = 3 (p1 - p2) + p3 - p0,
// integers of magnitude at most 1530.  Stored as c0, c1/2, c2/2, c3/2 (the
// halves are multiples of 1/2 below 766 in magnitude, so exact in fp16), the
// cubic is c0 + t (c1/2 + t (c2/2 + t c3/2)).  Entry (x, j) holds the four
// coefficients of rows y = sy0 - 1 + j and y + 1 (fp16, row y, then row y + 1:
// 16 bytes) for cell column sx0 + x, x < rw, j <= rh; the last column (never
// an interior cell) is zero.  A cell's four neighbourhood rows are then two
// 16-byte entries, (x, j) and (x, j + 2).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t h2bits(float lo, float hi)
{
    const uint32_t a = __builtin_bit_cast(uint16_t, (_Float16)lo);
    const uint32_t b = __builtin_bit_cast(uint16_t, (_Float16)hi);
    return a | (b << 16);
}

// Entry (x, j) from the two target rows' dwords at columns sx0 + x - 1 ..
// sx0 + x + 2: for x < rw - 1 they lie in the image (the ROI keeps a one-pixel
// frame); the buffer's range check would read anything past its last row as
// zero.  The target resource's base is the image, records (sy0 + rh + 1) rows.
__device__ __forceinline__ void row_poly_entry(__amdgpu_buffer_rsrc_t img, int stride, int sx0, int sy0, int rw,
                                               int x, int j, uint4* __restrict__ rp)
{
    uint4 e = make_uint4(0u, 0u, 0u, 0u);
    if (x < rw - 1) {
        const int o = (sy0 - 1 + j) * stride + (sx0 - 1 + x);
        uint32_t h[4];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(img, o, r * stride, 0);
            const int p0 = (int)(w & 255u), p1 = (int)((w >> 8) & 255u), p2 = (int)((w >> 16) & 255u),
                      p3 = (int)(w >> 24);
            h[2 * r] = h2bits((float)p1, 0.5f * (float)(p2 - p0));
            h[2 * r + 1] = h2bits(0.5f * (float)(2 * p0 - 5 * p1 + 4 * p2 - p3), 0.5f * (float)(3 * (p1 - p2) + p3 - p0));
        }
        e = make_uint4(h[0], h[1], h[2], h[3]);
    }
    rp[(size_t)j * rw + x] = e;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_poly_src(const uint8_t* tar, int stride, int sy0, int rh)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)tar, (short)0, (sy0 + rh + 1) * stride, 0x00020000);
}

// One entry per thread: x along the block's 256 threads and blockIdx.x, rows
// j = blockIdx.y, + gridDim.y, ...
__global__ __launch_bounds__(256) void k_row_poly(const uint8_t* __restrict__ tar, int stride, int sx0, int sy0,
                                                  int rw, int rh, uint4* __restrict__ rp)
{
    const __amdgpu_buffer_rsrc_t img = row_poly_src(tar, stride, sy0, rh);
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= rw) return;
    for (int j = blockIdx.y; j <= rh; j += gridDim.y) row_poly_entry(img, stride, sx0, sy0, rw, x, j, rp);
}

hipError_t launch_row_poly(const uint8_t* tar, int stride, int sx0, int sy0, int rw, int rh, uint4* rp,
                           hipStream_t st)
{
    const dim3 grid((rw + 255) / 256, std::min(rh + 1, 65535));
    hipLaunchKernelGGL(k_row_poly, grid, dim3(256), 0, st, tar, stride, sx0, sy0, rw, rh, rp);
    return hipGetLastError();
}

// Batched form: blockIdx.y is the pair; geometry and buffers come from its
// descriptor (read when the kernel runs, so a captured graph follows new
// targets).  Rows j = blockIdx.x, + gridDim.x, ..., the block's threads along
// x.  Pairs without a plane (rp null) are skipped.
__global__ __launch_bounds__(256) void k_batch_row_poly(const PairDesc* __restrict__ pairs)
{
    const IcgnParams P = const_pairs(pairs)[blockIdx.y].icgn;
    if (P.rp == nullptr) return;
    const __amdgpu_buffer_rsrc_t img = row_poly_src(P.tar, P.stride, P.sy0, P.rh);
    for (int j = blockIdx.x; j <= P.rh; j += gridDim.x)
        for (int x = threadIdx.x; x < P.rw; x += 256)
            row_poly_entry(img, P.stride, P.sx0, P.sy0, P.rw, x, j, const_cast<uint4*>(P.rp));
}

hipError_t launch_batch_row_poly(const PairDesc* pairs, int npairs, int max_rh, hipStream_t st)
{
    if (npairs <= 0) return hipSuccess;
    const int blocks = std::max(1, std::min(max_rh + 1, 1024));
    hipLaunchKernelGGL(k_batch_row_poly, dim3(blocks, npairs), dim3(256), 0, st, pairs);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-reference state, one POI per group of L lanes (a wavefront, or a
// half-wavefront for large POI sets, as the iterations; built once per
// reference and POI set; the reference rebuilds it on m_isRefImgUpdated,
// TW_paDIC_cuICGN2D.cu:883-903).  Sums follow ICGN2D_Compute's precompute
// (TW_paDIC_ICGN2D_CPU.cpp:516-595): SD = [Rx, Rx dx, Rx dy, Ry, Ry dx, Ry dy],
// H = sum SD^T SD.  Every SD is a half-integer multiple and every product is
// a multiple of 1/4 below 2^24, so the FP64 sums here are exact and
// independent of order: the group width does not change a bit of it.  The
// Cholesky factorisation is FP64.
// ---------------------------------------------------------------------------
template <int O>
__device__ __forceinline__ double add_lane_xor_d(double v)
{
    // lane ^ O within 32-lane groups: ds_swizzle bit mode (and-mask 0x1f, or-mask 0, xor-mask O)
    constexpr int pattern = 0x1f | (O << 10);
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), pattern);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), pattern);
    return v + __hiloint2double(hi, lo);
}

template <int L>
__device__ __forceinline__ double group_sum_d(double v)
{
    static_assert(L == 64 || L == 32, "lane groups of 64 or 32");
    if constexpr (L == 64) v += __shfl_xor(v, 32);
    v = add_lane_xor_d<16>(v);
    v = add_lane_xor_d<8>(v);
    v = add_lane_xor_d<4>(v);
    v = add_lane_xor_d<2>(v);
    return add_lane_xor_d<1>(v);
}

template <int L>
__device__ __forceinline__ void ref_state_poi(const IcgnParams& P, int poi, int lane)
{
    const int hl = lane & (L - 1);
    const int py = P.poi[2 * poi], px = P.poi[2 * poi + 1];
    double* rs = P.refstate + (size_t)poi * kRefStateStride;
    // subsets that leave the ROI are flagged by the iteration kernel; keep the
    // state well defined and touch no memory outside the ROI
    if (px - P.Sx < P.sx0 || py - P.Sy < P.sy0 || px + P.Sx >= P.sx0 + P.rw ||
        py + P.Sy >= P.sy0 + P.rh) {
        for (int i = hl; i < kRefStateStride; i += L) rs[i] = (i == RS_FLAG) ? (double)TWDIC_ICGN_OUT_OF_ROI : 0.0;
        return;
    }
    double h[21], a[6], br[6];
#pragma unroll
    for (int i = 0; i < 21; ++i) h[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) { a[i] = 0.0; br[i] = 0.0; }
    long long sr = 0, srr = 0;
    const int W = P.Wsub, s = P.s;
    const int q64 = L / W, r64 = L - q64 * W;
    int l = hl / W, m = hl - l * W;
    for (int p = hl; p < s; p += L) {
        int gy = py - P.Sy + l, gxp = px - P.Sx + m;
        int R = P.ref[(size_t)gy * P.stride + gxp];
        const uint32_t g = P.g2[(size_t)(gy - P.sy0) * P.rw + (gxp - P.sx0)];
        double rx = 0.5 * (double)(int16_t)(g & 0xffffu), ry = 0.5 * (double)(int16_t)(g >> 16);
        double dx = (double)(m - P.Sx), dy = (double)(l - P.Sy);
        double sd[6] = {rx, rx * dx, rx * dy, ry, ry * dx, ry * dy};
        int t = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int q = 0; q <= k; ++q) h[t++] += sd[k] * sd[q];
#pragma unroll
        for (int k = 0; k < 6; ++k) { a[k] += sd[k]; br[k] += sd[k] * (double)R; }
        sr += R;
        srr += (long long)R * R;
        m += r64; l += q64;
        if (m >= W) { m -= W; l += 1; }
    }
#pragma unroll
    for (int i = 0; i < 21; ++i) h[i] = group_sum_d<L>(h[i]);
#pragma unroll
    for (int i = 0; i < 6; ++i) { a[i] = group_sum_d<L>(a[i]); br[i] = group_sum_d<L>(br[i]); }
    double dsr = group_sum_d<L>((double)sr), dsrr = group_sum_d<L>((double)srr);  // exact (< 2^53)

    double mean, spread;
    mean_spread(dsr, dsrr, (double)s, mean, spread);
    double flag = 0.0;
    // TW_paDIC_ICGN2D_CPU.cpp:580-595: zero mean or zero spread -> DarkSubset.
    // Both sums are exact integers, so the test is exact: all-zero or uniform.
    if (dsr == 0.0 || dsrr * (double)s == dsr * dsr) flag = (double)TWDIC_ICGN_DARK_SUBSET;
    double Lf[21];
#pragma unroll
    for (int i = 0; i < 21; ++i) Lf[i] = h[i];
    if (flag == 0.0) {
        // Cholesky (spotrf 'L', TW_paDIC_ICGN2D_CPU.cpp:682,688-693) in FP64
        for (int j = 0; j < 6; ++j) {
            int jj = j * (j + 1) / 2;
            double d = Lf[jj + j];
            for (int k = 0; k < j; ++k) d -= Lf[jj + k] * Lf[jj + k];
            if (!(d > 0.0)) { flag = (double)TWDIC_ICGN_SINGULAR_HESSIAN; break; }
            d = sqrt(d);
            Lf[jj + j] = d;
            for (int i = j + 1; i < 6; ++i) {
                int ii = i * (i + 1) / 2;
                double v = Lf[ii + j];
                for (int k = 0; k < j; ++k) v -= Lf[ii + k] * Lf[jj + k];
                Lf[ii + j] = v / d;
            }
        }
    }
    if (hl == 0) {
#pragma unroll
        for (int i = 0; i < 21; ++i) rs[RS_L + i] = Lf[i];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            rs[RS_A + i] = a[i];
            rs[RS_B + i] = centred(br[i], mean, a[i]);
        }
        rs[RS_NORM] = spread;
        rs[RS_FLAG] = flag;
        rs[RS_MEAN] = mean;
#pragma unroll
        for (int i = 0; i < 6; ++i) rs[RS_RDIAG + i] = flag == 0.0 ? 1.0 / Lf[i * (i + 1) / 2 + i] : 0.0;
    }
}

template <int L>
__global__ __launch_bounds__(256) void k_ref_state(IcgnParams P)
{
    const int poi = blockIdx.x * (256 / L) + (int)(threadIdx.x / L);
    if (poi >= P.n) return;
    ref_state_poi<L>(P, poi, threadIdx.x & 63);
}

// Two POIs per wavefront from 4,096 POIs on, so the 35 group reductions are
// issued once for both (their own crossover: the sums are exact, so the
// width need not follow the iterations'); smaller sets one per wavefront,
// half the pixel rounds for a GPU left partly idle.
hipError_t launch_ref_state(const IcgnParams& p, hipStream_t st)
{
    if (p.n >= 4096)
        hipLaunchKernelGGL(k_ref_state<32>, dim3((p.n + 7) / 8), dim3(256), 0, st, p);
    else
        hipLaunchKernelGGL(k_ref_state<64>, dim3((p.n + 3) / 4), dim3(256), 0, st, p);
    return hipGetLastError();
}

// Persistent form over (pair, POI) work items, one item per wavefront grab.
__global__ __launch_bounds__(256) void k_batch_ref_state(const PairDesc* __restrict__ pairs,
                                                         const int2* __restrict__ items, int total, int* counter)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        int i = 0;
        if (lane == 0) i = atomicAdd(counter, 1);
        i = __builtin_amdgcn_readfirstlane(__shfl(i, 0));
        if (i >= total) break;
        const int2 it = items[i];
        const IcgnParams P = const_pairs(pairs)[__builtin_amdgcn_readfirstlane(it.x)].icgn;
        ref_state_poi<64>(P, __builtin_amdgcn_readfirstlane(it.y), lane);
    }
}

int persistent_blocks_state(int device)
{
    int cus = 0, per = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_ref_state, 256, 0) != hipSuccess || per < 1) per = 1;
    return (cus > 0 ? cus : 256) * per;
}

hipError_t launch_batch_ref_state(const PairDesc* pairs, const int2* items, int total, int* counter, int blocks,
                                  hipStream_t st)
{
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_ref_state, dim3(blocks), dim3(256), 0, st, pairs, items, total, counter);
    return hipGetLastError();
}

// Zero-fill for work counters (a kernel, so it is safe inside captured graphs).
__global__ void k_fill_zero(int* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_fill_zero(int* p, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_zero, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace twdic
