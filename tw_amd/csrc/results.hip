// results.hip -- the analysis step after the path: least-squares strain of a
// displacement field on the POI grid (TW/MATLAB/calcustrain4_x_0_L1_5_40.m).
// The per-POI CSV writer that goes with it is host code (twdic_api.cpp).
#include "twdic_internal.h"

namespace twdic {

namespace {

// Least-squares slope of f over a (2M+1) x (2M+1) window of the POI grid
// (calcustrain4_x_0_L1_5_40.m:23-33): sum over ii (outer) and jj (inner) of
// ii * f, scaled by 3 / (G M (M+1) (2M+1)^2); the y slope weights by jj.
// FP64, every product and sum rounded separately, in the script's order.
__global__ __launch_bounds__(256) void k_strain_ls(const float* __restrict__ f, int nx, int ny, int grid, int M,
                                                   double* __restrict__ dfdx, double* __restrict__ dfdy)
{
    const int ox = nx - 2 * M, oy = ny - 2 * M;
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= ox * oy) return;
    const int i = o % ox + M, j = o / ox + M;
    double sx = 0.0, sy = 0.0;
    for (int ii = -M; ii <= M; ++ii)
        for (int jj = -M; jj <= M; ++jj) {
            const double v = (double)f[(size_t)(j + jj) * nx + (i + ii)];
            sx = __dadd_rn(sx, __dmul_rn((double)ii, v));
            sy = __dadd_rn(sy, __dmul_rn((double)jj, v));
        }
    const double w = (double)(2 * M + 1);
    const double k = __ddiv_rn(3.0, __dmul_rn(__dmul_rn(__dmul_rn((double)grid, (double)M), (double)(M + 1)), __dmul_rn(w, w)));
    if (dfdx) dfdx[o] = __dmul_rn(k, sx);
    if (dfdy) dfdy[o] = __dmul_rn(k, sy);
}

}  // namespace

hipError_t launch_strain_ls(const float* f, int nx, int ny, int grid, int M, double* dfdx, double* dfdy,
                            hipStream_t st)
{
    const int n = (nx - 2 * M) * (ny - 2 * M);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_strain_ls, dim3((n + 255) / 256), dim3(256), 0, st, f, nx, ny, grid, M, dfdx, dfdy);
    return hipGetLastError();
}

}  // namespace twdic
