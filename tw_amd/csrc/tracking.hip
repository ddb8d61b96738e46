// tracking.hip -- POI tracking across reference resets, the step TW_Core runs
// after IC-GN (TW/TW_Core/cuda_utils.cu:174-239, called from
// FFTCCTWorkerThread::processFrameFFTCC_ICGN, fftcctworkerthread.cpp:408-438,
// 476-481).  One thread per POI; int() truncation toward zero as the reference.
#include "twdic_internal.h"

namespace twdic {

namespace {

// poi += (int(v), int(u)) in place (updatePOIpos_Kernel, cuda_utils.cu:174-186).
__global__ __launch_bounds__(256) void k_track_update_poi(int32_t* __restrict__ poi, int n,
                                                          const float* __restrict__ u, const float* __restrict__ v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int2 p = reinterpret_cast<int2*>(poi)[i];
    p.x += (int)v[i];
    p.y += (int)u[i];
    reinterpret_cast<int2*>(poi)[i] = p;
}

// out = poi + (int(v), int(u)) (accumulatePOI_Kernel, cuda_utils.cu:188-201).
__global__ __launch_bounds__(256) void k_track_current_poi(const int32_t* __restrict__ poi, int n,
                                                           const float* __restrict__ u, const float* __restrict__ v,
                                                           int32_t* __restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int2 p = reinterpret_cast<const int2*>(poi)[i];
    p.x += (int)v[i];
    p.y += (int)u[i];
    reinterpret_cast<int2*>(out)[i] = p;
}

// out = poi + (int(v) - int(acc_u), int(u) - int(acc_v)) (accumulatePOINew_Kernel,
// cuda_utils.cu:203-223): y takes the accumulated U and x the accumulated V,
// as the reference pairs them.
__global__ __launch_bounds__(256) void k_track_current_poi_acc(const int32_t* __restrict__ poi, int n,
                                                               const float* __restrict__ u, const float* __restrict__ v,
                                                               const float* __restrict__ acc_u,
                                                               const float* __restrict__ acc_v,
                                                               int32_t* __restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int2 p = reinterpret_cast<const int2*>(poi)[i];
    p.x += (int)v[i] - (int)acc_u[i];
    p.y += (int)u[i] - (int)acc_v[i];
    reinterpret_cast<int2*>(out)[i] = p;
}

// acc += (u, v) (accumulateUV_Kernel, cuda_utils.cu:225-239).
__global__ __launch_bounds__(256) void k_track_accumulate_uv(int n, const float* __restrict__ u,
                                                             const float* __restrict__ v, float* __restrict__ acc_u,
                                                             float* __restrict__ acc_v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    acc_u[i] += u[i];
    acc_v[i] += v[i];
}

}  // namespace

hipError_t launch_track_update_poi(int32_t* poi, int n, const float* u, const float* v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_track_update_poi, dim3((n + 255) / 256), dim3(256), 0, st, poi, n, u, v);
    return hipGetLastError();
}

hipError_t launch_track_current_poi(const int32_t* poi, int n, const float* u, const float* v, int32_t* out,
                                    hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_track_current_poi, dim3((n + 255) / 256), dim3(256), 0, st, poi, n, u, v, out);
    return hipGetLastError();
}

hipError_t launch_track_current_poi_acc(const int32_t* poi, int n, const float* u, const float* v, const float* acc_u,
                                        const float* acc_v, int32_t* out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_track_current_poi_acc, dim3((n + 255) / 256), dim3(256), 0, st, poi, n, u, v, acc_u, acc_v,
                       out);
    return hipGetLastError();
}

hipError_t launch_track_accumulate_uv(int n, const float* u, const float* v, float* acc_u, float* acc_v,
                                      hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_track_accumulate_uv, dim3((n + 255) / 256), dim3(256), 0, st, n, u, v, acc_u, acc_v);
    return hipGetLastError();
}

}  // namespace twdic
