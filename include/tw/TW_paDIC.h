// TW_paDIC.h -- handles and enums of the paDIC path
// (replaces TW/TW_Engine/TW_paDIC.h:13-93).
//
// The GPU handles keep the reference's public member names because TW_Core
// reads them directly (fftcctworkerthread.cpp:90,414,458-459).  Members of
// scratch that this implementation never materialises (the FFT subset and
// frequency-domain buffers, per-POI subsets, steepest-descent images, the
// explicit inverse Hessian) stay nullptr: the FFTs run in LDS and the IC-GN
// keeps per-POI sums and an FP64 Cholesky factor instead.
#ifndef TWAMD_TW_PADIC_H
#define TWAMD_TW_PADIC_H

#include "TW.h"

namespace TW {
namespace paDIC {

using cufftHandle = int;   // plan handles of the reference; unused here

struct GPUHandle_FFTCC {
    uchar1* m_d_fRefImg = nullptr;   // reference image (device, row stride = image width)
    uchar1* m_d_fTarImg = nullptr;   // target image
    int_t* m_d_iPOIXY = nullptr;     // POI positions, int32[n][2] = {y, x}
    real_t* m_d_fU = nullptr;        // displacement x (host-output API only)
    real_t* m_d_fV = nullptr;
    real_t* m_d_fZNCC = nullptr;
    real_t* m_d_fSubset1 = nullptr;  // not materialised (LDS-resident FFTs)
    real_t* m_d_fSubset2 = nullptr;
    real_t* m_d_fSubsetC = nullptr;
    real_t* m_d_fMod1 = nullptr;
    real_t* m_d_fMod2 = nullptr;
    cufftHandle m_forwardPlanXY = 0;
    cufftHandle m_reversePlanXY = 0;
    float2* m_dev_FreqDom1 = nullptr;
    float2* m_dev_FreqDom2 = nullptr;
    float2* m_dev_FreqDomfg = nullptr;
};

struct GPUHandle_ICGN {
    uchar1* m_d_fRefImg = nullptr;
    uchar1* m_d_fTarImg = nullptr;
    int_t* m_d_iPOIXY = nullptr;
    real_t* m_d_fU = nullptr;
    real_t* m_d_fV = nullptr;
    real_t* m_d_fRx = nullptr;                 // ROI-relative fp32 gradients of the reference
    real_t* m_d_fRy = nullptr;
    real_t* m_d_fTx = nullptr;                 // target gradients: written only by twdic_precompute_target on the
                                               // context (inspection); IC-GN never reads them
    real_t* m_d_fTy = nullptr;
    real_t* m_d_fTxy = nullptr;
    real_t4* m_d_f4InterpolationLUT = nullptr; // [roi_h][roi_w][16] fp32, element 4k+l ~ y^k x^l; written per
                                               // frame in the LUT and B-spline sources, unused by the fused one
    int* m_d_iIterationNums = nullptr;         // 1-based iteration count per POI
    real_t* m_d_fSubsetR = nullptr;            // not materialised
    real_t* m_d_fSubsetT = nullptr;
    real_t* m_d_fSubsetAveR = nullptr;
    real_t* m_d_fSubsetAveT = nullptr;
    real_t* m_d_invHessian = nullptr;
    real_t2* m_d_RDescent = nullptr;
    real_t* m_d_dP = nullptr;                  // final p = (u, ux, uy, v, vx, vy) per POI
    uint8_t* m_d_iFlags = nullptr;             // ICGN2DFlag per POI (extension)
};

enum class ICGN2DFlag { Success, DarkSubset, SingularHessian, SingularWarp, OutofROI };
enum class paDICThreadFlag { Single, Multicore };
enum class ICGN2DInterpolationFLag { Bicubic, BicubicSpline };

}  // namespace paDIC
}  // namespace TW

#endif
