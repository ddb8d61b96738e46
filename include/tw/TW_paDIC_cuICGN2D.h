// TW_paDIC_cuICGN2D.h -- GPU IC-GN, drop-in for TW_Engine's cuICGN2D
// (TW/TW_Engine/TW_paDIC_cuICGN2D.h:9-57), implemented on the C ABI in
// twdic.h with the CPU path's semantics (ICGN2D_CPU::ICGN2D_Compute,
// TW_paDIC_ICGN2D_CPU.cpp:481-925): per-POI ICGN2DFlag, 1-based iteration
// counts, Cholesky solve (FP64 here).
//
// Same constructor argument order as the reference, including the
// interpolation flag.  cuInitialize / cuCompute take device pointers as in the
// reference; the images are copied into the context (the reference borrows
// the pointers and frees them in its destructor, a double free with the
// FFT-CC owner -- this class never frees what it did not allocate).
// Extensions: ICGN2D_Algorithm mirrors ICGN2D_CPU's host-array interface
// (TW_paDIC_ICGN2D_CPU.h:56-70) so a CPU IC-GN worker can switch to the GPU;
// setInterpolationSource selects the LUT gather or the fused evaluation of
// the bicubic interpolant; ICGN2DInterpolationFLag::BicubicSpline uses the
// cubic B-spline LUT of the reference's CPU path (needs a 2-pixel frame right
// of and below the ROI).
#ifndef TWAMD_TW_PADIC_CUICGN2D_H
#define TWAMD_TW_PADIC_CUICGN2D_H

#include "TW_paDIC_ICGN2D.h"

struct twdic_ctx;

namespace TW {
namespace paDIC {

class TW_LIB_DLL_EXPORTS cuICGN2D : public ICGN2D {
public:
    GPUHandle_ICGN g_cuHandleICGN;

    cuICGN2D(int_t iImgWidth, int_t iImgHeight, int_t iStartX, int_t iStartY, int_t iROIWidth, int_t iROIHeight,
             int_t iSubsetX, int_t iSubsetY, int_t iNumberX, int_t iNumberY, int_t iNumIterations, real_t fDeltaP,
             ICGN2DInterpolationFLag Iflag);
    ~cuICGN2D();

    // reference image already on the device (row stride = image width)
    void cuInitialize(uchar1* d_fRefImg);
    // target on the device; d_fU/d_fV in: integer guesses, out: sub-pixel results
    void cuCompute(uchar1* d_fTarImg, int_t* d_iPOIXY, real_t* d_fU, real_t* d_fV);
    void cuFinalize();

    void Initialize(MatView refImg);
    void ResetRefImg(const MatView& refImg) override;
    void SetTarImg(const MatView& tarImg) override;

    // host-array form of ICGN2D_CPU::ICGN2D_Algorithm: fU/fV/iNumIterations of
    // m_iPOINumber entries, iPOIpos = int[2 * m_iPOINumber] {y, x}.  Returns the
    // per-POI flags through flags when non-null.
    void ICGN2D_Algorithm(real_t* fU, real_t* fV, int* iNumIterations, int* iPOIpos, const MatView& tarImg,
                          ICGN2DFlag* flags = nullptr);
    void ICGN2D_Algorithm(real_t* fU, real_t* fV, int* iNumIterations, int* iPOIpos, const MatView& tarImg,
                          float& fPrecomputeTime, float& fICGNTime);
    // the [head, tail) form of ICGN2D_CPU (TW_paDIC_ICGN2D_CPU.h:71-85): its
    // implementation (TW_paDIC_ICGN2D_CPU.cpp:198-254) runs every POI and
    // does not read head or tail, and so does this one
    void ICGN2D_Algorithm(real_t* fU, real_t* fV, int* iNumIterations, int* iPOIpos, const MatView& tarImg,
                          int_t head, int_t tail);

    // 0 = gather from the materialised LUT (the reference's scheme),
    // 1 = evaluate the same interpolant from the pixels (default)
    void setInterpolationSource(int source);
    twdic_ctx* context() const { return m_ctx; }

private:
    void prepare();

    ICGN2DInterpolationFLag m_Iflag;
    twdic_ctx* m_ctx = nullptr;
    int m_source = 1;
    int* m_d_poi = nullptr;         // device copy of host POIs (host-array API)
    real_t* m_d_scratch = nullptr;  // device u/v for the host-array API
};

}  // namespace paDIC
}  // namespace TW

#endif
