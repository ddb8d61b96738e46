// TW_paDIC_cuFFTCC2D.h -- GPU FFT-CC, drop-in for TW_Engine's cuFFTCC2D
// (TW/TW_Engine/TW_paDIC_cuFFTCC2D.h:18-170), implemented on the C ABI in
// twdic.h.  Same class name, constructor argument orders (note the whole-image
// constructor takes ROI width/height BEFORE the ROI start, unlike Fftcc2D) and
// method semantics: the cu* methods leave results on the device, the plain
// methods return hcreateptr host arrays [NumPOIY][NumPOIX].
//
// Differences from the reference, all deliberate:
//   * integer peaks are the exact argmax of the circular correlation (the
//     reference's cuFFT path can pick a different shift on near ties);
//   * errors throw std::runtime_error instead of exiting the process;
//   * the FFT-CC window edge defaults to 2 * iSubsetX as in the reference and
//     can be changed with setFFTWindow (any even N up to 64; every size runs
//     on the int8 matrix cores).
#ifndef TWAMD_TW_PADIC_CUFFTCC2D_H
#define TWAMD_TW_PADIC_CUFFTCC2D_H

#include "TW_paDIC_FFTCC2D.h"

struct twdic_ctx;

namespace TW {
namespace paDIC {

class TW_LIB_DLL_EXPORTS cuFFTCC2D : public Fftcc2D {
public:
    GPUHandle_FFTCC g_cuHandle;

    cuFFTCC2D(const int_t iROIWidth, const int_t iROIHeight, const int_t iSubsetX, const int_t iSubsetY,
              const int_t iGridSpaceX, const int_t iGridSpaceY, const int_t iMarginX, const int_t iMarginY);
    cuFFTCC2D(const int_t iImgWidth, const int_t iImgHeight, const int_t iROIWidth, const int_t iROIHeight,
              const int_t iStartX, const int_t iStartY, const int_t iSubsetX, const int_t iSubsetY,
              const int_t iGridSpaceX, const int_t iGridSpaceY, const int_t iMarginX, const int_t iMarginY);
    ~cuFFTCC2D();

    void setROI(const int_t& iStartX, const int_t& iStartY, const int_t& iROIWidth, const int_t& iROIHeight) override;
    // FFT-CC window edge N (even, <= 64); the reference fixes N = 2 * iSubsetX.
    void setFFTWindow(int_t N);

    // host-output API (TW_paDIC_cuFFTCC2D.cu:237-461)
    virtual void InitializeFFTCC(real_t**& fU, real_t**& fV, real_t**& fZNCC, const MatView& refImg);
    virtual void ComputeFFTCC(real_t**& fU, real_t**& fV, real_t**& fZNCC, const MatView& tarImg);
    virtual void DestroyFFTCC(real_t**& fU, real_t**& fV, real_t**& fZNCC);

    void ResetRefImg(const MatView& refImg) override;
    void SetTarImg(const MatView& tarImg) override;

    // device-output API (TW_paDIC_cuFFTCC2D.cu:464-677): the callee allocates
    // the device outputs in cuInitializeFFTCC and frees them in cuDestroyFFTCC
    virtual void cuInitializeFFTCC(real_t*& f_d_U, real_t*& f_d_V, real_t*& f_d_ZNCC, const MatView& refImg);
    virtual void cuComputeFFTCC(real_t*& f_d_U, real_t*& f_d_V, real_t*& f_d_ZNCC, const MatView& tarImg);
    virtual void cuDestroyFFTCC(real_t*& f_d_U, real_t*& f_d_V, real_t*& f_d_ZNCC);

    // the underlying context (stream, device buffers)
    twdic_ctx* context() const { return m_ctx; }

private:
    void ensureContext();
    void release();

    twdic_ctx* m_ctx = nullptr;
    int_t m_iFFTWindow;
    bool isLowLevelApiCalled = false;
    bool isDestroyed = false;
};

}  // namespace paDIC
}  // namespace TW

#endif
