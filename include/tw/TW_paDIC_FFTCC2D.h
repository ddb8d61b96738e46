// TW_paDIC_FFTCC2D.h -- FFT-CC configuration base class
// (replaces TW/TW_Engine/TW_paDIC_FFTCC2D.h:20-126 and TW_paDIC_FFTCC2D.cpp:6-94).
#ifndef TWAMD_TW_PADIC_FFTCC2D_H
#define TWAMD_TW_PADIC_FFTCC2D_H

#include "TW_paDIC.h"

namespace TW {
namespace paDIC {

class TW_LIB_DLL_EXPORTS Fftcc2D {
public:
    // whole-image configuration: POIs at StartX/Y + margin + subset + i * grid
    Fftcc2D(const int_t iImgWidth, const int_t iImgHeight, const int_t iStartX, const int_t iStartY,
            const int_t iROIWidth, const int_t iROIHeight, const int_t iSubsetX, const int_t iSubsetY,
            const int_t iGridSpaceX, const int_t iGridSpaceY, const int_t iMarginX, const int_t iMarginY);
    // ROI-only configuration: the images are the ROI, POIs at margin + subset + i * grid
    Fftcc2D(const int_t iROIWidth, const int_t iROIHeight, const int_t iSubsetX, const int_t iSubsetY,
            const int_t iGridSpaceX, const int_t iGridSpaceY, const int_t iMarginX, const int_t iMarginY);
    virtual ~Fftcc2D() {}

    Fftcc2D() = delete;
    Fftcc2D(const Fftcc2D&) = delete;
    Fftcc2D& operator=(const Fftcc2D&) = delete;

    virtual void ResetRefImg(const MatView& refImg) = 0;
    virtual void SetTarImg(const MatView& tarImg) = 0;

    inline int_t GetNumPOIsX() const { return m_iNumPOIX; }
    inline int_t GetNumPOIsY() const { return m_iNumPOIY; }
    inline int_t GetNumPOIs() const { return m_iNumPOIX * m_iNumPOIY; }
    inline int_t GetROISize() const { return m_iROIWidth * m_iROIHeight; }
    inline int_t GetImgSize() const
    {
        return (m_iImgHeight == -1 || m_iImgWidth == -1) ? GetROISize() : m_iImgHeight * m_iImgWidth;
    }
    virtual void setROI(const int_t& iStartX, const int_t& iStartY, const int_t& iROIWidth, const int_t& iROIHeight);

protected:
    bool recomputeNumPOI();

protected:
    bool m_isWholeImgUsed;
    int_t m_iImgWidth, m_iImgHeight;
    int_t m_iROIWidth, m_iROIHeight;
    int_t m_iStartX, m_iStartY;
    int_t m_iSubsetX, m_iSubsetY;
    int_t m_iGridSpaceX, m_iGridSpaceY;
    int_t m_iMarginX, m_iMarginY;
    int_t m_iNumPOIX = 0, m_iNumPOIY = 0;
};

}  // namespace paDIC
}  // namespace TW

#endif
