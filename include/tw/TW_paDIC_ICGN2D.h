// TW_paDIC_ICGN2D.h -- IC-GN configuration base class
// (replaces TW/TW_Engine/TW_paDIC_ICGN2D.h:18-86 and TW_paDIC_ICGN2D.cpp:6-70).
#ifndef TWAMD_TW_PADIC_ICGN2D_H
#define TWAMD_TW_PADIC_ICGN2D_H

#include "TW_paDIC.h"

namespace TW {
namespace paDIC {

class TW_LIB_DLL_EXPORTS ICGN2D {
public:
    ICGN2D(int_t iImgWidth, int_t iImgHeight, int_t iStartX, int_t iStartY, int_t iROIWidth, int_t iROIHeight,
           int_t iSubsetX, int_t iSubsetY, int_t iNumberX, int_t iNumberY, int_t iNumIterations, real_t fDeltaP);
    ICGN2D(const MatView& refImg, int_t iImgWidth, int_t iImgHeight, int_t iStartX, int_t iStartY,
           int_t iROIWidth, int_t iROIHeight, int_t iSubsetX, int_t iSubsetY, int_t iNumberX, int_t iNumberY,
           int_t iNumIterations, real_t fDeltaP);
    virtual ~ICGN2D() = 0;

    virtual void setROI(const int_t& iStartX, const int_t& iStartY, const int_t& iROIWidth, const int_t& iROIHeight);
    virtual void ResetRefImg(const MatView& refImg) = 0;
    virtual void SetTarImg(const MatView& tarImg) = 0;

protected:
    MatView m_refImg;
    MatView m_tarImg;
    bool m_isRefImgUpdated;
    int_t m_iImgWidth, m_iImgHeight;
    int_t m_iStartX, m_iStartY;
    int_t m_iROIWidth, m_iROIHeight;
    int_t m_iSubsetX, m_iSubsetY;
    int_t m_iSubsetH, m_iSubsetW, m_iSubsetSize;
    int_t m_iNumberX, m_iNumberY, m_iPOINumber;
    int_t m_iNumIterations;
    real_t m_fDeltaP;
};

}  // namespace paDIC
}  // namespace TW

#endif
