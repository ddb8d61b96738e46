// tw_classes.cpp -- the TW_Engine-compatible C++ classes (include/tw/) on top
// of the C ABI (include/twdic.h).  Errors from the ABI become exceptions:
// std::logic_error for invalid geometry (the reference's choice,
// TW_paDIC_cuFFTCC2D.cu:189-190), std::runtime_error otherwise.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "tw/TW_paDIC_cuFFTCC2D.h"
#include "tw/TW_paDIC_cuICGN2D.h"
#include "twdic.h"

namespace TW {
namespace paDIC {

namespace {

void check(twdic_status st, twdic_ctx* ctx, const char* what)
{
    if (st == TWDIC_OK) return;
    std::string msg = std::string(what) + ": status " + std::to_string((int)st);
    if (ctx) msg += std::string(": ") + twdic_last_error(ctx);
    if (st == TWDIC_EINVAL) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

template <typename T>
T* dev_alloc(twdic_ctx* ctx, size_t n)
{
    void* p = nullptr;
    check(twdic_malloc(ctx, &p, n * sizeof(T)), ctx, "twdic_malloc");
    check(twdic_memset(ctx, p, 0, n * sizeof(T)), ctx, "twdic_memset");
    return static_cast<T*>(p);
}

template <typename T>
void dev_free(twdic_ctx* ctx, T*& p)
{
    if (ctx && p) twdic_free(ctx, p);
    p = nullptr;
}

}  // namespace

// ------------------------------------------------------------------ Fftcc2D
Fftcc2D::Fftcc2D(const int_t iImgWidth, const int_t iImgHeight, const int_t iStartX, const int_t iStartY,
                 const int_t iROIWidth, const int_t iROIHeight, const int_t iSubsetX, const int_t iSubsetY,
                 const int_t iGridSpaceX, const int_t iGridSpaceY, const int_t iMarginX, const int_t iMarginY)
    : m_isWholeImgUsed(true), m_iImgWidth(iImgWidth), m_iImgHeight(iImgHeight), m_iROIWidth(iROIWidth),
      m_iROIHeight(iROIHeight), m_iStartX(iStartX), m_iStartY(iStartY), m_iSubsetX(iSubsetX),
      m_iSubsetY(iSubsetY), m_iGridSpaceX(iGridSpaceX), m_iGridSpaceY(iGridSpaceY), m_iMarginX(iMarginX),
      m_iMarginY(iMarginY)
{
}

Fftcc2D::Fftcc2D(const int_t iROIWidth, const int_t iROIHeight, const int_t iSubsetX, const int_t iSubsetY,
                 const int_t iGridSpaceX, const int_t iGridSpaceY, const int_t iMarginX, const int_t iMarginY)
    : m_isWholeImgUsed(false), m_iImgWidth(-1), m_iImgHeight(-1), m_iROIWidth(iROIWidth),
      m_iROIHeight(iROIHeight), m_iStartX(-1), m_iStartY(-1), m_iSubsetX(iSubsetX), m_iSubsetY(iSubsetY),
      m_iGridSpaceX(iGridSpaceX), m_iGridSpaceY(iGridSpaceY), m_iMarginX(iMarginX), m_iMarginY(iMarginY)
{
}

void Fftcc2D::setROI(const int_t& iStartX, const int_t& iStartY, const int_t& iROIWidth, const int_t& iROIHeight)
{
    m_iStartX = iStartX;
    m_iStartY = iStartY;
    m_iROIWidth = iROIWidth;
    m_iROIHeight = iROIHeight;
}

// TW_paDIC_FFTCC2D.cpp:88-94
bool Fftcc2D::recomputeNumPOI()
{
    if (m_iGridSpaceX < 1 || m_iGridSpaceY < 1) {   // would divide by zero: no grid
        m_iNumPOIX = m_iNumPOIY = 0;
        return false;
    }
    m_iNumPOIX = int_t(std::floor((m_iROIWidth - m_iSubsetX * 2 - m_iMarginX * 2) / real_t(m_iGridSpaceX))) + 1;
    m_iNumPOIY = int_t(std::floor((m_iROIHeight - m_iSubsetY * 2 - m_iMarginY * 2) / real_t(m_iGridSpaceY))) + 1;
    return m_iNumPOIX > 0 && m_iNumPOIY > 0;
}

// ---------------------------------------------------------------- cuFFTCC2D
cuFFTCC2D::cuFFTCC2D(const int_t iROIWidth, const int_t iROIHeight, const int_t iSubsetX, const int_t iSubsetY,
                     const int_t iGridSpaceX, const int_t iGridSpaceY, const int_t iMarginX, const int_t iMarginY)
    : Fftcc2D(iROIWidth, iROIHeight, iSubsetX, iSubsetY, iGridSpaceX, iGridSpaceY, iMarginX, iMarginY),
      m_iFFTWindow(2 * iSubsetX)
{
    if (!recomputeNumPOI()) throw std::logic_error("Number of POIs is below 0!");
}

cuFFTCC2D::cuFFTCC2D(const int_t iImgWidth, const int_t iImgHeight, const int_t iROIWidth, const int_t iROIHeight,
                     const int_t iStartX, const int_t iStartY, const int_t iSubsetX, const int_t iSubsetY,
                     const int_t iGridSpaceX, const int_t iGridSpaceY, const int_t iMarginX, const int_t iMarginY)
    : Fftcc2D(iImgWidth, iImgHeight, iStartX, iStartY, iROIWidth, iROIHeight, iSubsetX, iSubsetY, iGridSpaceX,
              iGridSpaceY, iMarginX, iMarginY),
      m_iFFTWindow(2 * iSubsetX)
{
    if (!recomputeNumPOI()) throw std::logic_error("Number of POIs is below 0!");
}

cuFFTCC2D::~cuFFTCC2D() { release(); }

void cuFFTCC2D::release()
{
    if (!m_ctx) return;
    dev_free(m_ctx, g_cuHandle.m_d_iPOIXY);
    dev_free(m_ctx, g_cuHandle.m_d_fU);
    dev_free(m_ctx, g_cuHandle.m_d_fV);
    dev_free(m_ctx, g_cuHandle.m_d_fZNCC);
    twdic_destroy(m_ctx);
    m_ctx = nullptr;
    g_cuHandle = GPUHandle_FFTCC();
}

void cuFFTCC2D::setROI(const int_t& iStartX, const int_t& iStartY, const int_t& iROIWidth, const int_t& iROIHeight)
{
    Fftcc2D::setROI(iStartX, iStartY, iROIWidth, iROIHeight);
    if (!recomputeNumPOI()) throw std::logic_error("Number of POIs is below 0!");
    release();
}

void cuFFTCC2D::setFFTWindow(int_t N)
{
    m_iFFTWindow = N;
    release();
}

void cuFFTCC2D::ensureContext()
{
    if (m_ctx) return;
    twdic_config c{};
    if (m_isWholeImgUsed) {
        c.img_w = m_iImgWidth; c.img_h = m_iImgHeight;
        c.roi_x = m_iStartX; c.roi_y = m_iStartY;
    } else {   // the images are the ROI (TW_paDIC_cuFFTCC2D.cu:575-587)
        c.img_w = m_iROIWidth; c.img_h = m_iROIHeight;
        c.roi_x = 0; c.roi_y = 0;
    }
    c.roi_w = m_iROIWidth; c.roi_h = m_iROIHeight;
    c.sx = m_iSubsetX; c.sy = m_iSubsetY;
    c.gx = m_iGridSpaceX; c.gy = m_iGridSpaceY;
    c.mx = m_iMarginX; c.my = m_iMarginY;
    c.fft_n = m_iFFTWindow;
    c.max_iter = 20; c.delta_p = 0.001f;
    c.interp = TWDIC_INTERP_FUSED;
    check(twdic_create(&c, 0, &m_ctx), nullptr, "twdic_create");
    const size_t n = (size_t)GetNumPOIs();
    g_cuHandle.m_d_iPOIXY = dev_alloc<int_t>(m_ctx, 2 * n);
    check(twdic_poi_grid(m_ctx, g_cuHandle.m_d_iPOIXY), m_ctx, "twdic_poi_grid");
    void* p;
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_REF, &p, nullptr), m_ctx, "get_buffer");
    g_cuHandle.m_d_fRefImg = static_cast<uchar1*>(p);
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_TAR, &p, nullptr), m_ctx, "get_buffer");
    g_cuHandle.m_d_fTarImg = static_cast<uchar1*>(p);
    isDestroyed = false;
}

void cuFFTCC2D::ResetRefImg(const MatView& refImg)
{
    ensureContext();
    check(twdic_set_reference(m_ctx, refImg.data, (int)refImg.step, 0), m_ctx, "ResetRefImg");
    // the image buffer is read by other objects' streams (cuICGN2D::cuInitialize):
    // complete the copy before returning, as the reference's cudaMemcpy does
    check(twdic_sync(m_ctx), m_ctx, "sync");
}

void cuFFTCC2D::SetTarImg(const MatView& tarImg)
{
    ensureContext();
    check(twdic_set_target(m_ctx, tarImg.data, (int)tarImg.step, 0), m_ctx, "SetTarImg");
    check(twdic_sync(m_ctx), m_ctx, "sync");
}

void cuFFTCC2D::cuInitializeFFTCC(real_t*& f_d_U, real_t*& f_d_V, real_t*& f_d_ZNCC, const MatView& refImg)
{
    isLowLevelApiCalled = true;
    ensureContext();
    const size_t n = (size_t)GetNumPOIs();
    f_d_U = dev_alloc<real_t>(m_ctx, n);
    f_d_V = dev_alloc<real_t>(m_ctx, n);
    f_d_ZNCC = dev_alloc<real_t>(m_ctx, n);
    ResetRefImg(refImg);
    check(twdic_sync(m_ctx), m_ctx, "sync");
}

void cuFFTCC2D::cuComputeFFTCC(real_t*& f_d_U, real_t*& f_d_V, real_t*& f_d_ZNCC, const MatView& tarImg)
{
    ensureContext();
    SetTarImg(tarImg);
    check(twdic_fftcc(m_ctx, g_cuHandle.m_d_iPOIXY, GetNumPOIs(), f_d_U, f_d_V, f_d_ZNCC, nullptr), m_ctx,
          "cuComputeFFTCC");
    check(twdic_sync(m_ctx), m_ctx, "sync");   // results ready on return (:640)
}

void cuFFTCC2D::cuDestroyFFTCC(real_t*& f_d_U, real_t*& f_d_V, real_t*& f_d_ZNCC)
{
    dev_free(m_ctx, f_d_U);
    dev_free(m_ctx, f_d_V);
    dev_free(m_ctx, f_d_ZNCC);
    release();
    isDestroyed = true;
}

void cuFFTCC2D::InitializeFFTCC(real_t**& fU, real_t**& fV, real_t**& fZNCC, const MatView& refImg)
{
    if (isLowLevelApiCalled) {   // TW_paDIC_cuFFTCC2D.cu:244-249
        std::fprintf(stderr, "The low-level GPU APIs are already initialized!\n");
        return;
    }
    ensureContext();
    hcreateptr<real_t>(fU, m_iNumPOIY, m_iNumPOIX);
    hcreateptr<real_t>(fV, m_iNumPOIY, m_iNumPOIX);
    hcreateptr<real_t>(fZNCC, m_iNumPOIY, m_iNumPOIX);
    const size_t n = (size_t)GetNumPOIs();
    g_cuHandle.m_d_fU = dev_alloc<real_t>(m_ctx, n);
    g_cuHandle.m_d_fV = dev_alloc<real_t>(m_ctx, n);
    g_cuHandle.m_d_fZNCC = dev_alloc<real_t>(m_ctx, n);
    ResetRefImg(refImg);
}

void cuFFTCC2D::ComputeFFTCC(real_t**& fU, real_t**& fV, real_t**& fZNCC, const MatView& tarImg)
{
    ensureContext();
    const int n = GetNumPOIs();
    SetTarImg(tarImg);
    check(twdic_fftcc(m_ctx, g_cuHandle.m_d_iPOIXY, n, g_cuHandle.m_d_fU, g_cuHandle.m_d_fV, g_cuHandle.m_d_fZNCC,
                      nullptr), m_ctx, "ComputeFFTCC");
    check(twdic_memcpy(m_ctx, fU[0], g_cuHandle.m_d_fU, n * sizeof(real_t), 1), m_ctx, "memcpy");
    check(twdic_memcpy(m_ctx, fV[0], g_cuHandle.m_d_fV, n * sizeof(real_t), 1), m_ctx, "memcpy");
    check(twdic_memcpy(m_ctx, fZNCC[0], g_cuHandle.m_d_fZNCC, n * sizeof(real_t), 1), m_ctx, "memcpy");
}

void cuFFTCC2D::DestroyFFTCC(real_t**& fU, real_t**& fV, real_t**& fZNCC)
{
    hdestroyptr(fU);
    hdestroyptr(fV);
    hdestroyptr(fZNCC);
    release();
    isDestroyed = true;
}

// ------------------------------------------------------------------- ICGN2D
ICGN2D::ICGN2D(int_t iImgWidth, int_t iImgHeight, int_t iStartX, int_t iStartY, int_t iROIWidth, int_t iROIHeight,
               int_t iSubsetX, int_t iSubsetY, int_t iNumberX, int_t iNumberY, int_t iNumIterations, real_t fDeltaP)
    : m_isRefImgUpdated(false), m_iImgWidth(iImgWidth), m_iImgHeight(iImgHeight), m_iStartX(iStartX),
      m_iStartY(iStartY), m_iROIWidth(iROIWidth), m_iROIHeight(iROIHeight), m_iSubsetX(iSubsetX),
      m_iSubsetY(iSubsetY), m_iNumberX(iNumberX), m_iNumberY(iNumberY), m_iNumIterations(iNumIterations),
      m_fDeltaP(fDeltaP)
{
    m_iPOINumber = m_iNumberX * m_iNumberY;     // TW_paDIC_ICGN2D.cpp:31-35
    m_iSubsetH = m_iSubsetY * 2 + 1;
    m_iSubsetW = m_iSubsetX * 2 + 1;
    m_iSubsetSize = m_iSubsetH * m_iSubsetW;
}

ICGN2D::ICGN2D(const MatView& refImg, int_t iImgWidth, int_t iImgHeight, int_t iStartX, int_t iStartY,
               int_t iROIWidth, int_t iROIHeight, int_t iSubsetX, int_t iSubsetY, int_t iNumberX, int_t iNumberY,
               int_t iNumIterations, real_t fDeltaP)
    : ICGN2D(iImgWidth, iImgHeight, iStartX, iStartY, iROIWidth, iROIHeight, iSubsetX, iSubsetY, iNumberX, iNumberY,
             iNumIterations, fDeltaP)
{
    m_refImg = refImg;
    m_isRefImgUpdated = true;
}

ICGN2D::~ICGN2D() {}

void ICGN2D::setROI(const int_t& iStartX, const int_t& iStartY, const int_t& iROIWidth, const int_t& iROIHeight)
{
    m_iStartX = iStartX;
    m_iStartY = iStartY;
    m_iROIWidth = iROIWidth;
    m_iROIHeight = iROIHeight;
}

// ----------------------------------------------------------------- cuICGN2D
cuICGN2D::cuICGN2D(int_t iImgWidth, int_t iImgHeight, int_t iStartX, int_t iStartY, int_t iROIWidth,
                   int_t iROIHeight, int_t iSubsetX, int_t iSubsetY, int_t iNumberX, int_t iNumberY,
                   int_t iNumIterations, real_t fDeltaP, ICGN2DInterpolationFLag Iflag)
    : ICGN2D(iImgWidth, iImgHeight, iStartX, iStartY, iROIWidth, iROIHeight, iSubsetX, iSubsetY, iNumberX, iNumberY,
             iNumIterations, fDeltaP),
      m_Iflag(Iflag)
{
    // Bicubic: the LUT gather or the fused evaluation (setInterpolationSource);
    // BicubicSpline: the B-spline LUT of the reference's CPU path
    // (TW_paDIC_ICGN2D_CPU.cpp:322-341; its GPU version is unfinished)
    if (Iflag != ICGN2DInterpolationFLag::Bicubic && Iflag != ICGN2DInterpolationFLag::BicubicSpline)
        throw std::invalid_argument("cuICGN2D: unknown interpolation flag");
    prepare();
}

cuICGN2D::~cuICGN2D() { cuFinalize(); }

void cuICGN2D::prepare()
{
    twdic_config c{};
    c.img_w = m_iImgWidth; c.img_h = m_iImgHeight;
    c.roi_x = m_iStartX; c.roi_y = m_iStartY; c.roi_w = m_iROIWidth; c.roi_h = m_iROIHeight;
    c.sx = m_iSubsetX; c.sy = m_iSubsetY;
    c.gx = 1; c.gy = 1; c.mx = 0; c.my = 0;     // POIs come from the caller
    c.fft_n = 2;
    c.max_iter = m_iNumIterations; c.delta_p = m_fDeltaP;
    c.interp = m_Iflag == ICGN2DInterpolationFLag::BicubicSpline ? TWDIC_INTERP_BSPLINE : m_source;
    check(twdic_create(&c, 0, &m_ctx), nullptr, "twdic_create");
    const size_t n = (size_t)m_iPOINumber;
    g_cuHandleICGN.m_d_iIterationNums = dev_alloc<int>(m_ctx, n);
    g_cuHandleICGN.m_d_dP = dev_alloc<real_t>(m_ctx, 6 * n);
    g_cuHandleICGN.m_d_iFlags = dev_alloc<uint8_t>(m_ctx, n);
    void* p;
    size_t b;
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_REF, &p, &b), m_ctx, "get_buffer");
    g_cuHandleICGN.m_d_fRefImg = static_cast<uchar1*>(p);
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_TAR, &p, &b), m_ctx, "get_buffer");
    g_cuHandleICGN.m_d_fTarImg = static_cast<uchar1*>(p);
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_TX, &p, &b), m_ctx, "get_buffer");
    g_cuHandleICGN.m_d_fTx = static_cast<real_t*>(p);
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_TY, &p, &b), m_ctx, "get_buffer");
    g_cuHandleICGN.m_d_fTy = static_cast<real_t*>(p);
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_TXY, &p, &b), m_ctx, "get_buffer");
    g_cuHandleICGN.m_d_fTxy = static_cast<real_t*>(p);
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_LUT, &p, &b), m_ctx, "get_buffer");
    g_cuHandleICGN.m_d_f4InterpolationLUT = static_cast<real_t4*>(p);
}

void cuICGN2D::setInterpolationSource(int source)
{
    if (source != TWDIC_INTERP_LUT && source != TWDIC_INTERP_FUSED)
        throw std::invalid_argument("cuICGN2D: unknown interpolation source");
    if (source == m_source) return;
    m_source = source;
    cuFinalize();
    prepare();
}

void cuICGN2D::cuFinalize()
{
    if (!m_ctx) return;
    dev_free(m_ctx, g_cuHandleICGN.m_d_iIterationNums);
    dev_free(m_ctx, g_cuHandleICGN.m_d_dP);
    dev_free(m_ctx, g_cuHandleICGN.m_d_iFlags);
    dev_free(m_ctx, m_d_poi);
    dev_free(m_ctx, m_d_scratch);
    twdic_destroy(m_ctx);
    m_ctx = nullptr;
    g_cuHandleICGN = GPUHandle_ICGN();
}

void cuICGN2D::cuInitialize(uchar1* d_fRefImg)
{
    check(twdic_set_reference(m_ctx, reinterpret_cast<const uint8_t*>(d_fRefImg), m_iImgWidth, 1), m_ctx,
          "cuInitialize");
    // the source belongs to another object and may be overwritten once this returns
    check(twdic_sync(m_ctx), m_ctx, "sync");
    m_isRefImgUpdated = true;
    void* p;
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_RX, &p, nullptr), m_ctx, "get_buffer");
    g_cuHandleICGN.m_d_fRx = static_cast<real_t*>(p);
    check(twdic_get_buffer(m_ctx, TWDIC_BUF_RY, &p, nullptr), m_ctx, "get_buffer");
    g_cuHandleICGN.m_d_fRy = static_cast<real_t*>(p);
}

void cuICGN2D::Initialize(MatView refImg) { ResetRefImg(refImg); }

void cuICGN2D::ResetRefImg(const MatView& refImg)
{
    if (refImg.cols != m_iImgWidth || refImg.rows != m_iImgHeight) {   // TW_paDIC_cuICGN2D.cu:783-787
        std::fprintf(stderr, "Image Dimension Not Match\n");
        return;
    }
    check(twdic_set_reference(m_ctx, refImg.data, (int)refImg.step, 0), m_ctx, "ResetRefImg");
    check(twdic_sync(m_ctx), m_ctx, "sync");
    m_refImg = refImg;
    m_isRefImgUpdated = true;
}

void cuICGN2D::SetTarImg(const MatView& tarImg)
{
    if (tarImg.cols != m_iImgWidth || tarImg.rows != m_iImgHeight) {
        std::fprintf(stderr, "Image Dimension Not Match\n");
        return;
    }
    // the reference copies into the reference buffer here (TW_paDIC_cuICGN2D.cu:805-808);
    // this class sets the target
    check(twdic_set_target(m_ctx, tarImg.data, (int)tarImg.step, 0), m_ctx, "SetTarImg");
    check(twdic_sync(m_ctx), m_ctx, "sync");
    m_tarImg = tarImg;
}

void cuICGN2D::cuCompute(uchar1* d_fTarImg, int_t* d_iPOIXY, real_t* d_fU, real_t* d_fV)
{
    g_cuHandleICGN.m_d_iPOIXY = d_iPOIXY;   // borrowed, as in the reference (:823-826)
    g_cuHandleICGN.m_d_fU = d_fU;
    g_cuHandleICGN.m_d_fV = d_fV;
    check(twdic_set_target(m_ctx, reinterpret_cast<const uint8_t*>(d_fTarImg), m_iImgWidth, 1), m_ctx, "cuCompute");
    check(twdic_icgn(m_ctx, d_iPOIXY, m_iPOINumber, d_fU, d_fV, g_cuHandleICGN.m_d_dP,
                     g_cuHandleICGN.m_d_iIterationNums, g_cuHandleICGN.m_d_iFlags, nullptr),
          m_ctx, "cuCompute");
    check(twdic_sync(m_ctx), m_ctx, "sync");   // (:961)
    m_isRefImgUpdated = false;
}

void cuICGN2D::ICGN2D_Algorithm(real_t* fU, real_t* fV, int* iNumIterations, int* iPOIpos, const MatView& tarImg,
                                ICGN2DFlag* flags)
{
    const size_t n = (size_t)m_iPOINumber;
    SetTarImg(tarImg);
    if (!m_d_poi) {
        m_d_poi = dev_alloc<int>(m_ctx, 2 * n);
        m_d_scratch = dev_alloc<real_t>(m_ctx, 2 * n);
    }
    // the POI buffer is rewritten in place on every call: drop the cached state
    check(twdic_memcpy(m_ctx, m_d_poi, iPOIpos, 2 * n * sizeof(int), 0), m_ctx, "memcpy");
    check(twdic_reset_poi_state(m_ctx), m_ctx, "reset");
    real_t* du = m_d_scratch;
    real_t* dv = m_d_scratch + n;
    check(twdic_memcpy(m_ctx, du, fU, n * sizeof(real_t), 0), m_ctx, "memcpy");
    check(twdic_memcpy(m_ctx, dv, fV, n * sizeof(real_t), 0), m_ctx, "memcpy");
    check(twdic_memcpy(m_ctx, g_cuHandleICGN.m_d_iIterationNums, iNumIterations, n * sizeof(int), 0), m_ctx, "memcpy");
    check(twdic_icgn(m_ctx, m_d_poi, (int)n, du, dv, g_cuHandleICGN.m_d_dP, g_cuHandleICGN.m_d_iIterationNums,
                     g_cuHandleICGN.m_d_iFlags, nullptr),
          m_ctx, "ICGN2D_Algorithm");
    check(twdic_memcpy(m_ctx, fU, du, n * sizeof(real_t), 1), m_ctx, "memcpy");
    check(twdic_memcpy(m_ctx, fV, dv, n * sizeof(real_t), 1), m_ctx, "memcpy");
    check(twdic_memcpy(m_ctx, iNumIterations, g_cuHandleICGN.m_d_iIterationNums, n * sizeof(int), 1), m_ctx, "memcpy");
    if (flags) {
        std::vector<uint8_t> f(n);
        check(twdic_memcpy(m_ctx, f.data(), g_cuHandleICGN.m_d_iFlags, n, 1), m_ctx, "memcpy");
        for (size_t i = 0; i < n; ++i) flags[i] = static_cast<ICGN2DFlag>(f[i]);
    }
}

void cuICGN2D::ICGN2D_Algorithm(real_t* fU, real_t* fV, int* iNumIterations, int* iPOIpos, const MatView& tarImg,
                                float& fPrecomputeTime, float& fICGNTime)
{
    // the precompute is fused into the per-frame work; its time is folded into
    // fICGNTime (TW_paDIC_ICGN2D_CPU.cpp:145-192 times the two separately)
    auto t0 = std::chrono::steady_clock::now();
    ICGN2D_Algorithm(fU, fV, iNumIterations, iPOIpos, tarImg, nullptr);
    auto t1 = std::chrono::steady_clock::now();
    fPrecomputeTime = 0.f;
    fICGNTime = std::chrono::duration<float, std::milli>(t1 - t0).count();
}

void cuICGN2D::ICGN2D_Algorithm(real_t* fU, real_t* fV, int* iNumIterations, int* iPOIpos, const MatView& tarImg,
                                int_t head, int_t tail)
{
    (void)head;
    (void)tail;
    ICGN2D_Algorithm(fU, fV, iNumIterations, iPOIpos, tarImg, nullptr);
}

}  // namespace paDIC
}  // namespace TW
