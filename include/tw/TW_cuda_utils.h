// TW_cuda_utils.h -- the POI-tracking helpers TW_Core calls around a reference
// reset (TW/TW_Core/cuda_utils.cuh:44-75), same names and signatures, over
// the C ABI in twdic.h.  Device pointers in, device pointers out; each call
// completes before it returns (the reference launches on the legacy default
// stream).  int() truncates toward zero, as the reference.
//
// Not provided: the colour-map and min/max helpers (display only).
#ifndef TWAMD_TW_CUDA_UTILS_H
#define TWAMD_TW_CUDA_UTILS_H

#include "TW.h"

// iPOIpos[i] += (int(fV[i]), int(fU[i])) for the iNumberX * iNumberY POIs
// ({y, x} pairs).  Call twdic_reset_poi_state (or reset the IC-GN reference)
// on a context that caches per-POI state for this array.
void cuUpdatePOIpos(TW::real_t* fU, TW::real_t* fV, int iNumberX, int iNumberY, int* iPOIpos);

// iPOIXY[i] = iCurrentPOIXY[i] + (int(fCurrentV[i]), int(fCurrentU[i])).
void cuAccumulatePOI(TW::real_t* fCurrentU, TW::real_t* fCurrentV, int* iCurrentPOIXY, int iNumPOI, int* iPOIXY);

// iPOIXY[i] = iCurrentPOIXY[i] + (int(fV[i]) - int(fAccumulateU[i]),
// int(fU[i]) - int(fAccumulateV[i])): the accumulated U is taken from y and
// the accumulated V from x, as the reference's accumulatePOINew_Kernel
// (cuda_utils.cu:203-223) pairs them.
void cuAccumulatePOI(TW::real_t* fU, TW::real_t* fV, TW::real_t* fAccumulateU, TW::real_t* fAccumulateV,
                     int* iCurrentPOIXY, int iNumPOI, int* iPOIXY);

// fCurrentU[i] += fU[i], fCurrentV[i] += fV[i] (the first pair accumulates).
void cuAccumulateUV(TW::real_t* fCurrentU, TW::real_t* fCurrentV, int iPOINum, TW::real_t* fU, TW::real_t* fV);

#endif
