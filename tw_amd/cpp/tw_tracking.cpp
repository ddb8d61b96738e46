// tw_tracking.cpp -- TW_Core's POI-tracking helpers (include/tw/TW_cuda_utils.h)
// over the C ABI.
#include <stdexcept>
#include <string>

#include "tw/TW_cuda_utils.h"
#include "twdic.h"

namespace {
void check(twdic_status st, const char* what)
{
    if (st != TWDIC_OK) throw std::runtime_error(std::string(what) + ": status " + std::to_string((int)st));
}
}  // namespace

void cuUpdatePOIpos(TW::real_t* fU, TW::real_t* fV, int iNumberX, int iNumberY, int* iPOIpos)
{
    check(twdic_track_update_poi(nullptr, iPOIpos, iNumberX * iNumberY, fU, fV), "cuUpdatePOIpos");
}

void cuAccumulatePOI(TW::real_t* fCurrentU, TW::real_t* fCurrentV, int* iCurrentPOIXY, int iNumPOI, int* iPOIXY)
{
    check(twdic_track_current_poi(nullptr, iCurrentPOIXY, iNumPOI, fCurrentU, fCurrentV, iPOIXY), "cuAccumulatePOI");
}

void cuAccumulatePOI(TW::real_t* fU, TW::real_t* fV, TW::real_t* fAccumulateU, TW::real_t* fAccumulateV,
                     int* iCurrentPOIXY, int iNumPOI, int* iPOIXY)
{
    check(twdic_track_current_poi_acc(nullptr, iCurrentPOIXY, iNumPOI, fU, fV, fAccumulateU, fAccumulateV, iPOIXY),
          "cuAccumulatePOI");
}

void cuAccumulateUV(TW::real_t* fCurrentU, TW::real_t* fCurrentV, int iPOINum, TW::real_t* fU, TW::real_t* fV)
{
    check(twdic_track_accumulate_uv(nullptr, iPOINum, fU, fV, fCurrentU, fCurrentV), "cuAccumulateUV");
}
