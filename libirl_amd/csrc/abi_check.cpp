// abi_check.cpp -- compile-time check that the exported structs keep the
// reference's x86-64 layout.  The sizes and offsets below were measured by
// compiling the reference headers (irl.h, osemhooks.h, drftabio.h,
// irlprivate.h) with gcc on x86-64.
#include <cstddef>

#include "../../include/irl.h"
#include "../../include/irlprj.h"
#include "../../include/osemhooks.h"

static_assert(sizeof(IrlParms_t) == 72, "IrlParms_t size");
static_assert(offsetof(IrlParms_t, pchNormImageBase) == 40, "IrlParms_t layout");
static_assert(offsetof(IrlParms_t, SrfCollapseFac) == 48, "IrlParms_t layout");
static_assert(offsetof(IrlParms_t, fIntrinsicFWHM) == 64, "IrlParms_t layout");
static_assert(sizeof(Options_t) == 36, "Options_t size");
static_assert(offsetof(Options_t, fAtnMapThresh) == 20, "Options_t layout");
static_assert(offsetof(Options_t, iAxialAvgLength) == 32, "Options_t layout");
static_assert(sizeof(PrjView_t) == 16, "PrjView_t size");
static_assert(offsetof(PrjView_t, CFCR) == 12, "PrjView_t layout");
static_assert(sizeof(sOsemData) == 48, "sOsemData size");
static_assert(offsetof(sOsemData, pfEst) == 16, "sOsemData layout");
static_assert(offsetof(sOsemData, pfUpdate) == 40, "sOsemData layout");
static_assert(sizeof(sAdvOptions) == 8, "sAdvOptions size");
static_assert(sizeof(Drf_t) == 24, "Drf_t size");
static_assert(offsetof(Drf_t, Psf) == 16, "Drf_t layout");
static_assert(sizeof(DrfTab_t) == 72, "DrfTab_t size");
static_assert(offsetof(DrfTab_t, psBckDrfTab) == 24, "DrfTab_t layout");
static_assert(offsetof(DrfTab_t, fMaxFracErr) == 48, "DrfTab_t layout");
static_assert(offsetof(DrfTab_t, psDrfs) == 56, "DrfTab_t layout");
static_assert(offsetof(DrfTab_t, bFFTconvolve) == 64, "DrfTab_t layout");
static_assert(sizeof(Srf_t) == 16, "Srf_t size");
static_assert(offsetof(Srf_t, pvScatterData) == 8, "Srf_t layout");
