// prjbck_api.cpp -- projector-level C entry points (include/irlprj.h).
//
// InitPrj / Project / BackProject / FreePrj keep the workspace semantics of
// prjbck.c:29-453; rot3dpar and ConvolveDrfYpad run one rotation or one plane
// convolution on the GPU.  The per-row DRF plane indices are computed here
// with the reference's float arithmetic (prjbck.c:186-204, 343-360).
#include <algorithm>
#include <cstring>
#include <map>

#include "../../include/irl_amd.h"
#include "esse.h"
#include "irlprj.h"

using namespace irlamd;

namespace {

struct CachedDrf {
    PaddedDrf pad;
    int nDist;
    Drf_t *drfs;
};

std::map<const DrfTab_t *, CachedDrf> g_drfCache;

const PaddedDrf &cached_drf(const DrfTab_t *t)
{
    auto it = g_drfCache.find(t);
    if (it != g_drfCache.end() && it->second.nDist == t->iNumDistances &&
        it->second.drfs == t->psDrfs)
        return it->second.pad;
    CachedDrf c{pad_drf_table(t), t->iNumDistances, t->psDrfs};
    g_drfCache[t] = std::move(c);
    return g_drfCache[t].pad;
}

void report(const IrlError &e, const char *fn)
{
    set_error_string(e.msg);
    set_last_error(e.code);
    fprintf(stderr, "%s: %s\n", fn, e.msg.c_str());
}

// Builds the per-row tables and rotation constants of one call.
PrjCall make_call(int Nz, int N, int NRot, int model, float pw, float fAngle, float fCFCR,
                  const DrfTab_t *t, float primary, float atnFac, int pad, int avg,
                  int atnRot, int actRot, std::vector<int> &rk, std::vector<int> &rp)
{
    PrjCall c{};
    c.Nz = Nz;
    c.N = N;
    c.NRot = NRot;
    c.model = model;
    ViewTab vt;
    view_trig(fAngle, vt);
    c.fsin = vt.fsin;
    c.fcos = vt.fcos;
    c.bsin = vt.bsin;
    c.bcos = vt.bcos;
    c.fang = vt.fang;
    c.bang = vt.bang;
    // prjbck.c:127, 317: with scatter modelling the primary factor is applied
    // by the scatter source instead.
    c.primary = (model & MODEL_SRF) ? 1.0f : primary;
    c.primaryFac = primary;
    c.esse = nullptr;
    c.atnFac = atnFac;
    c.avg = avg;
    c.atnRotated = atnRot;
    c.actRotated = actRot;
    rk.assign(NRot, -1);
    rp.assign(NRot, 0);
    c.drf = nullptr;
    c.drfKey = nullptr;
    if ((model & MODEL_DRF) && t) {
        c.drf = &cached_drf(t);
        c.drfKey = t;
        for (int r = 0; r < NRot; ++r) {
            bool skip;
            const int d = drf_row_index(t, fCFCR, pw, NRot, r, &skip);
            if (skip) continue;
            rk[r] = d;
            int p = std::max(pad, 0);
            p = std::min(p, t->psDrfs[d].MaxOffsY - 1);
            if (t->bFFTconvolve) p = 0;
            rp[r] = p;
        }
    } else if (model & MODEL_DRF) {
        fail(ETYPE_ILLEGAL_VALUE, "DRF modelling requested without a DRF table");
    }
    c.rowKernel = rk.data();
    c.rowPad = rp.data();
    return c;
}

}  // namespace

extern "C" {

// prjbck.c:42-66
void InitPrj(int iNumSlices, int iNumPixels, int iNumRotPixs, int iModel)
{
    set_last_error(0);
    try {
        g_drfCache.clear();
        prjws_init(iNumSlices, iNumPixels, iNumRotPixs, iModel);
    } catch (const IrlError &e) {
        report(e, "InitPrj");
    }
}

// prjbck.c:432-453
void FreePrj(void)
{
    try {
        g_drfCache.clear();
        prjws_free();
    } catch (const IrlError &e) {
        report(e, "FreePrj");
    }
}

// prjbck.c:68-252
void Project(int iNumSlices, int iNumPixels, int iNumRotPixs, int iPrjModel, float fPixelWidth,
             int iAngle, float fAngle, float fCFCR, DrfTab_t *psDrfTab, Srf_t *psSrf,
             float *atnmatrix, float *reconmatrix, float *calcprj, float fPrimaryFac,
             float fAtnFac, int iAxialPadLength, int iAxialAvgLength, int bAtnMapRotated,
             int bActImgRotated, int bFastRotate)
{
    (void)iAngle;
    try {
        std::vector<int> rk, rp;
        PrjCall c = make_call(iNumSlices, iNumPixels, iNumRotPixs, iPrjModel, fPixelWidth,
                              fAngle, fCFCR, psDrfTab, fPrimaryFac, fAtnFac, iAxialPadLength,
                              iAxialAvgLength, bAtnMapRotated, bActImgRotated, rk, rp);
        c.fast = bFastRotate != 0;   // prjbck.c:130-138, 154-161
        if (iPrjModel & MODEL_SRF) c.esse = psSrf ? psSrf->pvScatterData : nullptr;
        prjws_project(c, atnmatrix, reconmatrix, calcprj);
    } catch (const IrlError &e) {
        report(e, "Project");
    }
}

// prjbck.c:254-427
void BackProject(int iNumSlices, int iNumPixels, int iNumRotPixs, int iBckModel,
                 float fPixelWidth, int iAngle, float fAngle, float fCFCR, DrfTab_t *psDrfTab,
                 Srf_t *psSrf, float *atnmatrix, float *ratioprj, float *imagematrix,
                 float fPrimaryFac, float fAtnFac, int iAxialPadLength, int iAxialAvgLength,
                 int bAtnMapRotated, int bFastRotate)
{
    (void)iAngle;
    try {
        // prjbck.c:308-310: a separate backprojection table wins when present
        if (psDrfTab && psDrfTab->psBckDrfTab) psDrfTab = psDrfTab->psBckDrfTab;
        std::vector<int> rk, rp;
        PrjCall c = make_call(iNumSlices, iNumPixels, iNumRotPixs, iBckModel, fPixelWidth,
                              fAngle, fCFCR, psDrfTab, fPrimaryFac, fAtnFac, iAxialPadLength,
                              iAxialAvgLength, bAtnMapRotated, 0, rk, rp);
        c.fast = bFastRotate != 0;   // prjbck.c:327-334, 408-416
        if (iBckModel & MODEL_SRF) c.esse = psSrf ? psSrf->pvScatterData : nullptr;
        prjws_backproject(c, atnmatrix, ratioprj, imagematrix);
    } catch (const IrlError &e) {
        report(e, "BackProject");
    }
}

// rot3dpar.c:19-198
void rot3dpar(float fAngle, float *OldMatrix, float *NewMatrix, int PB_FLAG, int NumSlices,
              int NumPixels, int NumRotPixs)
{
    try {
        dev_rot3dpar(fAngle, OldMatrix, NewMatrix, PB_FLAG, NumSlices, NumPixels, NumRotPixs);
    } catch (const IrlError &e) {
        report(e, "rot3dpar");
    }
}

// fastrotate.c:330-563 (the circle extents of vInitFastRotate,
// fastrotate.c:573-589, are computed per call)
void vShearsRotate(float fAngle, float *pfOldMatrix, float *pfRotMatrix, int PB_FLAG,
                   int NumSlices, int NumPixels)
{
    (void)PB_FLAG;
    try {
        if (NumSlices <= 0 || NumPixels <= 0)
            fail(ETYPE_ILLEGAL_VALUE, "vShearsRotate: bad sizes %d %d", NumSlices, NumPixels);
        dev_shears_rotate(fAngle, pfOldMatrix, pfRotMatrix, NumSlices, NumPixels);
    } catch (const IrlError &e) {
        report(e, "vShearsRotate");
    }
}

// convolve_ypad.c:22-96
void ConvolveDrfYpad(float *pfIn, int inX, int inY, Drf_t *psDrf, int bFFTConvolve, int iYpad,
                     int iYavg, float *pfOut)
{
    try {
        if (!psDrf) fail(ETYPE_ILLEGAL_VALUE, "can't convolve using NULL DRF");
        DrfTab_t one;
        memset(&one, 0, sizeof one);
        one.iNumDistances = 1;
        one.psDrfs = psDrf;
        PaddedDrf d = pad_drf_table(&one);
        int pad = std::min(std::max(iYpad, 0), psDrf->MaxOffsY - 1);
        if (bFFTConvolve) {
            // the FFT branch is the same linear convolution without axial pad
            pad = 0;
            iYavg = 0;
        }
        dev_convolve_plane(pfIn, inX, inY, d, pad, iYavg, pfOut);
    } catch (const IrlError &e) {
        report(e, "ConvolveDrfYpad");
    }
}

// ConvolveDrfYpad with no device allocation (no counterpart in the
// reference): the same kernel and arguments, staged in the library's
// file-scope device array.
int ConvolveDrfYpadNoAlloc(float *pfIn, int inX, int inY, Drf_t *psDrf, int bFFTConvolve,
                           int iYpad, int iYavg, float *pfOut)
{
    set_last_error(0);
    try {
        if (!psDrf) fail(ETYPE_ILLEGAL_VALUE, "can't convolve using NULL DRF");
        if (psDrf->bFFTdone) fail(ETYPE_ILLEGAL_VALUE, "DRF is already Fourier transformed");
        if (psDrf->MaxOffsX < 1 || psDrf->MaxOffsY < 1)
            fail(ETYPE_ILLEGAL_VALUE, "DRF has illegal offsets %d x %d", psDrf->MaxOffsX,
                 psDrf->MaxOffsY);
        int pad = std::min(std::max(iYpad, 0), psDrf->MaxOffsY - 1);
        if (bFFTConvolve) {
            pad = 0;
            iYavg = 0;
        }
        dev_convolve_plane_noalloc(pfIn, inX, inY, psDrf->Psf, 2 * psDrf->MaxOffsX - 1,
                                   2 * psDrf->MaxOffsY - 1, pad, iYavg, pfOut);
        return 0;
    } catch (const IrlError &e) {
        report(e, "ConvolveDrfYpadNoAlloc");
        return e.code;
    }
}

}  // extern "C"

extern "C" {

// krnlprj3d.c:956-1009.  Returns NULL when the kernel's scatter factor is 0
// (no scatter data, krnlprj3d.c:428), as the reference does.
void *pvInitESSEScatterModel(char *pchKrnlFile, char *pchParString, float fPixSize, int nPix,
                             int nSlices, int iCollapseFac, int bDebugFlag)
{
    (void)bDebugFlag;
    set_last_error(0);
    try {
        if (iCollapseFac < 1 || iCollapseFac > 4)
            fail(ETYPE_ILLEGAL_VALUE, "collapse factor %d: must be 1 to 4", iCollapseFac);
        KrnlData k = read_krnl_file(pchKrnlFile);
        EsseModel m = make_esse_model(k, pchParString, fPixSize, nPix, nSlices, iCollapseFac);
        if (!m.enabled) return nullptr;
        return esse_handle_create(m);
    } catch (const IrlError &e) {
        report(e, "pvInitESSEScatterModel");
        return nullptr;
    }
}

// krnlprj3d.c:676-954 on host arrays [ny][nz][nx]: nx bins, ny depth rows,
// nz slices.
void CalcEffScatterSource(void *pvScatterData, float *pfAct, float *pfMap, int nx, int ny,
                          int nz, float *pfEffSrc)
{
    try {
        esse_handle_effsrc(pvScatterData, pfAct, pfMap, nx, ny, nz, pfEffSrc);
    } catch (const IrlError &e) {
        report(e, "CalcEffScatterSource");
    }
}

// krnlprj3d.c:377-401
void *pvDeinitScatterPrj(void *pvScatterData)
{
    esse_handle_free(pvScatterData);
    return nullptr;
}

}  // extern "C"
