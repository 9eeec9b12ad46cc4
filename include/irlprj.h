/*
 * irlprj.h -- projector-level C entry points of libirl_amd.
 *
 * These are the reference's internal projector/backprojector functions,
 * exported with host pointers so that a caller (or a parity test) can drive
 * one view at a time exactly as vOsemRecon does.  Each call copies its inputs
 * to the GPU, runs the gfx950 kernels and copies the result back.
 *
 *   Drf_t, DrfTab_t        replace drftabio.h:7-28 (same layout)
 *   Srf_t                  replaces irlprivate.h:116-120 (same layout)
 *   InitPrj                replaces irlprivate.h:461-466 (prjbck.c:42-66)
 *   Project                replaces irlprivate.h:468-489 (prjbck.c:68-252)
 *   BackProject            replaces irlprivate.h:491-511 (prjbck.c:254-427)
 *   FreePrj                replaces irlprivate.h:517     (prjbck.c:432-453)
 *   rot3dpar               replaces irlprivate.h:521-529 (rot3dpar.c:19-198)
 *   vShearsRotate          replaces irlprivate.h:255-262 (fastrotate.c:330-563)
 *   gendrf                 replaces irlprivate.h:272-285 (gendrf.c:221-427)
 *   ConvolveDrfYpad        replaces irlprivate.h:233-242 (convolve_ypad.c:22-96)
 *   sReadDrfTab            replaces drftabio.h:37-41     (drftabio.c:150-275)
 *   vWriteDrfTab           replaces drftabio.h:30-35     (drftabio.c:57-148)
 *   pvInitESSEScatterModel replaces krnlprj.h:9-10       (krnlprj3d.c:956-1009)
 *   CalcEffScatterSource   replaces krnlprj.h:12-13      (krnlprj3d.c:676-954)
 *   pvDeinitScatterPrj     replaces krnlprj.h:15         (krnlprj3d.c:377-401)
 *
 * Workspace semantics follow prjbck.c: after Project the rotated estimate
 * and rotated attenuation map stay in the library's workspace, so a later
 * Project(..., bActImgRotated=1) or BackProject(..., bAtnMapRotated=1) for
 * the same angle reuses them, and BackProject leaves its rotated planes in
 * the rotated-estimate workspace.  A DrfTab_t passed to Project/BackProject
 * is copied to the GPU on first use and cached until FreePrj/InitPrj; it
 * must not be modified in between.  bFastRotate selects the three-shear
 * rotation (vShearsRotate, which needs iNumRotPixs == iNumPixels) instead of
 * rot3dpar, as in the reference.
 */
#ifndef IRLPRJ_H
#define IRLPRJ_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int MaxOffsX;  /* PSF is (2*MaxOffsX-1) x (2*MaxOffsY-1), x fastest */
    int MaxOffsY;
    int bFFTdone;
    float *Psf;
} Drf_t;

typedef struct DrfTab_s {
    int iNumDistances;
    float fStartDistance;
    float fDrfSpacing;
    float fPixSize;
    int bIncKrnl;
    int bDrfFromFile;
    struct DrfTab_s *psBckDrfTab;
    float fGap, fHoleDiam, fColThick, fIntrinsicFWHM, fMaxFracErr;
    Drf_t *psDrfs;
    int bFFTconvolve;
} DrfTab_t;

typedef struct {
    int iSrfCollapseFac;
    int iSrfParmLineNum;
    void *pvScatterData;
} Srf_t;

void InitPrj(int iNumSlices, int iNumPixels, int iNumRotPixs, int iModel);

void Project(int iNumSlices, int iNumPixels, int iNumRotPixs, int iPrjModel,
             float fPixelWidth, int iAngle, float fAngle, float fCFCR,
             DrfTab_t *psDrfTab, Srf_t *psSrf, float *atnmatrix, float *reconmatrix,
             float *calcprj, float fPrimaryFac, float fAtnFac, int iAxialPadLength,
             int iAxialAvgLength, int bAtnMapRotated, int bActImgRotated,
             int bFastRotate);

void BackProject(int iNumSlices, int iNumPixels, int iNumRotPixs, int iBckModel,
                 float fPixelWidth, int iAngle, float fAngle, float fCFCR,
                 DrfTab_t *psDrfTab, Srf_t *psSrf, float *atnmatrix, float *ratioprj,
                 float *imagematrix, float fPrimaryFac, float fAtnFac,
                 int iAxialPadLength, int iAxialAvgLength, int bAtnMapRotated,
                 int bFastRotate);

void FreePrj(void);

void rot3dpar(float fAngle, float *OldMatrix, float *NewMatrix, int PB_FLAG,
              int NumSlices, int NumPixels, int NumRotPixs);

/* Three-shear rotation of a [y][z][x] volume of NumPixels x NumSlices x
 * NumPixels (PB_FLAG is unused, as in the reference).  The circle extents
 * that the reference sets up in vInitFastRotate are computed per call. */
void vShearsRotate(float fAngle, float *pfOldMatrix, float *pfRotMatrix, int PB_FLAG,
                   int NumSlices, int NumPixels);

/* Table of NumDists computed parallel-hole GRFs (round holes of diameter
 * fHoleDiam, length fColThick, gap fGap to the image plane, Gaussian
 * intrinsic blur fIntrinsic FWHM, sampled on fBinSize pixels) for distances
 * minDist + k*fDrfSpacing; each truncated to the smallest centred square
 * holding 1 - fMaxErr of its area (fMaxErr <= 0: NOffs) and renormalised.
 * DrfTab[k].Psf is malloc'ed.  Only Type 0 (PARALLEL) is provided. */
void gendrf(int Type, float Rfan, float fColThick, float fHoleDiam, float fGap, float fIntrinsic,
            float fMaxErr, float fBinSize, float fDrfSpacing, float minDist, int NumDists,
            int NOffs, Drf_t *DrfTab);

void ConvolveDrfYpad(float *pfIn, int inX, int inY, Drf_t *psDrf, int bFFTConvolve,
                     int iYpad, int iYavg, float *pfOut);

/* No counterpart in the reference.  ConvolveDrfYpad for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc, processes with HIP graph capture in progress elsewhere): the
 * same kernel and result, with the PSF, the planes and the row tables staged
 * in a file-scope device array of the library (one per GPU, 540 KB, held
 * until the process ends) instead of allocated buffers.  Limits: inX * inY
 * <= 65536 values, PSF at most 63 x 63 (MaxOffsX, MaxOffsY <= 32).  Runs on
 * the current device's null stream and returns once pfOut holds the result.
 * Calls that overlap on one GPU share its array and race.  Returns 0, or
 * nonzero with pchIrlErrorString() set (sizes beyond the limits, NULL or
 * invalid DRF). */
int ConvolveDrfYpadNoAlloc(float *pfIn, int inX, int inY, Drf_t *psDrf, int bFFTConvolve,
                           int iYpad, int iYavg, float *pfOut);

DrfTab_t sReadDrfTab(char *pchInFname, char **ppchInTitle, char **ppchInColName);

void vWriteDrfTab(DrfTab_t *psOutDrfTab, char *pchOutFname, char *pchTitle,
                  char *pchColName);

/* ESSE scatter model.  pvInitESSEScatterModel reads a KRNL file and prepares
 * the model for nPix x nPix x nSlices volumes; the parameter string is
 * "xypadfac zpadfac numexpterms geometry" ("2 2 2 3" in IrlOsem).  The
 * returned handle goes into Srf_t.pvScatterData for Project/BackProject with
 * MODEL_SRF.  Images are [ny][nz][nx]: nx bins, ny depth rows, nz slices. */
void *pvInitESSEScatterModel(char *pchKrnlFile, char *pchParString, float fPixSize, int nPix,
                             int nSlices, int iCollapseFac, int bDebugFlag);
void CalcEffScatterSource(void *pvScatterData, float *pfAct, float *pfMap, int nx, int ny,
                          int nz, float *pfEffSrc);
void *pvDeinitScatterPrj(void *pvScatterData);

#ifdef __cplusplus
}
#endif

#endif /* IRLPRJ_H */
