/*
 * osemhooks.h -- OSEM hook, advanced-option and detector API of libirl_amd.
 *
 * Drop-in for the reference header osemhooks.h:
 *   sOsemData             replaces osemhooks.h:18-27
 *   sAdvOptions           replaces osemhooks.h:29-33
 *   OsemSetAdvOptions     replaces osemhooks.h:35     (osemrecon.c:44-48)
 *   OsemSet*Hook          replace  osemhooks.h:37-44  (osemrecon.c:50-90)
 *   iSetupExtraPrjViewData replaces osemhooks.h:53-54 (IrlSetup.c:350-402)
 *   iSetupDetectors       replaces osemhooks.h:62     (IrlSetup.c:441-462)
 *   pvStoreDetectorData   replaces osemhooks.h:65-79  (IrlSetup.c:415-439)
 *
 * Hooks receive host pointers to the estimate, update, sensitivity and
 * scatter buffers in the internal [y][z][x] order.  The library copies the
 * device buffers to these host mirrors before each registered hook and back
 * afterwards, so a hook may read and modify them.  With no hook registered
 * nothing leaves the GPU inside the iteration loop.
 */
#ifndef OSEMHOOKS_H
#define OSEMHOOKS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _OsemData {
    int iIterationIdx;
    int iSubsetIdx;
    int iAngleIdx;
    float *pfEst;
    float *pfSct;
    float *pfNorm;
    float *pfUpdate;
} sOsemData, *psOsemData;

typedef struct _AdvOptions {
    int bFastRotate;      /* 1 (default): three-shear rotation (fastrotate.c), 0: bilinear (rot3dpar.c) */
    int iStartIteration;
} sAdvOptions, *psAdvOptions;

int OsemSetAdvOptions(sAdvOptions *pOpt);

int OsemSetPrevIterHook(void (*pPrevIterHookFunc)(sOsemData *psData, void *ctx), void *ctx);
int OsemSetPostIterHook(void (*pPostIterHookFunc)(sOsemData *psData, void *ctx), void *ctx);
int OsemSetPrevSubsetHook(void (*pPrevSubsetHookFunc)(sOsemData *psData, void *ctx), void *ctx);
int OsemSetPostSubsetHook(void (*pPostSubsetHookFunc)(sOsemData *psData, void *ctx), void *ctx);
int OsemSetPrevAngleHook(void (*pPrevAngleHookFunc)(sOsemData *psData, void *ctx), void *ctx);
int OsemSetPostAngleHook(void (*pPostAngleHookFunc)(sOsemData *psData, void *ctx), void *ctx);

int iSetupExtraPrjViewData(int iNumPrjViews, float *pfAcqTimes, int *piDetectorNums);
int iSetupDetectors(int iNumDetectors, void **ppsDetectors);
void *pvStoreDetectorData(float fHoleDiam, float fHoleLen, float fBackToDet,
                          float fIntrinsicFWHM, float fSensitivity, int bFFTconvolve,
                          int bUseGrfInBck, char *pchDrfTabFname);

#ifdef __cplusplus
}
#endif

#endif /* OSEMHOOKS_H */
