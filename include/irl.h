/*
 * irl.h -- public C API of libirl_amd, the MI355X implementation of libirl's
 * OSEM reconstruction path.
 *
 * Drop-in for the reference header irl.h: the structs below have the same
 * field order, types and sizes (checked by static_assert in the library
 * against offsets measured on the reference header), and the functions have
 * the same names, arguments, return codes and side effects on the caller's
 * arrays.
 *
 *   IrlParms_t            replaces irl.h:39-79
 *   Options_t             replaces irl.h:81-124
 *   PrjView_t             replaces irl.h:127-143
 *   IrlOsem               replaces irl.h:145-213 (irlosem.c:22-285)
 *   IrlGenprj             replaces irl.h:215-288 (irlgenprj.c:23-309)
 *   ComputeOrbit          replaces irl.h:291-316 (computeorbit.c:21-150)
 *   pchIrlErrorString     replaces irl.h:318     (irlosem.c:287-290)
 *
 * Volumes are x-fastest, then y, then z (slice) ("external" order); projection
 * arrays are bin-fastest, then slice, then view, with NumPixels*NumSlices*
 * NumViews elements.
 */
#ifndef IRL_H
#define IRL_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef TRUE
#define TRUE 1
#endif
#ifndef FALSE
#define FALSE 0
#endif

/* effect masks (irl.h:29-31) */
#define MODEL_ATN 1
#define MODEL_DRF 2
#define MODEL_SRF 4

typedef struct {
    int NumPixels;          /* transaxial matrix size = number of bins */
    int NumSlices;
    int NumViews;
    float BinWidth;         /* cm; pixel width and slice thickness equal it */
    float fAtnScaleFac;     /* scales the attenuation map to 1/pixel */
    int iNumSrfIterations;  /* iterations that model scatter (ESSE) */
    float fScatEstFac;      /* multiplies pfScatterEstimate */
    int NumAngPerSubset;    /* views per subset; NumViews/NumAngPerSubset is 1 or even */
    int NumIterations;
    char *pchNormImageBase; /* if non-NULL, sensitivity images go to <base>.norm.<s>.dat */
    int SrfCollapseFac;     /* 1..4 */
    float fHoleDiam;        /* collimator parameters for a computed GRF */
    float fHoleLen;
    float fBackToDet;
    float fIntrinsicFWHM;
} IrlParms_t;

typedef struct {
    int bModelDrf;
    int bReconIsInitEst;
    int bUseGrfInBck;
    int bFFTConvolve;
    int bUseContourSupport;
    float fAtnMapThresh;
    int iMsgLevel;
    int iAxialPadLength;
    int iAxialAvgLength;
} Options_t;

typedef struct {
    float Angle;  /* radians; negative direction for a standard orbit */
    float Left;   /* first measured bin location */
    float Right;  /* last measured bin location */
    float CFCR;   /* centre of rotation to collimator face, cm */
} PrjView_t;

int IrlOsem(IrlParms_t *psIrlParms, Options_t *psOptions, PrjView_t *psViews,
            char *pchDrfTabFname, char *pchScatKrnlFname,
            void (*pIterationCallbackFunc)(int IterationNumber, float *pfCurrentEst),
            float *pfScatterEstimate, float *pfAtnMap, float *pfPrjImage,
            float *pfReconImage, char *pchLogFname, char *pchMessageFname);

int IrlGenprj(IrlParms_t *psIrlParms, Options_t *psOptions, PrjView_t *psPrjViews,
              char *pchDrfTabFname, char *pchScatKrnlFname,
              void (*pPrjViewCallbackFunc)(int PrjViewNumber, float *pfCurrentPrjView),
              float *pfScatterEstimate, float *pfActImage, float *pfAtnMap,
              float *pfPrjImage, float fPrimaryFac, char *pchLogFname,
              char *pchMessageFname);

/* Camera-to-centre distances of an orbit from a 2-D image (irl.h:291-316,
 * computeorbit.c:21-150).  OrbitType 0 sets every CFCR to fMargin; 1 moves
 * each view to the thresholded object's edge plus the margin (pfMargins[v]
 * when given); 2 additionally ties views v and v+NumViews/2 to the larger
 * radius (L-mode).  fThresh > 0 is a fraction of the image maximum, <= 0 an
 * absolute threshold of -fThresh; pixels at or below it are zeroed in
 * pfPixels, which is modified.  Returns 0, 1 (allocation error) or 2
 * (orbit type not 0, 1 or 2). */
int ComputeOrbit(float *pfPixels, IrlParms_t *psParms, float fMargin, float *pfMargins,
                 float fThresh, int OrbitType, PrjView_t *psViews);

char *pchIrlErrorString(void);

#ifdef __cplusplus
}
#endif

#endif /* IRL_H */
