/*
 * irl_amd.h -- device-resident reconstruction API of libirl_amd.
 *
 * No counterpart in the reference: IrlOsem (irl.h) runs setup, the whole
 * iteration loop and teardown in one call.  These entry points split it so a
 * caller can keep one reconstruction resident on the GPU and drive it one
 * OSEM sub-iteration at a time (benchmarks, schedulers, learned-scatter
 * loops).  IrlOsem itself is built on them.  Inputs are the same arrays
 * IrlOsem takes, in the same external order; they are copied, never
 * modified.  Functions return 0 on success; on failure they return nonzero
 * and pchIrlErrorString() holds the message.
 */
#ifndef IRL_AMD_H
#define IRL_AMD_H

#include "irl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct irlamd_recon irlamd_recon;

/* Setup of IrlOsem (irlosem.c:143-236): parameters, views, DRF table, scaling
 * and clamping of copies of the inputs, initial estimate, sensitivity
 * images.  pfReconImage is read only when bReconIsInitEst is set. */
irlamd_recon *irlamd_recon_create(IrlParms_t *psIrlParms, Options_t *psOptions,
                                  PrjView_t *psViews, const char *pchDrfTabFname,
                                  const char *pchScatKrnlFname,
                                  const float *pfScatterEstimate, const float *pfAtnMap,
                                  const float *pfPrjImage, const float *pfReconImage);

/* Number of subsets and views per subset. */
int irlamd_recon_subsets(const irlamd_recon *r, int *piNumSubsets, int *piViewsPerSubset);

/* One sub-iteration (all views of subset iSubset, then the update). */
int irlamd_recon_subiteration(irlamd_recon *r, int iSubset);

/* Waits for queued GPU work. */
int irlamd_recon_synchronize(irlamd_recon *r);

/* Current estimate in external [z][y][x] order. */
int irlamd_recon_get_estimate(irlamd_recon *r, float *pfReconImage);

/* Sensitivity inverse of subset s, internal [y][z][x] order. */
int irlamd_recon_get_norm(irlamd_recon *r, int iSubset, float *pfNorm);

/* Per-kernel HIP-event timing on the reconstruction's stream. */
int irlamd_recon_set_timing(irlamd_recon *r, int bOn);
int irlamd_recon_reset_timing(irlamd_recon *r);
int irlamd_recon_stage_count(irlamd_recon *r);
int irlamd_recon_stage(irlamd_recon *r, int i, char *pchName, int iNameCap,
                       double *pdMsTotal, int *piLaunches);

void irlamd_recon_destroy(irlamd_recon *r);

/* Nonzero when the last call into the library failed. */
int irlamd_last_error(void);

/* Arithmetic of the DRF convolution for reconstructions created afterwards
 * (irlamd_recon_create, IrlOsem, IrlGenprj): 0 = FP32 (the default, the
 * reference's arithmetic), 1 = BF16 MFMA with FP32 accumulation, using hi+lo
 * split BF16 operands when the estimated per-plane relative error exceeds
 * 1e-3, 2 = BF16 hi+lo always, 3 = plain BF16 always.  Modes 1-3 need
 * NumSlices % 4 == 0.  Returns the previous mode, or -1 for an invalid one. */
int irlamd_set_precision(int iMode);

#ifdef __cplusplus
}
#endif

#endif /* IRL_AMD_H */
