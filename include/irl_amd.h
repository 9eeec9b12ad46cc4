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

/* With graph replay on (irlamd_set_graphs): captures and instantiates every
 * subset's sub-iteration for the scatter modes of the remaining iterations
 * now, so that later sub-iterations replay from their first occurrence
 * (learned-scatter checks still run eagerly).  Executes no GPU work. */
int irlamd_recon_prepare_graphs(irlamd_recon *r);

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
 * reference's arithmetic), 1 = BF16 MFMA with FP32 accumulation, per DRF
 * table entry the cheapest of plain BF16, hi+lo split weights, and hi+lo
 * split weights and data whose relative error, measured at setup on a probe
 * plane against the FP32 kernel, is at most 1e-3 (irlamd_recon_bf16_planes),
 * 2 = BF16 hi+lo weights and data always, 3 = plain BF16 always.  Modes 1-3 need
 * NumSlices % 4 == 0.  Returns the previous mode, or -1 for an invalid one. */
int irlamd_set_precision(int iMode);

/* BF16 DRF arithmetic of a reconstruction, per DRF table entry (the last one
 * is the centred delta used for rows added unconvolved): piModes[k] 0 =
 * plain BF16, 1 = PSF weights as hi+lo BF16, 2 = weights and data as hi+lo;
 * pfErr[2k], pfErr[2k+1] = the relative L2 errors of plain BF16 and of the
 * weight split measured on a probe plane against the FP32 kernel (mode 1
 * takes the cheapest arithmetic within 1e-3; zeros for modes 2 and 3).
 * Writes at most iCap entries; returns the number of entries (0 on the
 * FP32 path), or -1 on error. */
int irlamd_recon_bf16_planes(irlamd_recon *r, int *piModes, float *pfErr, int iCap);

/* Reconstructions created afterwards replay each sub-iteration from a
 * captured HIP graph from the second occurrence of its subset and scatter
 * mode on (same kernels, arguments and order as the eager launches; subsets
 * with a learned-scatter check and timed runs stay eager).  Returns the
 * previous setting. */
int irlamd_set_graphs(int bOn);

/* Reconstructions created afterwards time the candidate tile shapes of the
 * DRF kernels on their first subset at setup and keep the fastest (results
 * do not depend on the tiles).  Returns the previous setting. */
int irlamd_set_autotune(int bOn);
/* Tile shapes in use (0: the size heuristic): FP32 slice groups per block or
 * BF16 slice tiles per wave, of the projector and the backprojector. */
int irlamd_recon_tiles(irlamd_recon *r, int *piPrj, int *piBck);
/* Sets the DRF tile shapes (0: size heuristic; BF16: 2, 4 or 8 slice tiles
 * per wave, FP32: 4 or 8 slice groups per block); captured sub-iteration
 * graphs are dropped.  Results do not depend on the tiles.  No counterpart
 * in the reference (tuning of this library's kernels). */
int irlamd_recon_set_tiles(irlamd_recon *r, int iPrj, int iBck);
/* Rows per chunk of the spatial projector (each chunk's rows summed into one
 * partial plane, the chunks added in order by the ratio step): NumRotPixs/32,
 * or with the BF16 DRF twice that where the projector's launch at its
 * largest tile still has a block for every CU.  Fixed by the geometry (it
 * groups the float sums of the projection).  No counterpart in the
 * reference (tuning of this library's kernels). */
int irlamd_recon_prj_rows_per_chunk(irlamd_recon *r);

/* Scatter model of reconstructions created afterwards: 1 (the default) =
 * where the grid allows (equal in-plane FFT lengths of 128 or 160, two
 * Taylor terms, no collapse, at least 1024 planes per batch), the slice and
 * bin passes run as 2-D transforms per plane and the effective source is
 * formed in the last pass (depth from the attenuation factors, single-
 * precision logarithm and exponential); 0 = the line passes and the separate
 * source kernel (running depth sum, FP64 exponential, as krnlprj3d.c:903-947).
 * The two agree to rounding.  Returns the previous setting. */
int irlamd_set_esse_fused(int bOn);

/* FP32 DRF convolution algorithm of reconstructions created afterwards:
 * 0 = the reference's choice (the default): the 2-D FFT kernel for tables
 * flagged bFFTconvolve (convolve_ypad.c:62-94, conv2dPsfFFT.c:53-118), the
 * spatial kernel otherwise; 1 = the FFT kernel for every table; 2 = the
 * faster of the two per direction, timed at setup on the first subset;
 * -1 = the spatial kernel always.  The FFT kernel needs slice-fastest
 * planes with NumRotPixs + KX - 1 and NumSlices + KY - 1 at most 160 (the
 * spatial kernel runs otherwise); its results equal the spatial kernel's to
 * FP32 rounding of the transforms.  BF16 modes ignore the setting.
 * Returns the previous mode, or -2 for an invalid one. */
int irlamd_set_drf_fft(int iMode);
/* DRF algorithm in use per direction (0 spatial, 1 FFT) and the FFT length
 * (0: none built). */
int irlamd_recon_drf_algorithms(irlamd_recon *r, int *piPrj, int *piBck, int *piLength);

/* Several reconstructions (any sizes and view counts) stepped together: each
 * step runs the next sub-iteration of every member (subsets in order, the
 * iteration and scatter schedule of each as in IrlOsem).  Members whose
 * sub-iteration can be replayed run from captured graphs: by default
 * (irlamd_batch_set_parent_graph mode 2) as parallel branches of one graph
 * captured with every member's launches, one launch per step; mode 1: the
 * members' graphs as child graphs of one parent graph; mode 0: each member's
 * graph on its own stream (they overlap on the GPU).  The others
 * (first occurrence of a subset and mode, learned-scatter checks) run
 * eagerly on their own streams after them.  The members stay owned by the
 * caller and must outlive the batch; results equal those of stepping each
 * member alone. */
typedef struct irlamd_batch irlamd_batch;
irlamd_batch *irlamd_batch_create(irlamd_recon **pRecons, int iNum);
int irlamd_batch_set_parent_graph(irlamd_batch *b, int iMode);
int irlamd_batch_step(irlamd_batch *b);
int irlamd_batch_synchronize(irlamd_batch *b);
/* One-launch-per-step graph launches, sub-iterations replayed from graphs and
 * sub-iterations run eagerly so far. */
int irlamd_batch_stats(irlamd_batch *b, int *piGraphLaunches, int *piReplays, int *piEager);
void irlamd_batch_destroy(irlamd_batch *b);

/* ---------------------------------------------------------------------------
 * Learned scatter (no counterpart in the reference).  A small 3-D CNN on the
 * half-resolution grid predicts, once per sub-iteration, the two FFT
 * convolution terms of the ESSE effective source (CalcEffScatterSource,
 * krnlprj3d.c:826-948) from the current estimate and the attenuation map;
 * each view's source is then their first-order combination with the view's
 * attenuation depth.  The per-view ESSE stays the fallback and the check:
 * on the first scatter-modelled subset (and every iCheckEvery-th one after
 * it) ESSE also runs for the subset's first view, and a relative L2
 * difference of the two projected sources (primary + scatter, summed over
 * depth before the collimator blur) above fMaxNrmse switches the
 * reconstruction back to ESSE, from that subset on.
 * Handles are independent of reconstructions; the network is shared (kept
 * alive) by every reconstruction created while it is set.
 * ------------------------------------------------------------------------- */
typedef struct irlamd_cnn irlamd_cnn;

/* Trains a network on iNumPairs (x, mu) -> ESSE-term pairs generated on the
 * GPU from seeded phantoms with the scatter kernel file's ESSE model at
 * iNumPixels^2 x iNumSlices voxels of fPixSize cm (the last iNumVal pairs are
 * held out).  pfStats (5 floats, may be NULL): final training loss, held-out
 * relative L2 of the two outputs, seconds, output scale. */
irlamd_cnn *irlamd_cnn_train(const char *pchScatKrnlFname, int iNumPixels, int iNumSlices,
                             float fPixSize, int iNumPairs, int iNumVal, int iEpochs,
                             int iBatch, float fLearnRate, unsigned uSeed, float *pfStats);
/* Untrained network (He-normal weights from uSeed) with output scale fKappa. */
irlamd_cnn *irlamd_cnn_create(float fKappa, unsigned uSeed);
int irlamd_cnn_save(irlamd_cnn *h, const char *pchFname);
irlamd_cnn *irlamd_cnn_load(const char *pchFname);
void irlamd_cnn_free(irlamd_cnn *h);

/* Parameters in packed order: per layer W[27][Cout][Cin] (tap = 9(dy+1) +
 * 3(dx+1) + dz+1) then b[Cout]; layers 2->32, 32->32, 32->32, 32->2. */
int irlamd_cnn_num_params(void);
int irlamd_cnn_get_params(irlamd_cnn *h, float *pfParams, float *pfKappa);
int irlamd_cnn_set_params(irlamd_cnn *h, const float *pfParams);

/* Slices per wave of the inference path's MFMA layers (64, 32 or 16; the
 * output does not depend on it), for launches made or captured afterwards.
 * No counterpart in the reference (tuning of this library's kernels). */
int irlamd_cnn_set_row_length(irlamd_cnn *h, int iRowLength);

/* Network on a host input [Y][X][Z][2] -> output [Y][X][Z][2]; bBf16 selects
 * the inference arithmetic (BF16 MFMA layers), else FP32 throughout. */
int irlamd_cnn_forward(irlamd_cnn *h, const float *pfIn, int iY, int iX, int iZ, int bBf16,
                       float *pfOut);
/* FP32 loss mean((out - target)^2) of one sample and its parameter
 * gradients (packed order). */
int irlamd_cnn_loss_grad(irlamd_cnn *h, const float *pfIn, const float *pfTarget, int iY,
                         int iX, int iZ, double *pdLoss, float *pfGrads);
/* Source volumes g0, g1 ([y][x][z], slices fastest) of an estimate x and an
 * attenuation map mu (per pixel) for a scatter model with fDens =
 * 1/(mu0 pixsize) and fCntrPixFac. */
int irlamd_cnn_sources(irlamd_cnn *h, const float *pfX, const float *pfMu, int iNumPixels,
                       int iNumSlices, float fDens, float fCntrPixFac, float fPrimary,
                       float *pfG0, float *pfG1);

/* Network used by reconstructions created afterwards (NULL: ESSE only). */
int irlamd_set_scatter_cnn(irlamd_cnn *h, int iCheckEvery, float fMaxNrmse);
/* Diagnostic: the same first-order source with the terms from one FFT
 * convolution of the unrotated estimate instead of the network. */
int irlamd_set_scatter_cnn_exact(int bExact);
/* Slices per wave of the scatter network's MFMA layers that a reconstruction
 * uses (its own choice: timed at setup with irlamd_set_autotune(1), else 64;
 * 0 without the learned scatter). */
int irlamd_recon_cnn_row_length(irlamd_recon *r);

/* State of the learned scatter of a reconstruction (or, with r = NULL, of the
 * last IrlOsem call): whether it is still in use, checks run, subsets run on
 * the network, last and worst checked relative L2 of the projected source
 * (each check: the larger of two views, the subset's first and middle one),
 * and the last and worst of its scatter part alone. */
int irlamd_recon_cnn_status(irlamd_recon *r, int *piActive, int *piChecks, int *piUsed,
                            float *pfLastNrmse, float *pfWorstNrmse, float *pfLastScatterNrmse,
                            float *pfWorstScatterNrmse);

#ifdef __cplusplus
}
#endif

#endif /* IRL_AMD_H */
