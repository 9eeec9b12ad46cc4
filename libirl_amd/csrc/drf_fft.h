// drf_fft.h -- DRF convolution by 2-D FFT on gfx950 (HIP translation units
// only): the device form of the reference's Fourier-domain branch
// (convolve_ypad.c:60-94, conv2dPsfFFT.c:53-118).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "drf_dev.h"

namespace irlamd {

// FFT length for NRot x Nz planes and KX x KY PSFs: the smallest supported
// P with P >= NRot + KX - 1 and P >= Nz + KY - 1 (the circular convolution
// then equals the zero-padded linear one on the kept outputs), or 0 when
// none fits one block's LDS.
int drf_fft_size(int NRot, int Nz, int KX, int KY);

// Spectra of one PSF table and the line-transform twiddles (device).
struct DrfFftTables {
    int P = 0, nE = 0;
    float2 *H = nullptr;    // [nE][P/2+1][P] in the kernels' register order
    // PSFs symmetric in both axes have real spectra, even in both
    // frequencies: [nE][P/2+1][P/2+1] real (a quarter of the bytes)
    float *Hr = nullptr;
    float2 *tw = nullptr;   // [P]
    // psf: device [nE][KX][KYP] (transposed, PSFs centred in KX x KY);
    // symmetric: every PSF equals its mirror images in x and in y
    void build(hipStream_t s, int P, const float *psf, int nE, int KX, int KY, int KYP,
               bool symmetric);
    void release();
    DrfFftTables() = default;
    DrfFftTables(const DrfFftTables &) = delete;
    DrfFftTables &operator=(const DrfFftTables &) = delete;
    ~DrfFftTables() { release(); }
};

// The convolutions of k_drf (same arguments, same results to FP32 rounding
// of the transforms): PRJ sums the rows of each chunk (a.rchunk, a.nchunk),
// BCK writes every row's plane.
void launch_drf_fft(hipStream_t s, bool bck, const DrfArgs &a, int B, const DrfFftTables &t);

// The scatter model's slice and bin passes per plane (both lengths P):
// forward act [planes][nm][nf] -> buf [planes][P][P/2+1]; inverse of terms
// 0 and 1 of outk [B][2][ny][P][P/2+1] with the effective source formed from
// them (d from the attenuation factors, as EsseDev::convolveSource).
bool esse_fft2_supported(int Pf, int Pm);
void launch_esse_fwd2d(hipStream_t s, int P, const float *act, float2 *buf, long planes, int nm,
                       int nf, const float2 *tw);
void launch_esse_inv2d_src(hipStream_t s, int P, const float2 *outk, int B, int ny, int nm, int nf,
                           float scale, const float *act, const float *map, const float *afac,
                           const int *bview, float fAtnCnvFac, float fMu, float cntr,
                           float fDensFac, float primary, float *out, const float2 *tw);

}  // namespace irlamd
