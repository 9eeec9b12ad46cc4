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

}  // namespace irlamd
