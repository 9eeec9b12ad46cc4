// drf_dev.h -- arguments of the DRF convolution kernels and the staging
// of a (padded) input plane (HIP translation units only).
#pragma once

#include <hip/hip_runtime.h>

namespace irlamd {

// Element addressing is by strides, so the kernels serve the engine's
// slice-fastest planes and the projector-level entry points' bin-fastest
// planes:  in (b,row,z,x) = in[b*inB + row*inRow + z*zs + x*xs],
// afac (v,row,z,x) = afac[v*afB + row*afRow + z*zs + x*xs],
// out (b,row|chunk,z,x) = out[b*outB + row*outRow + z*outZS + x*outXS].
struct DrfArgs {
    int NRot, Nz;                 // plane width (bins) and height (slices)
    int nrows;                    // planes (rows) per batch item
    int KX;                       // padded PSF width (odd)
    const float *psf;             // [nDist][KX][KYP] transposed, KYP = bucket padded to 4
    const int *rowKernel;         // [V][NRot]
    const int *rowPad;            // [V][NRot]
    int avg;                      // axial average rows (clamped to [0, Nz])
    const int *bview;             // batch slot -> view
    const float *in;              // PRJ: rotated, attenuated planes; BCK: ratio plane
    long inB, inRow;              // BCK: inRow = 0 (one ratio plane per view)
    const float *afac;            // factors multiplied into the input, or null
    long afB, afRow;
    int zs, xs;                   // in/afac slice and bin strides
    float scale;                  // BCK: output scale (fScaleFac)
    float *out;                   // PRJ: partial sums per chunk; BCK: planes
    long outB, outRow;
    int outZS, outXS;
    int vec4;                     // outZS == 1 and 16-byte aligned rows: float4 stores
    int rchunk, nchunk;           // PRJ row chunking
    int xtiles, ztiles;
    const int2 *colRange;         // per table entry: first and last PSF column holding a
                                  // nonzero tap (null: all KX columns)
    // BF16 kernel: rows taken by this launch (-1 all; 0 rows whose entry is
    // not data-split; 1 the data-split ones), and PRJ partials added to
    // instead of written (the second of two launches over the rows)
    int rowSel;
    int accumulate;
    // BF16 PRJ: the input planes of rows whose entry does not split its data
    // hold bf16 values (the first half of the plane, same indexing), the
    // others FP32
    int inBf16;
    // BF16 PRJ: per row the bins [head, head + npix) of the inscribed
    // circle, outside which every producer of the input planes writes +0
    // (the input is then staged as zeros without being read); null: all bins
    const int *circHead, *circNpix;
};

__device__ __forceinline__ float drf_in(const DrfArgs &a, int b, int v, int row, int g, int gx)
{
    const long off = (long)g * a.zs + (long)gx * a.xs;
    float r = a.in[(long)b * a.inB + (long)row * a.inRow + off];
    if (a.afac) r = r * a.afac[(long)v * a.afB + (long)row * a.afRow + off];
    return r;
}

// Value of padded-plane row g (global slice index, may be outside [0,Nz)) at
// bin gx, following pfPad2d/pfAvgRows/pfCopyRows (pad2d.c:21-141).
__device__ inline float drf_stage_val(const DrfArgs &a, int b, int v, int row, int g, int gx, int pad)
{
    if (gx < 0 || gx >= a.NRot) return 0.0f;
    if (g >= 0 && g < a.Nz) return drf_in(a, b, v, row, g, gx);
    if (pad <= 0 || a.avg <= 0) return 0.0f;
    if (g < -pad || g >= a.Nz + pad) return 0.0f;
    const int base = g < 0 ? 0 : a.Nz - a.avg;
    float s = 0.0f;
    for (int k = 0; k < a.avg; ++k) s += drf_in(a, b, v, row, base + k, gx);
    if (a.avg > 1) s /= (float)a.avg;
    return s;
}

}  // namespace irlamd
