// drf_fft.hip -- the per-plane DRF convolution of k_drf by 2-D FFT on
// gfx950: the device form of the reference's Fourier-domain branch
// (convolve_ypad.c:60-94, conv2dPsfFFT.c:53-118, PSF wrapped to the origin
// as PadSwap does, padswap.c:21-119), on a P x P grid just large enough for
// the circular convolution to equal the zero-padded linear one on the kept
// outputs (the reference pads to 2N).  The axial padding of the spatial
// branch (pad2d.c:21-141) is applied to the input as k_drf does, so the
// results equal k_drf's to FP32 rounding of the transforms.
//
// One block transforms whole planes in LDS.  Forward: the real slice
// lines of two bins at a time are one complex line (real and imaginary
// parts), split into their half spectra (k <= P/2) afterwards; then the
// P/2 + 1 complex lines along the bins.  Thread (line, tt) then holds
// S[kz][kx = tt + T w] in registers.  PRJ multiplies every row's spectrum
// by its PSF spectrum and accumulates in registers, so one inverse serves
// a chunk of rows; BCK multiplies and inverts per row.  Inverse: complex
// lines along the bins, then the half spectra of two bins completed by
// Hermitian symmetry as one complex line, whose real and imaginary outputs
// are the two bins' slices.  Line transforms are fft_line (fft_dev.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "drf_fft.h"
#include "engine.h"
#include "fft_dev.h"

namespace irlamd {

#define DFFT_CHECK(expr)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::irlamd::fail(ETYPE_DEVICE, "HIP error '%s' in %s (%s:%d)",                 \
                           hipGetErrorString(e_), #expr, __FILE__, __LINE__);            \
    } while (0)

template <int P>
struct DrfFftShape {
    static constexpr int T = FftPlan<P>::T, R = FftPlan<P>::R;
    static constexpr int LS = FftLine<P>::LS;          // LDS line stride (float2)
    static constexpr int WC = P / 2 + 1;               // half spectrum along the slices
    static constexpr int NL = (WC * T + 63) / 64 * 64 / T;  // line slots per block
    static constexpr int NT = NL * T;                  // threads per block
    // line slots, then the twiddles
    static constexpr size_t LDS = ((size_t)NL * LS + P) * sizeof(float2);
    static constexpr int KH = (P / 2) / T + 1;         // registers that can hold k <= P/2
    static_assert(NT <= 1024 && R % 2 == 0, "block size, 128-bit spectrum loads");
};

// First line slot of the calling thread's wave (uniform per wave).
template <int P>
__device__ __forceinline__ int drf_fft_wave_first_line()
{
    const int first = __builtin_amdgcn_readfirstlane((int)threadIdx.x);
    return (first / 64) * (64 / DrfFftShape<P>::T);
}

// Axial pad of a bin pair's slice line: FFT index n in [Nz, Nz + pad) holds
// the mean of the last avg slices, n in [P - pad, P) (slices -pad..-1) that
// of the first avg (pad2d.c:21-141, as drf_stage_val).
template <int P>
__device__ __forceinline__ void drf_fft_pad(const DrfArgs &a, int b, int v, int row, int pad,
                                            float2 *s, int c0, int c1, bool zact, int tt)
{
    using D = DrfFftShape<P>;
    if (pad <= 0 || a.avg <= 0 || !zact) return;
    const float lo0 = drf_stage_val(a, b, v, row, -1, c0, pad);
    const float lo1 = drf_stage_val(a, b, v, row, -1, c1, pad);
    const float hi0 = drf_stage_val(a, b, v, row, a.Nz, c0, pad);
    const float hi1 = drf_stage_val(a, b, v, row, a.Nz, c1, pad);
#pragma unroll
    for (int w = 0; w < D::R; ++w) {
        const int n = tt + D::T * w;
        if (n >= a.Nz && n < a.Nz + pad) s[w] = make_float2(hi0, hi1);
        if (n >= P - pad) s[w] = make_float2(lo0, lo1);
    }
}

// One 4-byte load per 128-byte line of plane (b, row) of the input and of
// its factors (nothing for row < 0): pulls the block's next plane toward its
// L2 while the current one is transformed.  The caller folds the returned
// value in only after the row's later loads, so these stay in flight.
__device__ __forceinline__ float drf_fft_touch(const DrfArgs &a, int b, int v, int row)
{
    float r = 0.0f;
    if (row < 0) return r;
    const long span = (long)(a.NRot - 1) * a.xs + a.Nz;
    const float *p = a.in + (long)b * a.inB + (long)row * a.inRow;
    const float *f = a.afac ? a.afac + (long)v * a.afB + (long)row * a.afRow : nullptr;
    // projector: only the bins of the row's circle are read
    long e0 = 0, e1 = span;
    if (a.circHead) {
        e0 = (long)a.circHead[row] * a.xs;
        e1 = min(span, (long)(a.circHead[row] + a.circNpix[row]) * a.xs);
    }
    for (long e = e0 + (long)threadIdx.x * 32; e < e1; e += (long)blockDim.x * 32) {
        r += p[e];
        if (f) r += f[e];
    }
    return r;
}

template <int P>
__device__ __forceinline__ void fft2_forward_rest(float2 *s, float2 *sm, int li, int tt,
                                                  const float2 *__restrict__ tw, bool zIdle,
                                                  int nbins);

// Forward 2-D transform of plane (b, row) of the DRF input (with axial pad
// 'pad').  On exit thread (line li, tt), li < P/2 + 1, holds
// S[kz = li][kx = tt + T w] in s[w].  Every thread of the block calls it.
// After the plane's loads it issues drf_fft_touch of row 'next' into
// 'pending'.
template <int P>
__device__ __forceinline__ void drf_fft_forward(const DrfArgs &a, int b, int v, int row, int pad,
                                                float2 *s, float2 *sm, int li, int tt,
                                                const float2 *__restrict__ tw, int next,
                                                float &pending)
{
    using D = DrfFftShape<P>;
    float2 *slot = sm + li * D::LS;
    const int c0 = 2 * li, c1 = c0 + 1;
    const bool zact = c0 < a.NRot, has1 = c1 < a.NRot;
    // projector: the input is +0 outside the row's circle, so bins there are
    // not read (their values are the zeros the selects below give)
    bool rd0 = zact, rd1 = has1;
    if (a.circHead) {
        const int h = a.circHead[row], e = h + a.circNpix[row];
        rd0 = rd0 && c0 >= h && c0 < e;
        rd1 = rd1 && c1 >= h && c1 < e;
    }
    // waves beyond the last bin pair skip the slice transform (the condition
    // is uniform per wave; the line transforms synchronise within a wave)
    const bool zIdle = drf_fft_wave_first_line<P>() >= (a.NRot + 1) / 2;
    // slice lines of bins c0 (real part) and c1 (imaginary part), slices
    // contiguous (zs == 1): FFT index n < Nz holds slice n
    if (zIdle) {
#pragma unroll
        for (int w = 0; w < D::R; ++w) s[w] = make_float2(0.0f, 0.0f);
    } else {
        // loads at clamped positions (always inside the plane), then selects
        const int q0 = min(c0, a.NRot - 1), q1 = min(c1, a.NRot - 1);
        const float *base = a.in + (long)b * a.inB + (long)row * a.inRow;
        const float *i0 = base + (long)q0 * a.xs, *i1 = base + (long)q1 * a.xs;
        const float *fb = a.afac ? a.afac + (long)v * a.afB + (long)row * a.afRow : nullptr;
        const float *g0 = fb ? fb + (long)q0 * a.xs : nullptr, *g1 = fb ? fb + (long)q1 * a.xs : nullptr;
#pragma unroll
        for (int w = 0; w < D::R; ++w) {
            // uniform per register: wholly inside, straddling or beyond the slices
            const int n0 = D::T * w;
            if (n0 >= a.Nz) {
                s[w] = make_float2(0.0f, 0.0f);
                continue;
            }
            const bool whole = n0 + D::T <= a.Nz;
            const int n = tt + n0;
            const int nc = whole ? n : min(n, a.Nz - 1);
            float re = 0.0f, im = 0.0f;
            if (rd0) re = i0[nc];
            if (rd1) im = i1[nc];
            if (fb) {
                re *= g0[nc];
                im *= g1[nc];
            }
            const bool in = whole || n < a.Nz;
            s[w] = make_float2((rd0 && in) ? re : 0.0f, (rd1 && in) ? im : 0.0f);
        }
        drf_fft_pad<P>(a, b, v, row, pad, s, c0, c1, zact, tt);
    }
    __builtin_amdgcn_sched_barrier(0);   // the next plane's loads after this one's
    pending = drf_fft_touch(a, b, v, next);
    fft2_forward_rest<P>(s, sm, li, tt, tw, zIdle, a.NRot);
}

// The forward 2-D transform from the loaded slice lines on: thread (line
// li, tt) holds slices tt + T w of bins 2 li (real part) and 2 li + 1
// (imaginary part) in s[w], zero where there is no such bin or slice (zIdle:
// the wave has no bin pair).  On exit it holds S[kz = li][kx = tt + T w].
template <int P>
__device__ __forceinline__ void fft2_forward_rest(float2 *s, float2 *sm, int li, int tt,
                                                  const float2 *__restrict__ tw, bool zIdle,
                                                  int nbins)
{
    using D = DrfFftShape<P>;
    float2 *slot = sm + li * D::LS;
    const int c0 = 2 * li, c1 = c0 + 1;
    const bool zact = c0 < nbins;
    if (!zIdle) fft_line<P, false, true>(s, slot, tt, tw);
    // Between phases the lane indices pass through an empty asm, so each
    // phase computes its own LDS addresses instead of the compiler keeping
    // those of every phase live at once.
    asm volatile("" : "+v"(tt), "+v"(li));
    slot = sm + li * D::LS;
    // Z = A + iB: A[k] = (Z[k] + conj Z[P-k]) / 2, B[k] = (Z[k] - conj Z[P-k]) / 2i.
    // Z[P - k], k = tt + T w, is in lane T - tt of the line, register R-1-w
    // (tt > 0), or in this thread, register R - w (tt = 0): a lane shuffle
    // within the wave, no LDS.
    const int lane = (int)(threadIdx.x & 63);
    const int src = (lane & ~(D::T - 1)) | ((D::T - tt) & (D::T - 1));
    float2 A[D::KH], Bq[D::KH];
#pragma unroll
    for (int w = 0; w < D::KH; ++w) {
        const float2 zk = s[w];
        float2 zm = make_float2(__shfl(s[D::R - 1 - w].x, src, 64), __shfl(s[D::R - 1 - w].y, src, 64));
        if (tt == 0) zm = s[w == 0 ? 0 : D::R - w];
        A[w] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        Bq[w] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    }
    // every wave's slice transforms are done with their line slots
    __syncthreads();
    if (zact) {
#pragma unroll
        for (int w = 0; w < D::KH; ++w) {
            const int k = tt + D::T * w;
            if (k <= P / 2) {
                sm[k * D::LS + fft_skew(c0)] = A[w];
                if (c1 < nbins) sm[k * D::LS + fft_skew(c1)] = Bq[w];
            }
        }
    }
    __syncthreads();
    asm volatile("" : "+v"(tt), "+v"(li));
    slot = sm + li * D::LS;
    // bin lines: line kz = li, bins >= NRot are zero
    const bool xact = li < D::WC;
#pragma unroll
    for (int w = 0; w < D::R; ++w) {
        const int x = tt + D::T * w;
        const float2 val = slot[fft_skew(x)];   // inside the slot for every x < P
        s[w] = (xact && x < nbins) ? val : make_float2(0.0f, 0.0f);
    }
    fft_line<P, false, true>(s, slot, tt, tw);
}

// Inverse 2-D transform of the half spectrum held as drf_fft_forward leaves
// it (already scaled by 1/P^2).  On exit thread (line li, tt) holds slice
// n = tt + T w of bin 2 li in g[w].x and of bin 2 li + 1 in g[w].y.
template <int P>
__device__ __forceinline__ void fft2_inverse(int nbins, float2 *g, float2 *sm, int li, int tt,
                                             const float2 *__restrict__ tw);

template <int P>
__device__ __forceinline__ void drf_fft_inverse(const DrfArgs &a, float2 *g, float2 *sm, int li,
                                                int tt, const float2 *__restrict__ tw)
{
    fft2_inverse<P>(a.NRot, g, sm, li, tt, tw);
}

// The inverse transform for planes of nbins bins (see drf_fft_inverse).
template <int P>
__device__ __forceinline__ void fft2_inverse(int nbins, float2 *g, float2 *sm, int li, int tt,
                                             const float2 *__restrict__ tw)
{
    using D = DrfFftShape<P>;
    float2 *slot = sm + li * D::LS;
    fft_line<P, true, true>(g, slot, tt, tw);
    asm volatile("" : "+v"(tt), "+v"(li));
    slot = sm + li * D::LS;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < D::R; ++w) {
        const int x = tt + D::T * w;
        if (x < nbins) slot[fft_skew(x)] = g[w];
    }
    __syncthreads();
    asm volatile("" : "+v"(tt), "+v"(li));
    slot = sm + li * D::LS;
    const int c0 = 2 * li, c1 = c0 + 1;
    const bool has0 = c0 < nbins, has1 = c1 < nbins;
    // waves with no bin pair skip the slice transform; every wave meets the
    // block barrier below (the line transform synchronises within a wave)
    const bool idle = drf_fft_wave_first_line<P>() >= (nbins + 1) / 2;
    if (idle) {
#pragma unroll
        for (int w = 0; w < D::R; ++w) g[w] = make_float2(0.0f, 0.0f);
    } else {
        // reads at clamped bins (inside line slot kk), then selects
        const int s0 = fft_skew(min(c0, P - 1)), s1 = fft_skew(min(c1, P - 1));
#pragma unroll
        for (int w = 0; w < D::R; ++w) {
            const int n = tt + D::T * w;
            const bool upper = n > P / 2;
            const int kk = upper ? P - n : n;
            const float2 A0 = sm[kk * D::LS + s0], B0 = sm[kk * D::LS + s1];
            const float2 A = has0 ? A0 : make_float2(0.0f, 0.0f);
            const float2 B = has1 ? B0 : make_float2(0.0f, 0.0f);
            const float sg = upper ? -1.0f : 1.0f;   // conj of the mirrored half
            g[w] = make_float2(A.x - sg * B.y, sg * A.y + B.x);
        }
    }
    __syncthreads();   // every read of the line slots is done
    if (!idle) fft_line<P, true, true>(g, slot, tt, tw);
}

// PRJ: block = (batch item b, chunk ch) sums rows [ch rchunk, ...) of
// DRF_row (*) in[b][row] and writes the chunk's partial projection.
template <int P>
__global__ __launch_bounds__(DrfFftShape<P>::NT) void k_drf_fft_prj(DrfArgs a,
                                                                    const float2 *__restrict__ H,
                                                                    const float *__restrict__ Hr,
                                                                    const float2 *__restrict__ twg)
{
    using D = DrfFftShape<P>;
    extern __shared__ float2 sm[];
    const int li = threadIdx.x / D::T, tt = threadIdx.x % D::T;
    float2 *tw = sm + D::NL * D::LS;
    for (int i = threadIdx.x; i < P; i += D::NT) tw[i] = twg[i];
    __syncthreads();
    const int b = blockIdx.x / a.nchunk, ch = blockIdx.x % a.nchunk;
    const int r0 = ch * a.rchunk, r1 = min(r0 + a.rchunk, a.nrows);
    const int v = a.bview[b];
    const bool xact = li < D::WC;
    const float unit = 1.0f / ((float)P * (float)P);
    float2 acc[D::R], s[D::R];
    float keep = 0.0f;
#pragma unroll
    for (int w = 0; w < D::R; ++w) acc[w] = make_float2(0.0f, 0.0f);
    for (int row = r0; row < r1; ++row) {
        const int kidx = a.rowKernel[(long)v * a.nrows + row];
        const int pad = (kidx < 0 || a.avg <= 0) ? 0 : a.rowPad[(long)v * a.nrows + row];
        // the lane indices pass through an empty asm each row, so the
        // compiler recomputes the transforms' LDS addresses instead of
        // keeping hundreds of them live across the loop
        int lq = li, tq = tt;
        asm volatile("" : "+v"(lq), "+v"(tq));
        float pending;
        drf_fft_forward<P>(a, b, v, row, pad, s, sm, lq, tq, tw, row + 1 < r1 ? row + 1 : -1,
                           pending);
        if (xact) {
            if (kidx >= 0 && Hr) {
                // real spectrum, even in kx: value at min(kx, P - kx)
                const float *h = Hr + ((long)kidx * D::WC + li) * D::WC;
#pragma unroll
                for (int w = 0; w < D::R; ++w) {
                    const int kx = tt + D::T * w;
                    const float hv = h[kx <= P / 2 ? kx : P - kx];
                    acc[w].x = fmaf(s[w].x, hv, acc[w].x);
                    acc[w].y = fmaf(s[w].y, hv, acc[w].y);
                }
            } else if (kidx >= 0) {
                // this thread's R values are contiguous: R/2 128-bit loads
                const float4 *h = (const float4 *)(H + ((long)kidx * D::WC + li) * P + tt * D::R);
#pragma unroll
                for (int w2 = 0; w2 < D::R / 2; ++w2) {
                    const float4 hq = h[w2];
                    const float2 hv[2] = {make_float2(hq.x, hq.y), make_float2(hq.z, hq.w)};
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int w = 2 * w2 + e;
                        acc[w].x = fmaf(s[w].x, hv[e].x, fmaf(-s[w].y, hv[e].y, acc[w].x));
                        acc[w].y = fmaf(s[w].x, hv[e].y, fmaf(s[w].y, hv[e].x, acc[w].y));
                    }
                }
            } else {
                // distance <= 0: added unconvolved
#pragma unroll
                for (int w = 0; w < D::R; ++w) {
                    acc[w].x = fmaf(s[w].x, unit, acc[w].x);
                    acc[w].y = fmaf(s[w].y, unit, acc[w].y);
                }
            }
        }
        keep += pending;
        __syncthreads();   // the next row rewrites the line slots
    }
    if (a.nrows < 0) a.out[0] = keep;   // never: keeps the touch loads
    drf_fft_inverse<P>(a, acc, sm, li, tt, tw);
    const int c0 = 2 * li, c1 = c0 + 1;
    float *o = a.out + (long)b * a.outB + (long)ch * a.outRow;
#pragma unroll
    for (int w = 0; w < D::R; ++w) {
        const int n = tt + D::T * w;
        if (n < a.Nz) {
            if (c0 < a.NRot) o[(long)c0 * a.outXS + n] = acc[w].x;
            if (c1 < a.NRot) o[(long)c1 * a.outXS + n] = acc[w].y;
        }
    }
}

// BCK: block = (batch item b, rows [r0, r0 + rows)): planes[row] =
// scale * DRF_row (*) in[b][row].  The input spectrum is reused across the
// block's rows while the input does not depend on the row (no factors)
// and the axial pad stays the same.
template <int P>
__global__ __launch_bounds__(DrfFftShape<P>::NT) void k_drf_fft_bck(DrfArgs a,
                                                                    const float2 *__restrict__ H,
                                                                    const float *__restrict__ Hr,
                                                                    const float2 *__restrict__ twg,
                                                                    int rows, int nblk)
{
    using D = DrfFftShape<P>;
    extern __shared__ float2 sm[];
    const int li = threadIdx.x / D::T, tt = threadIdx.x % D::T;
    float2 *tw = sm + D::NL * D::LS;
    for (int i = threadIdx.x; i < P; i += D::NT) tw[i] = twg[i];
    __syncthreads();
    const int b = blockIdx.x / nblk;
    const int r0 = (blockIdx.x % nblk) * rows, r1 = min(r0 + rows, a.nrows);
    const int v = a.bview[b];
    const bool xact = li < D::WC;
    const bool reuse = a.afac == nullptr && a.inRow == 0;
    const float unit = 1.0f / ((float)P * (float)P);
    const int c0 = 2 * li, c1 = c0 + 1;
    float2 spec[D::R], g[D::R];
    float keep = 0.0f;
    int have = -1;
    for (int row = r0; row < r1; ++row) {
        const int kidx = a.rowKernel[(long)v * a.nrows + row];
        const int pad = (kidx < 0 || a.avg <= 0) ? 0 : a.rowPad[(long)v * a.nrows + row];
        int lq = li, tq = tt;
        asm volatile("" : "+v"(lq), "+v"(tq));
        float pending = 0.0f;
        if (!reuse || pad != have) {
            drf_fft_forward<P>(a, b, v, row, pad, spec, sm, lq, tq, tw,
                               (!reuse && row + 1 < r1) ? row + 1 : -1, pending);
            have = pad;
            __syncthreads();   // the inverse rewrites the line slots
        }
        if (xact) {
            if (kidx >= 0 && Hr) {
                const float *h = Hr + ((long)kidx * D::WC + li) * D::WC;
#pragma unroll
                for (int w = 0; w < D::R; ++w) {
                    const int kx = tt + D::T * w;
                    const float hv = h[kx <= P / 2 ? kx : P - kx];
                    g[w] = make_float2(spec[w].x * hv, spec[w].y * hv);
                }
            } else if (kidx >= 0) {
                const float4 *h = (const float4 *)(H + ((long)kidx * D::WC + li) * P + tt * D::R);
#pragma unroll
                for (int w2 = 0; w2 < D::R / 2; ++w2) {
                    const float4 hq = h[w2];
                    const float2 hv[2] = {make_float2(hq.x, hq.y), make_float2(hq.z, hq.w)};
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int w = 2 * w2 + e;
                        g[w].x = fmaf(spec[w].x, hv[e].x, -spec[w].y * hv[e].y);
                        g[w].y = fmaf(spec[w].x, hv[e].y, spec[w].y * hv[e].x);
                    }
                }
            } else {
#pragma unroll
                for (int w = 0; w < D::R; ++w) g[w] = make_float2(spec[w].x * unit, spec[w].y * unit);
            }
        } else {
#pragma unroll
            for (int w = 0; w < D::R; ++w) g[w] = make_float2(0.0f, 0.0f);
        }
        drf_fft_inverse<P>(a, g, sm, lq, tq, tw);
        float *o = a.out + (long)b * a.outB + (long)row * a.outRow;
#pragma unroll
        for (int w = 0; w < D::R; ++w) {
            const int n = tt + D::T * w;
            if (n < a.Nz) {
                if (c0 < a.NRot) o[(long)c0 * a.outXS + n] = g[w].x * a.scale;
                if (c1 < a.NRot) o[(long)c1 * a.outXS + n] = g[w].y * a.scale;
            }
        }
        keep += pending;
        __syncthreads();   // the next row rewrites the line slots
    }
    if (a.nrows < 0) a.out[0] = keep;   // never: keeps the touch loads
}

// H[e][kz][tt][w] = (1/P^2) sum_{i,j} psf_e[j][i] W^(kz (j - CY) + kx (i - CX)),
// kx = tt + T w, W = exp(-2 pi i / P): the transform of the PSF with its
// centre at the origin, scaled for the unnormalised inverse.  FP64 sums;
// root[m] = W^m in double.
template <int P>
__global__ void k_drf_spectra(const float *__restrict__ psf, int nE, int KX, int KY, int KYP,
                              const double2 *__restrict__ root, float2 *__restrict__ H)
{
    using D = DrfFftShape<P>;
    const long per = (long)P * D::WC;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nE * per) return;
    const int e = (int)(id / per);
    const int rem = (int)(id % per);
    const int w = rem % D::R, tt = (rem / D::R) % D::T, kz = rem / P;
    const int kx = tt + D::T * w;
    const int CX = (KX - 1) / 2, CY = (KY - 1) / 2;
    const float *p = psf + (long)e * KX * KYP;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < KX; ++i) {
        double cr = 0.0, ci = 0.0;
        for (int j = 0; j < KY; ++j) {
            const double val = p[i * KYP + j];
            if (val == 0.0) continue;
            const double2 r = root[((kz * (j - CY)) % P + P) % P];
            cr += val * r.x;
            ci += val * r.y;
        }
        const double2 r = root[((kx * (i - CX)) % P + P) % P];
        re += cr * r.x - ci * r.y;
        im += cr * r.y + ci * r.x;
    }
    const double sc = 1.0 / ((double)P * P);
    H[id] = make_float2((float)(re * sc), (float)(im * sc));
}

// Hr[e][kz][kx] = (1/P^2) sum_{i,j} psf_e[j][i] cos(2 pi kz (j - CY)/P)
// cos(2 pi kx (i - CX)/P), kz, kx <= P/2: the spectrum above for a PSF
// symmetric in both axes (the sine terms cancel in pairs).
template <int P>
__global__ void k_drf_spectra_real(const float *__restrict__ psf, int nE, int KX, int KY,
                                   int KYP, const double2 *__restrict__ root,
                                   float *__restrict__ Hr)
{
    constexpr int WC = P / 2 + 1;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nE * WC * WC) return;
    const int e = (int)(id / (WC * WC));
    const int kz = (int)(id / WC % WC), kx = (int)(id % WC);
    const int CX = (KX - 1) / 2, CY = (KY - 1) / 2;
    const float *p = psf + (long)e * KX * KYP;
    double acc = 0.0;
    for (int i = 0; i < KX; ++i) {
        double c = 0.0;
        for (int j = 0; j < KY; ++j) {
            const double val = p[i * KYP + j];
            if (val != 0.0) c += val * root[((kz * (j - CY)) % P + P) % P].x;
        }
        acc += c * root[((kx * (i - CX)) % P + P) % P].x;
    }
    Hr[id] = (float)(acc / ((double)P * P));
}

// supported transform lengths, ascending
static const int kDrfFftSizes[] = {32, 64, 128, 160};

int drf_fft_size(int NRot, int Nz, int KX, int KY)
{
    const int need = std::max(NRot + KX - 1, Nz + KY - 1);
    for (int P : kDrfFftSizes)
        if (P >= need) return P;
    return 0;
}

void DrfFftTables::release()
{
    if (H) (void)hipFree(H);
    if (Hr) (void)hipFree(Hr);
    if (tw) (void)hipFree(tw);
    H = tw = nullptr;
    Hr = nullptr;
    P = nE = 0;
}

void DrfFftTables::build(hipStream_t s, int P_, const float *psf, int nE_, int KX, int KY, int KYP,
                         bool symmetric)
{
    release();
    if (nE_ <= 0) fail(ETYPE_ILLEGAL_VALUE, "FFT DRF: empty PSF table");
    const long per = (long)P_ * (P_ / 2 + 1);
    std::vector<double2> hroot(P_);
    for (int m = 0; m < P_; ++m) {
        const double ang = -2.0 * M_PI * m / P_;
        hroot[m] = make_double2(cos(ang), sin(ang));
    }
    const std::vector<float2> htw = twiddles(P_);
    double2 *root = nullptr;
    DFFT_CHECK(hipMalloc((void **)&root, P_ * sizeof(double2)));
    const long wc = P_ / 2 + 1;
    if (symmetric)
        DFFT_CHECK(hipMalloc((void **)&Hr, (size_t)nE_ * wc * wc * sizeof(float)));
    else
        DFFT_CHECK(hipMalloc((void **)&H, (size_t)nE_ * per * sizeof(float2)));
    DFFT_CHECK(hipMalloc((void **)&tw, P_ * sizeof(float2)));
    P = P_;
    nE = nE_;
    DFFT_CHECK(hipMemcpyAsync(root, hroot.data(), P_ * sizeof(double2), hipMemcpyHostToDevice, s));
    DFFT_CHECK(hipMemcpyAsync(tw, htw.data(), P_ * sizeof(float2), hipMemcpyHostToDevice, s));
    const long n = symmetric ? (long)nE_ * wc * wc : (long)nE_ * per;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
#define DFFT_SPECTRA(PP)                                                                       \
    case PP:                                                                                   \
        if (symmetric)                                                                         \
            hipLaunchKernelGGL(k_drf_spectra_real<PP>, grid, block, 0, s, psf, nE_, KX, KY,    \
                               KYP, root, Hr);                                                 \
        else                                                                                   \
            hipLaunchKernelGGL(k_drf_spectra<PP>, grid, block, 0, s, psf, nE_, KX, KY, KYP,    \
                               root, H);                                                       \
        break;
    switch (P_) {
        DFFT_SPECTRA(32)
        DFFT_SPECTRA(64)
        DFFT_SPECTRA(128)
        DFFT_SPECTRA(160)
    default: fail(ETYPE_ILLEGAL_VALUE, "FFT DRF: unsupported length %d", P_);
    }
#undef DFFT_SPECTRA
    DFFT_CHECK(hipGetLastError());
    DFFT_CHECK(hipStreamSynchronize(s));
    DFFT_CHECK(hipFree(root));
}

template <int P>
static void launch_p(hipStream_t s, bool bck, const DrfArgs &a, int B, const DrfFftTables &t)
{
    using D = DrfFftShape<P>;
    // every line pair, every half-spectrum line and every kept output fit the block
    if (a.NRot + 1 > 2 * D::WC || a.NRot > P || a.Nz > P)
        fail(ETYPE_ILLEGAL_VALUE, "FFT DRF: %d x %d planes do not fit length %d", a.NRot, a.Nz, P);
    if (a.zs != 1 || a.outZS != 1)
        fail(ETYPE_ILLEGAL_VALUE, "FFT DRF: slice-fastest planes only");
    if (bck) {
        // 8 rows per block: without per-row factors one input spectrum
        // serves them all, with factors the next row's plane is prefetched
        const int rows = std::min(8, a.nrows);
        const int nblk = (a.nrows + rows - 1) / rows;
        const long nb = (long)B * nblk;
        if (nb <= 0 || nb > INT_MAX) fail(ETYPE_ILLEGAL_VALUE, "bad FFT DRF launch size");
        hipLaunchKernelGGL(k_drf_fft_bck<P>, dim3((unsigned)nb), dim3(D::NT), D::LDS, s, a, t.H,
                           t.Hr, t.tw, rows, nblk);
    } else {
        const long nb = (long)B * a.nchunk;
        if (nb <= 0 || nb > INT_MAX || a.rchunk <= 0)
            fail(ETYPE_ILLEGAL_VALUE, "bad FFT DRF launch size");
        hipLaunchKernelGGL(k_drf_fft_prj<P>, dim3((unsigned)nb), dim3(D::NT), D::LDS, s, a, t.H,
                           t.Hr, t.tw);
    }
    DFFT_CHECK(hipGetLastError());
}

void launch_drf_fft(hipStream_t s, bool bck, const DrfArgs &a, int B, const DrfFftTables &t)
{
    if ((!t.H && !t.Hr) || !t.tw) fail(ETYPE_ILLEGAL_VALUE, "FFT DRF: tables not built");
    switch (t.P) {
    case 32: launch_p<32>(s, bck, a, B, t); break;
    case 64: launch_p<64>(s, bck, a, B, t); break;
    case 128: launch_p<128>(s, bck, a, B, t); break;
    case 160: launch_p<160>(s, bck, a, B, t); break;
    default: fail(ETYPE_ILLEGAL_VALUE, "FFT DRF: unsupported length %d", t.P);
    }
}

}  // namespace irlamd

namespace irlamd {

// ---------------------------------------------------------------------------
// Scatter model (esse.hip): the slice and bin passes of its 3-D FFT per
// (batch item, depth) plane in LDS with the machinery above, where both
// in-plane lengths are P.  The depth pass between them stays in esse.hip.
// ---------------------------------------------------------------------------

// Forward: act [planes][nm][nf] real (slices fastest) -> buf
// [planes][P][P/2+1] (bin frequency, slice frequency).  A block takes rpb
// consecutive planes and pulls the next one toward its L2 while transforming.
template <int P>
__global__ __launch_bounds__(DrfFftShape<P>::NT) void k_esse_fwd2d(
    const float *__restrict__ act, float2 *__restrict__ buf, long planes, int nm, int nf, int rpb,
    const float2 *__restrict__ twg)
{
    using D = DrfFftShape<P>;
    extern __shared__ float2 sm[];
    const int li = threadIdx.x / D::T, tt = threadIdx.x % D::T;
    float2 *tw = sm + D::NL * D::LS;
    for (int i = threadIdx.x; i < P; i += D::NT) tw[i] = twg[i];
    __syncthreads();
    const long p0 = (long)blockIdx.x * rpb, p1 = min(p0 + rpb, planes);
    const long pl = (long)nm * nf;
    const bool zIdle = drf_fft_wave_first_line<P>() >= (nm + 1) / 2;
    float keep = 0.0f;
    for (long p = p0; p < p1; ++p) {
        int lq = li, tq = tt;
        asm volatile("" : "+v"(lq), "+v"(tq));
        float2 s[D::R];
        if (zIdle) {
#pragma unroll
            for (int w = 0; w < D::R; ++w) s[w] = make_float2(0.0f, 0.0f);
        } else {
            const int c0 = 2 * lq, c1 = c0 + 1;
            const bool zact = c0 < nm, has1 = c1 < nm;
            const float *i0 = act + p * pl + (long)min(c0, nm - 1) * nf;
            const float *i1 = act + p * pl + (long)min(c1, nm - 1) * nf;
#pragma unroll
            for (int w = 0; w < D::R; ++w) {
                const int n0 = D::T * w;
                if (n0 >= nf) {
                    s[w] = make_float2(0.0f, 0.0f);
                    continue;
                }
                const bool whole = n0 + D::T <= nf;
                const int n = tq + n0;
                const int nc = whole ? n : min(n, nf - 1);
                const float re = i0[nc], im = i1[nc];
                const bool in = whole || n < nf;
                s[w] = make_float2((zact && in) ? re : 0.0f, (has1 && in) ? im : 0.0f);
            }
        }
        __builtin_amdgcn_sched_barrier(0);   // the next plane's loads after this one's
        float pending = 0.0f;
        if (p + 1 < p1)
            for (long e = (long)threadIdx.x * 32; e < pl; e += (long)D::NT * 32)
                pending += act[(p + 1) * pl + e];
        fft2_forward_rest<P>(s, sm, lq, tq, tw, zIdle, nm);
        if (lq < D::WC) {
            float2 *o = buf + p * (long)P * D::WC + lq;
#pragma unroll
            for (int w = 0; w < D::R; ++w) o[(long)(tq + D::T * w) * D::WC] = s[w];
        }
        keep += pending;
        __syncthreads();   // the next plane rewrites the line slots
    }
    if (planes < 0) buf[0].x = keep;   // never: keeps the touch loads
}

// Inverse of terms 0 and 1 of every (batch, depth) plane of outk
// [B][2][ny][P][P/2+1] and the effective source on the plane's voxels (the
// epilogue of k_fft_f_inv_src: eff = (E0 - d E1) exp(d fMu) (+ cntr act) *
// map fDens, clamped at 0, out = (act primary + eff) afac, d = -ln(afac) /
// mu0, single-precision logarithm and exponential), out [B][ny][nm][nf].
// act, map and afac as in k_fft_f_inv_src.
template <int P>
__global__ __launch_bounds__(DrfFftShape<P>::NT) void k_esse_inv2d_src(
    const float2 *__restrict__ outk, int B, int ny, int nm, int nf, int rpb, float scale,
    const float *__restrict__ act, const float *__restrict__ map, const float *__restrict__ afac,
    const int *__restrict__ bview, float fAtnCnvFac, float fMu, float cntr, float fDensFac,
    float primary, float *__restrict__ out, const float2 *__restrict__ twg)
{
    using D = DrfFftShape<P>;
    extern __shared__ float2 sm[];
    const int li = threadIdx.x / D::T, tt = threadIdx.x % D::T;
    float2 *tw = sm + D::NL * D::LS;
    for (int i = threadIdx.x; i < P; i += D::NT) tw[i] = twg[i];
    __syncthreads();
    const long planes = (long)B * ny;
    const long p0 = (long)blockIdx.x * rpb, p1 = min(p0 + rpb, planes);
    const long pl = (long)nm * nf, vol = (long)ny * pl;
    const long spec = (long)P * D::WC;   // one term's plane
    for (long p = p0; p < p1; ++p) {
        const int b = (int)(p / ny), y = (int)(p % ny);
        float2 e0[D::R], g[D::R];
#pragma unroll 1
        for (int t = 0; t < 2; ++t) {
            int lq = li, tq = tt;
            asm volatile("" : "+v"(lq), "+v"(tq));
            const float2 *src = outk + (((long)b * 2 + t) * ny + y) * spec + lq;
            if (lq < D::WC) {
#pragma unroll
                for (int w = 0; w < D::R; ++w) g[w] = src[(long)(tq + D::T * w) * D::WC];
            } else {
#pragma unroll
                for (int w = 0; w < D::R; ++w) g[w] = make_float2(0.0f, 0.0f);
            }
            fft2_inverse<P>(nm, g, sm, lq, tq, tw);
            if (t == 0) {
#pragma unroll
                for (int w = 0; w < D::R; ++w) e0[w] = g[w];
            }
        }
        // voxels (b, y, bin 2 li (.x) and 2 li + 1 (.y), slice tt + T w)
        const int c0 = 2 * li;
        const long vo = (long)bview[b] * vol + (long)y * pl, bo = p * pl;
#pragma unroll
        for (int w = 0; w < D::R; ++w) {
            const int n = tt + D::T * w;
            if (n >= nf) continue;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = c0 + h;
                if (c >= nm) continue;
                const long o = (long)c * nf + n;
                const float e0v = (h ? e0[w].y : e0[w].x) * scale;
                const float e1v = (h ? g[w].y : g[w].x) * scale;
                const float f = afac[vo + o], mp = map[vo + o], av = act[bo + o];
                // f == 0 (depth past float range): d = 0 keeps eff finite, out is 0
                const float d = f > 0.0f ? -logf(f) * fAtnCnvFac : 0.0f;
                float eff = e0v - d * e1v;
                eff = eff * expf(d * fMu);
                if (cntr > 0) eff = eff + av * cntr;
                eff = eff * (mp * fDensFac);
                if (eff < 0) eff = 0.0f;
                out[bo + o] = (av * primary + eff) * f;
            }
        }
        __syncthreads();   // the next plane rewrites the line slots
    }
}

bool esse_fft2_supported(int Pf, int Pm) { return Pf == Pm && (Pf == 128 || Pf == 160); }

// planes per block: enough that the blocks fit one pass over the 256 CUs
// (one block per CU), at most 8
static int esse_planes_per_block(long planes)
{
    return (int)std::max(1L, std::min(8L, (planes + 255) / 256));
}

void launch_esse_fwd2d(hipStream_t s, int P, const float *act, float2 *buf, long planes, int nm,
                       int nf, const float2 *tw)
{
    const int rpb = esse_planes_per_block(planes);
    const long nb = (planes + rpb - 1) / rpb;
    if (nb <= 0 || nb > INT_MAX || nm > P || nf > P) fail(ETYPE_ILLEGAL_VALUE, "bad ESSE 2-D FFT size");
#define EFWD(PP)                                                                                \
    hipLaunchKernelGGL(k_esse_fwd2d<PP>, dim3((unsigned)nb), dim3(DrfFftShape<PP>::NT),         \
                       DrfFftShape<PP>::LDS, s, act, buf, planes, nm, nf, rpb, tw)
    if (P == 128)
        EFWD(128);
    else if (P == 160)
        EFWD(160);
    else
        fail(ETYPE_ILLEGAL_VALUE, "ESSE 2-D FFT length %d", P);
#undef EFWD
    DFFT_CHECK(hipGetLastError());
}

void launch_esse_inv2d_src(hipStream_t s, int P, const float2 *outk, int B, int ny, int nm, int nf,
                           float scale, const float *act, const float *map, const float *afac,
                           const int *bview, float fAtnCnvFac, float fMu, float cntr,
                           float fDensFac, float primary, float *out, const float2 *tw)
{
    const int rpb = esse_planes_per_block((long)B * ny);
    const long nb = ((long)B * ny + rpb - 1) / rpb;
    if (nb <= 0 || nb > INT_MAX || nm > P || nf > P) fail(ETYPE_ILLEGAL_VALUE, "bad ESSE 2-D FFT size");
#define EINV(PP)                                                                                \
    hipLaunchKernelGGL(k_esse_inv2d_src<PP>, dim3((unsigned)nb), dim3(DrfFftShape<PP>::NT),     \
                       DrfFftShape<PP>::LDS, s, outk, B, ny, nm, nf, rpb, scale, act, map, afac, \
                       bview, fAtnCnvFac, fMu, cntr, fDensFac, primary, out, tw)
    if (P == 128)
        EINV(128);
    else if (P == 160)
        EINV(160);
    else
        fail(ETYPE_ILLEGAL_VALUE, "ESSE 2-D FFT length %d", P);
#undef EINV
    DFFT_CHECK(hipGetLastError());
}

}  // namespace irlamd
