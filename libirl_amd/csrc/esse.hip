// esse.hip -- ESSE effective scatter source on gfx950 (krnlprj3d.c:549-954).
//
// The reference convolves the (optionally collapsed) rotated activity with
// two or three kernels by real FFTs on a 2x padded grid.  Here the same
// circular convolutions are evaluated as linear convolutions with the folded
// kernels on power-of-two grids (esse_host.cpp), by a batched hand-written
// FFT over all views of a subset.  Axes: "f" = the volume layout's fastest
// axis, "m" = the other in-plane axis, y = depth.
//   F forward : real lines along f -> half spectrum (kf <= Pf/2, the input
//               is real), lines beyond the volume skipped
//   M forward : lines along m
//   Y conv    : forward along depth, multiply by every kernel spectrum,
//               inverse along depth, keep the cropped depth range
//   M inverse, F inverse (complex-to-real): back to the cropped volume
// Each line FFT is a Stockham transform held in registers (FftPlan: R
// values per thread, P/R threads per line, one LDS exchange between passes;
// radix-16 passes for powers of two, radices 4, 4, 10 for P = 160 and 4, 4,
// 12 for P = 192, the 10- and 12-point DFTs done directly).  Lines along strided axes
// are grouped so that consecutive lanes take consecutive columns
// (coalesced); lines along the fastest axis put consecutive lanes on
// consecutive elements.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "drf_fft.h"
#include "esse_dev.h"
#include "fft_dev.h"

namespace irlamd {

#define ESSE_CHECK(expr)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::irlamd::fail(ETYPE_DEVICE, "HIP error '%s' in %s (%s:%d)",                 \
                           hipGetErrorString(e_), #expr, __FILE__, __LINE__);            \
    } while (0)

static long cdiv(long a, long b) { return (a + b - 1) / b; }

// F forward: real lines r = (plane, m), m < nm, of nf values (zero to Pf)
// -> half spectrum buf[(plane*Pm + m)*Wc + k], k < Wc.
template <int P>
__global__ __launch_bounds__(256) void k_fft_f_fwd(const float *__restrict__ in,
                                                   float2 *__restrict__ buf, long rows, int nm,
                                                   int nf, int Pm, int Wc,
                                                   const float2 *__restrict__ tw)
{
    using F = FftLine<P>;
    constexpr int LB = 256 / F::T;
    extern __shared__ float2 sm[];
    const int li = threadIdx.x / F::T, tt = threadIdx.x % F::T;
    const long r = (long)blockIdx.x * LB + li;
    const bool act = r < rows;
    float2 a[F::R];
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        a[w] = (act && n < nf) ? make_float2(in[r * nf + n], 0.0f) : make_float2(0.0f, 0.0f);
    }
    fft_line<P, false>(a, sm + li * F::LS, tt, tw);
    if (!act) return;
    const long plane = r / nm, m = r - plane * nm;
    float2 *dst = buf + (plane * Pm + m) * Wc;
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        if (n < Wc) dst[n] = a[w];
    }
}

// F inverse (complex to real): half spectra of the lines (plane, m < nm)
// -> real lines out[r*nf + n], n < nf, times scale.  X[k] for k > P/2 is
// conj(X[P-k]).
template <int P>
__global__ __launch_bounds__(256) void k_fft_f_inv(const float2 *__restrict__ buf,
                                                   float *__restrict__ out, long rows, int nm,
                                                   int nf, int Pm, int Wc,
                                                   const float2 *__restrict__ tw, float scale)
{
    using F = FftLine<P>;
    constexpr int LB = 256 / F::T;
    extern __shared__ float2 sm[];
    const int li = threadIdx.x / F::T, tt = threadIdx.x % F::T;
    const long r = (long)blockIdx.x * LB + li;
    const bool act = r < rows;
    const long plane = act ? r / nm : 0, m = act ? r - plane * nm : 0;
    const float2 *src = buf + (plane * Pm + m) * Wc;
    float2 a[F::R];
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        float2 v = make_float2(0.0f, 0.0f);
        if (act) {
            if (n <= P / 2) {
                v = src[n];
            } else {
                v = src[P - n];
                v.y = -v.y;
            }
        }
        a[w] = v;
    }
    fft_line<P, true>(a, sm + li * F::LS, tt, tw);
    if (!act) return;
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        if (n < nf) out[r * nf + n] = a[w].x * scale;
    }
}

// F inverse of terms 0 and 1 of line (b, y, m) as one complex line (the
// two half spectra completed by Hermitian symmetry as its real and
// imaginary parts), then the effective source of CalcEffScatterSource for
// two Taylor terms (krnlprj3d.c:903-947, as k_esse_source) on the line's
// voxels, slice-fastest layout, no collapse:
//   eff = (E0 - d E1) exp(d fMu) (+ cntr act) * map fDens, clamped at 0,
//   out = (act primary + eff) afac,
// with the depth d = S / mu0 from the factors afac = exp(-S), logarithm and
// exponential in single precision (k_esse_source takes the exponential in
// double as the reference; the two differ by rounding).
template <int P>
__global__ __launch_bounds__(256) void k_fft_f_inv_src(
    const float2 *__restrict__ buf, long rows, int nm, int nf, int ny, int Pm, int Wc,
    long tstride, const float2 *__restrict__ tw, float scale, const float *__restrict__ act,
    const float *__restrict__ map, const float *__restrict__ afac, const int *__restrict__ bview,
    float fAtnCnvFac, float fMu, float cntr, float fDensFac, float primary,
    float *__restrict__ out)
{
    using F = FftLine<P>;
    constexpr int LB = 256 / F::T;
    extern __shared__ float2 sm[];
    const int li = threadIdx.x / F::T, tt = threadIdx.x % F::T;
    const long r = (long)blockIdx.x * LB + li;
    const bool act_ = r < rows;
    const long rr = act_ ? r : 0;
    const long b = rr / ((long)ny * nm), yl = rr - b * ny * nm;   // yl = y nm + m
    // term t of line (b, y, m): ((b T + t) ny + y) Pm + m, T = 2
    const long y = yl / nm, m = yl - y * nm;
    const float2 *s0 = buf + (((b * 2) * ny + y) * Pm + m) * Wc;
    const float2 *s1 = s0 + tstride;
    float2 a[F::R];
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        const bool up = n > P / 2;
        const int k = up ? P - n : n;
        float2 A = s0[k], B = s1[k];
        if (up) {
            A.y = -A.y;
            B.y = -B.y;
        }
        a[w] = make_float2(A.x - B.y, A.y + B.x);
    }
    fft_line<P, true, true>(a, sm + li * F::LS, tt, tw);
    if (!act_) return;
    const long vol = (long)ny * nm * nf;
    const long vo = (long)bview[b] * vol + yl * nf;   // this line in the view's volumes
    const long bo = b * vol + yl * nf;                // and in the batch's
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        if (n >= nf) continue;
        const float e0 = a[w].x * scale, e1 = a[w].y * scale;
        const float f = afac[vo + n], mp = map[vo + n], av = act[bo + n];
        // f == 0 (depth past float range): d = 0 keeps eff finite, out is 0
        const float d = f > 0.0f ? -logf(f) * fAtnCnvFac : 0.0f;
        float eff = e0 - d * e1;
        eff = eff * expf(d * fMu);
        if (cntr > 0) eff = eff + av * cntr;
        eff = eff * (mp * fDensFac);
        if (eff < 0) eff = 0.0f;
        out[bo + n] = (av * primary + eff) * f;
    }
}

// Strided lines: element n of line (bi, c), c < W, is at bi*bstride +
// n*nstride + c.  Input n >= nIn is zero; n < nOut is written to dst (may
// alias src).  Blocks take LB consecutive lines of the flattened (bi, c)
// order, so only the very last block has idle lanes.
template <int P, bool INV>
__global__ __launch_bounds__(256) void k_fft_strided(const float2 *src, float2 *dst, long bstride,
                                                     long nstride, int W, int nIn, int nOut,
                                                     long nlines, const float2 *__restrict__ tw)
{
    using F = FftLine<P>;
    constexpr int LB = 256 / F::T;
    extern __shared__ float2 sm[];
    const int li = threadIdx.x % LB, tt = threadIdx.x / LB;
    // lines (base, column) flattened, LB consecutive ones per block
    const long line = (long)blockIdx.x * LB + li;
    const bool act = line < nlines;
    const long bi = act ? line / W : 0;
    const int c = act ? (int)(line - bi * W) : 0;
    const long off = bi * bstride + c;
    float2 a[F::R];
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        a[w] = (act && n < nIn) ? src[off + n * nstride] : make_float2(0.0f, 0.0f);
    }
    fft_line<P, INV>(a, sm + li * F::LS, tt, tw);
    if (!act) return;
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        if (n < nOut) dst[off + n * nstride] = a[w];
    }
}

// Logical block index of a launch whose blocks are dispatched round-robin
// over the chip's 8 XCDs (each with its own L2): logically consecutive
// blocks land on the same XCD, so blocks that share inputs share its L2.
__device__ __forceinline__ long xcd_block()
{
    constexpr int NXCD = 8;
    const long n = gridDim.x, p = blockIdx.x;
    const long per = n / NXCD, rem = n % NXCD;
    const long xcd = p % NXCD, k = p / NXCD;
    return (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + k;
}

// Y conv of the kernel terms: lines along depth at (b, m, column c).
// Forward FFT (rows y >= ny are zero) once, then per term: multiply by the
// term's spectrum, inverse FFT, keep y < ny.  Between the phases the lane
// indices pass through an empty asm, so each phase computes its own LDS
// addresses instead of the compiler keeping those of every phase live (that
// kept the kernel at two waves per SIMD).  The batch index runs fastest over
// the logical blocks, so the B blocks that read the same spectrum columns
// run together on one XCD and share them through its L2.
template <int P, int NT>
__global__ __launch_bounds__(256) void k_fft_y_conv(const float2 *__restrict__ buf,
                                                    const float2 *__restrict__ khat,
                                                    long khatStride, float2 *__restrict__ outk,
                                                    long outTStride, long outBStride, int B,
                                                    int ny, int Pm, int Wc,
                                                    const float2 *__restrict__ tw)
{
    using F = FftLine<P>;
    constexpr int LB = 256 / F::T;
    extern __shared__ float2 sm[];
    int li = threadIdx.x % LB, tt = threadIdx.x / LB;
    // columns (m, kc) flattened: a block's LB lanes take LB consecutive
    // elements of a depth row, across m boundaries, none idle but the last
    const long ystride = (long)Pm * Wc;
    long bid = xcd_block();
    const int b = (int)(bid % B);
    const long col = (bid / B) * LB + li;
    const bool act = col < ystride;
    const float2 *src = buf + (long)b * ny * ystride + col;
    float2 a[F::R];
#pragma unroll
    for (int w = 0; w < F::R; ++w) {
        const int n = tt + F::T * w;
        a[w] = (act && n < ny) ? src[n * ystride] : make_float2(0.0f, 0.0f);
    }
    fft_line<P, false>(a, sm + li * F::LS, tt, tw);
#pragma unroll 1
    for (int t = 0; t < NT; ++t) {
        asm volatile("" : "+v"(tt), "+v"(li));
        const float2 *kt = khat + t * khatStride + col;
        float2 c[F::R];
#pragma unroll
        for (int w = 0; w < F::R; ++w) {
            const int n = tt + F::T * w;
            c[w] = act ? cmul(a[w], kt[n * ystride]) : make_float2(0.0f, 0.0f);
        }
        fft_line<P, true>(c, sm + li * F::LS, tt, tw);
        if (act) {
            float2 *dst = outk + (long)b * outBStride + t * outTStride + col;
#pragma unroll
            for (int w = 0; w < F::R; ++w) {
                const int n = tt + F::T * w;
                if (n < ny) dst[n * ystride] = c[w];
            }
        }
    }
}

// Volume layouts: zfast = [b][ny][nx][nz] (the engine's slice-fastest
// order), else [b][ny][nz][nx] (the reference's image order).
struct VolStrides {
    long sy;
    int sz, sx;
};

__device__ __forceinline__ VolStrides vol_strides(int nz, int nx, bool zfast)
{
    return VolStrides{(long)nz * nx, zfast ? 1 : nx, zfast ? nz : 1};
}

// Decodes the linear index of an element of a [b][ny](nz,nx) volume.
__device__ __forceinline__ void vol_coords(long i, int ny, int nz, int nx, bool zfast, long &b,
                                           int &y, int &z, int &x)
{
    long r;
    if (zfast) {
        z = (int)(i % nz);
        r = i / nz;
        x = (int)(r % nx);
        r /= nx;
    } else {
        x = (int)(i % nx);
        r = i / nx;
        z = (int)(r % nz);
        r /= nz;
    }
    y = (int)(r % ny);
    b = r / ny;
}

// CollapseByFac (krnlprj3d.c:549-593): sum of c^3 blocks, summed with the
// depth index slowest and the bin index fastest.
__global__ void k_collapse(const float *__restrict__ in, float *__restrict__ out, long nOut,
                           int ny, int nz, int nx, int c, bool zfast)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nOut) return;
    long b;
    int yc, zc, xc;
    vol_coords(i, ny / c, nz / c, nx / c, zfast, b, yc, zc, xc);
    const VolStrides st = vol_strides(nz, nx, zfast);
    const float *src = in + b * (long)ny * nz * nx;
    float s = 0.0f;
    for (int k = 0; k < c; ++k)
        for (int j = 0; j < c; ++j)
            for (int ii = 0; ii < c; ++ii)
                s += src[(long)(yc * c + k) * st.sy + (long)(zc * c + j) * st.sz +
                         (long)(xc * c + ii) * st.sx];
    out[i] = s;
}

// ExpandByFac replication (krnlprj3d.c:610-630): value / c / c / c.
__global__ void k_expand_rep(const float *__restrict__ in, float *__restrict__ out, long nOut,
                             int ny, int nz, int nx, int c, bool zfast)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nOut) return;
    long b;
    int y, z, x;
    vol_coords(i, ny, nz, nx, zfast, b, y, z, x);
    const int nxc = nx / c, nzc = nz / c, nyc = ny / c;
    const VolStrides st = vol_strides(nzc, nxc, zfast);
    out[i] = in[b * (long)nyc * nzc * nxc + (long)(y / c) * st.sy + (long)(z / c) * st.sz +
                (long)(x / c) * st.sx] / c / c / c;
}

// ExpandByFac in-place box smoothing along one axis (krnlprj3d.c:632-670):
// v[k] = (sum of the window k-jj..k+jj, whose earlier entries are already
// smoothed) / (2jj+1), k walking upward.  One thread per line; axis 0 =
// depth rows, 1 = slices, 2 = bins of [bt][ny](nz,nx) volumes.
__global__ void k_smooth(float *__restrict__ img, long BT, int ny, int nz, int nx, int axis,
                         int jj, bool zfast)
{
    const long pl = (long)nz * nx, vol = pl * ny;
    const VolStrides st = vol_strides(nz, nx, zfast);
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long base, stride;
    int len;
    if (axis == 0) {
        if (t >= BT * pl) return;
        base = (t / pl) * vol + t % pl;
        stride = pl;
        len = ny;
    } else if (axis == 1) {
        if (t >= BT * ny * nx) return;
        const long bt = t / ((long)ny * nx), r = t % ((long)ny * nx);
        base = bt * vol + (r / nx) * pl + (r % nx) * st.sx;
        stride = st.sz;
        len = nz;
    } else {
        if (t >= BT * ny * nz) return;
        const long bt = t / ((long)ny * nz), r = t % ((long)ny * nz);
        base = bt * vol + (r / nz) * pl + (r % nz) * st.sz;
        stride = st.sx;
        len = nx;
    }
    float *p = img + base;
    for (int k = jj; k < len - jj; ++k) {
        float s = 0.0f;
        for (int kk = k - jj; kk <= k + jj; ++kk) s += p[(long)kk * stride];
        p[(long)k * stride] = s / (2 * jj + 1);
    }
}

// Effective source, primaries and attenuation of one column (krnlprj3d.c:
// 826-948; prjbck.c:170-173, 219-222).  Walks the depth rows with the
// half-pixel running sum of mu, as the reference's three depth loops do.
// Every operand has the same layout; only the depth stride nz*nx matters.
template <int T>
__global__ __launch_bounds__(256) void k_esse_source(
    const float *__restrict__ act, const float *__restrict__ map, const float *__restrict__ terms,
    int B, int ny, int nz, int nx, float fAtnCnvFac, float fMu, float cntr, float fDensFac,
    float primary, const float *__restrict__ afac, const int *__restrict__ bview, int mapByView,
    float *__restrict__ out)
{
    const long cols = (long)nz * nx;
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= cols * B) return;
    const int b = (int)(gid / cols);
    const long c = gid % cols;
    const long vol = cols * ny;
    const float *A = act + b * vol + c;
    const float *Mp = map + (mapByView ? (long)bview[b] : (long)b) * vol + c;
    const float *E = terms + (long)b * T * vol + c;
    const float *F = afac ? afac + (long)bview[b] * vol + c : nullptr;
    float *O = out + b * vol + c;
    float sum = 0.0f;
    // four depth rows per step: their loads are issued before the running
    // sum, which is the only serial dependence, walks through them
    constexpr int U = 4;
    for (int y0 = 0; y0 < ny; y0 += U) {
        float m[U], e0[U], e1[U], e2[U], a[U], f[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const long o = (long)min(y0 + k, ny - 1) * cols;
            m[k] = Mp[o];
            e0[k] = E[o];
            e1[k] = T > 1 ? E[vol + o] : 0.0f;
            e2[k] = T > 2 ? E[2 * vol + o] : 0.0f;
            a[k] = A[o];
            f[k] = F ? F[o] : 1.0f;
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            if (y0 + k >= ny) break;
            sum = (float)((double)sum + 0.5 * (double)m[k]);
            const float d = sum * fAtnCnvFac;
            float eff = e0[k];
            if (T > 1) eff = eff - d * e1[k];
            if (T > 2) eff = (float)((double)eff + 0.5 * (double)d * (double)d * (double)e2[k]);
            eff = (float)((double)eff * exp((double)(d * fMu)));
            sum = (float)((double)sum + 0.5 * (double)m[k]);
            if (cntr > 0) eff = eff + a[k] * cntr;
            eff = eff * (m[k] * fDensFac);
            if (eff < 0) eff = 0.0f;
            const float q = a[k] * primary + eff;
            O[(long)(y0 + k) * cols] = F ? q * f[k] : q;
        }
    }
}

// W_P^k = exp(-2 pi i k/P), k < P, in double rounded to float.
template <typename T>
static T *dalloc(size_t n)
{
    T *p = nullptr;
    if (n) ESSE_CHECK(hipMalloc((void **)&p, n * sizeof(T)));
    return p;
}

// threads per line of each supported length (FftPlan<P>::T)
static int fft_threads(int P)
{
    switch (P) {
    case 16: return 1;
    case 32: return 2;
    case 64: return 4;
    case 128: return 8;
    case 160: return 8;
    case 192: return 16;
    case 256: return 16;
    case 512: return 32;
    }
    fail(ETYPE_ILLEGAL_VALUE, "unsupported ESSE FFT length %d", P);
}

static int lines_per_block(int P) { return 256 / fft_threads(P); }
static size_t fft_lds(int P) { return (size_t)lines_per_block(P) * (P + P / 16 + 1) * sizeof(float2); }

// Half-spectrum width along the fastest axis: kf <= P/2, padded to 16.
static int half_width(int P) { return std::min(P, P / 2 + 1); }

#define ESSE_P_DISPATCH(P, LAUNCH)                                                           \
    do {                                                                                      \
        switch (P) {                                                                          \
        case 16: LAUNCH(16); break;                                                           \
        case 32: LAUNCH(32); break;                                                           \
        case 64: LAUNCH(64); break;                                                           \
        case 128: LAUNCH(128); break;                                                         \
        case 160: LAUNCH(160); break;                                                         \
        case 192: LAUNCH(192); break;                                                         \
        case 256: LAUNCH(256); break;                                                         \
        case 512: LAUNCH(512); break;                                                         \
        default: fail(ETYPE_ILLEGAL_VALUE, "unsupported ESSE FFT length %d", P);               \
        }                                                                                     \
        ESSE_CHECK(hipGetLastError());                                                        \
    } while (0)

void EsseDev::fft_f_fwd(const float *in, float2 *out, long rows, int nm, int nf)
{
    const dim3 grid((unsigned)cdiv(rows, lines_per_block(Pf_)));
#define L_(PP) hipLaunchKernelGGL(k_fft_f_fwd<PP>, grid, dim3(256), fft_lds(PP), stream_, in, out, \
                                  rows, nm, nf, Pm_, Wc_, twf_)
    ESSE_P_DISPATCH(Pf_, L_);
#undef L_
}

void EsseDev::fft_f_inv(const float2 *in, float *out, long rows, int nm, int nf, float scale)
{
    const dim3 grid((unsigned)cdiv(rows, lines_per_block(Pf_)));
#define L_(PP) hipLaunchKernelGGL(k_fft_f_inv<PP>, grid, dim3(256), fft_lds(PP), stream_, in, out, \
                                  rows, nm, nf, Pm_, Wc_, twf_, scale)
    ESSE_P_DISPATCH(Pf_, L_);
#undef L_
}

void EsseDev::fft_strided(int P, bool inv, const float2 *src, float2 *dst, long nbases,
                          long bstride, long nstride, int W, int nIn, int nOut, const float2 *tw)
{
    const long nlines = nbases * W;
    const dim3 grid((unsigned)cdiv(nlines, lines_per_block(P)));
#define L_(PP)                                                                                  \
    do {                                                                                        \
        if (inv)                                                                                \
            hipLaunchKernelGGL((k_fft_strided<PP, true>), grid, dim3(256), fft_lds(PP), stream_, \
                               src, dst, bstride, nstride, W, nIn, nOut, nlines, tw);           \
        else                                                                                    \
            hipLaunchKernelGGL((k_fft_strided<PP, false>), grid, dim3(256), fft_lds(PP),        \
                               stream_, src, dst, bstride, nstride, W, nIn, nOut, nlines, tw);  \
    } while (0)
    ESSE_P_DISPATCH(P, L_);
#undef L_
}

EsseDev::EsseDev(const EsseModel &m, int maxBatch, hipStream_t stream, bool zfast)
    : m_(m), maxB_(maxBatch), stream_(stream), zfast_(zfast)
{
    if (!m_.enabled) return;
    const int Px = m_.Px, Py = m_.Py, Pz = m_.Pz, T = m_.nterms;
    // The FFT passes run along the layout's fastest axis first ("f"), then
    // the other in-plane axis ("m"), then depth.
    Pf_ = zfast ? Pz : Px;
    Pm_ = zfast ? Px : Pz;
    Wc_ = half_width(Pf_);
    auto tf = twiddles(Pf_), ty = twiddles(Py), tm = twiddles(Pm_);
    twf_ = dalloc<float2>(tf.size());
    twy_ = dalloc<float2>(ty.size());
    twm_ = dalloc<float2>(tm.size());
    ESSE_CHECK(hipMemcpyAsync(twf_, tf.data(), tf.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
    ESSE_CHECK(hipMemcpyAsync(twy_, ty.data(), ty.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
    ESSE_CHECK(hipMemcpyAsync(twm_, tm.data(), tm.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
    const long spec = (long)Py * Pm_ * Wc_;
    khat_ = dalloc<float2>(spec * T);
    buf_ = dalloc<float2>((size_t)maxB_ * m_.nyc * Pm_ * Wc_);
    outk_ = dalloc<float2>((size_t)maxB_ * T * m_.nyc * Pm_ * Wc_);
    if (m_.c > 1) {
        coll_ = dalloc<float>((size_t)maxB_ * m_.nyc * m_.nzc * m_.nxc);
        res_ = dalloc<float>((size_t)maxB_ * T * m_.nyc * m_.nzc * m_.nxc);
    }
    // Kernel spectra [T][Py][Pm][Wc]: 3-D forward FFT of each folded kernel
    // grid (held [Py][Pz][Px], transposed for the slice-fastest layout).
    const long grid = (long)Px * Py * Pz;
    float *dgrid = dalloc<float>(grid);
    float2 *tmp = dalloc<float2>(spec);
    std::vector<float> h(grid);
    for (int t = 0; t < T; ++t) {
        const float *g = m_.grid[t].data();
        if (zfast) {
            for (int y = 0; y < Py; ++y)
                for (int z = 0; z < Pz; ++z)
                    for (int x = 0; x < Px; ++x)
                        h[((size_t)y * Px + x) * Pz + z] = g[((size_t)y * Pz + z) * Px + x];
            g = h.data();
        }
        ESSE_CHECK(hipMemcpyAsync(dgrid, g, grid * sizeof(float), hipMemcpyHostToDevice, stream_));
        fft_f_fwd(dgrid, tmp, (long)Py * Pm_, Pm_, Pf_);
        fft_strided(Pm_, false, tmp, tmp, Py, (long)Pm_ * Wc_, Wc_, Wc_, Pm_, Pm_, twm_);
        fft_strided(Py, false, tmp, khat_ + spec * t, 1, 0, (long)Pm_ * Wc_, Pm_ * Wc_, Py, Py,
                    twy_);
        // the host grid is reused by the next term: wait for the copy
        ESSE_CHECK(hipStreamSynchronize(stream_));
    }
    ESSE_CHECK(hipStreamSynchronize(stream_));
    (void)hipFree(dgrid);
    (void)hipFree(tmp);
}

EsseDev::~EsseDev()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void *p : {(void *)khat_, (void *)buf_, (void *)outk_, (void *)coll_, (void *)res_,
                    (void *)twf_, (void *)twy_, (void *)twm_})
        if (p) (void)hipFree(p);
}

// Slice and bin forward passes of planes [planes][nm][nf] into buf_: one
// 2-D transform per plane in LDS where both lengths allow it (drf_fft.hip),
// else the two line passes.
// (the plane kernels take one CU per block: for fewer than 1024 planes, as
// a single view's check, the line passes keep more of the chip busy)
static bool plane_kernels(int Pf, int Pm, long planes)
{
    return esse_fft2_supported(Pf, Pm) && planes >= 1024;
}

void EsseDev::forward2(const float *in, long planes, int nm, int nf)
{
    if (fused_ && plane_kernels(Pf_, Pm_, planes)) {
        launch_esse_fwd2d(stream_, Pf_, in, buf_, planes, nm, nf, twf_);
        return;
    }
    fft_f_fwd(in, buf_, planes * nm, nm, nf);
    fft_strided(Pm_, false, buf_, buf_, planes, (long)Pm_ * Wc_, Wc_, Wc_, nm, Pm_, twm_);
}

void EsseDev::convolve(const float *act, int B, float *terms)
{
    if (!m_.enabled) return;
    if (B > maxB_) fail(ETYPE_ILLEGAL_VALUE, "ESSE batch %d exceeds %d", B, maxB_);
    const int Py = m_.Py, Pm = Pm_, Wc = Wc_, T = m_.nterms;
    const int ny = m_.nyc, c = m_.c;
    const int nf = zfast_ ? m_.nzc : m_.nxc, nm = zfast_ ? m_.nxc : m_.nzc;
    const float *src = act;
    if (c > 1) {
        const long n = (long)B * ny * nf * nm;
        hipLaunchKernelGGL(k_collapse, dim3(cdiv(n, 256)), dim3(256), 0, stream_, act, coll_, n,
                           m_.ny, m_.nz, m_.nx, c, zfast_);
        ESSE_CHECK(hipGetLastError());
        src = coll_;
    }
    const long plane = (long)Pm * Wc;
    forward2(src, B * ny, nm, nf);
    {
        const dim3 grid((unsigned)((long)B * cdiv((long)Pm * Wc, lines_per_block(Py))));
        const long khs = (long)Py * plane, ots = (long)ny * plane, obs = (long)T * ny * plane;
#define L_(PP)                                                                                 \
    if (T == 1)                                                                                \
        hipLaunchKernelGGL((k_fft_y_conv<PP, 1>), grid, dim3(256), fft_lds(PP), stream_, buf_, \
                           khat_, khs, outk_, ots, obs, B, ny, Pm, Wc, twy_);                  \
    else if (T == 2)                                                                           \
        hipLaunchKernelGGL((k_fft_y_conv<PP, 2>), grid, dim3(256), fft_lds(PP), stream_, buf_, \
                           khat_, khs, outk_, ots, obs, B, ny, Pm, Wc, twy_);                  \
    else                                                                                       \
        hipLaunchKernelGGL((k_fft_y_conv<PP, 3>), grid, dim3(256), fft_lds(PP), stream_, buf_, \
                           khat_, khs, outk_, ots, obs, B, ny, Pm, Wc, twy_)
        ESSE_P_DISPATCH(Py, L_);
#undef L_
    }
    fft_strided(Pm, true, outk_, outk_, (long)B * T * ny, plane, Wc, Wc, Pm, nm, twm_);
    const float scale = (float)(1.0 / ((double)m_.Px * Py * m_.Pz));
    float *dst = c > 1 ? res_ : terms;
    fft_f_inv(outk_, dst, (long)B * T * ny * nm, nm, nf, scale);
    if (c > 1) {
        const int NY = m_.ny, NZ = m_.nz, NX = m_.nx, jj = c / 2;
        const long n = (long)B * T * NY * NZ * NX;
        hipLaunchKernelGGL(k_expand_rep, dim3(cdiv(n, 256)), dim3(256), 0, stream_, res_, terms, n,
                           NY, NZ, NX, c, zfast_);
        const long BT = (long)B * T;
        for (int axis = 0; axis < 3; ++axis) {
            const long lines = axis == 0 ? BT * NZ * NX : axis == 1 ? BT * NY * NX : BT * NY * NZ;
            hipLaunchKernelGGL(k_smooth, dim3(cdiv(lines, 256)), dim3(256), 0, stream_, terms, BT,
                               NY, NZ, NX, axis, jj, zfast_);
        }
        ESSE_CHECK(hipGetLastError());
    }
}

bool EsseDev::convolveSource(const float *act, const float *map, int B, float primary,
                             const float *afac, const int *bview, float *out)
{
    if (!canConvolveSource(map, afac, bview)) return false;
    if (B > maxB_) fail(ETYPE_ILLEGAL_VALUE, "ESSE batch %d exceeds %d", B, maxB_);
    const int Py = m_.Py, Pm = Pm_, Wc = Wc_, T = 2;
    const int ny = m_.nyc, nf = m_.nzc, nm = m_.nxc;
    const long plane = (long)Pm * Wc;
    forward2(act, B * ny, nm, nf);
    {
        const dim3 grid((unsigned)((long)B * cdiv((long)Pm * Wc, lines_per_block(Py))));
        const long khs = (long)Py * plane, ots = (long)ny * plane, obs = (long)T * ny * plane;
#define L_(PP)                                                                                 \
    hipLaunchKernelGGL((k_fft_y_conv<PP, 2>), grid, dim3(256), fft_lds(PP), stream_, buf_,     \
                       khat_, khs, outk_, ots, obs, B, ny, Pm, Wc, twy_)
        ESSE_P_DISPATCH(Py, L_);
#undef L_
    }
    const float scale = (float)(1.0 / ((double)m_.Px * Py * m_.Pz));
    const float fAtnCnvFac = 1.0 / (m_.mu0);
    const float fMu = m_.mu0 - m_.muavg;
    const float fDens = fAtnCnvFac / m_.pixsize;
    if (plane_kernels(Pf_, Pm_, (long)B * ny)) {
        // bin and slice inverses of both terms per plane, source in the epilogue
        launch_esse_inv2d_src(stream_, Pf_, outk_, B, ny, nm, nf, scale, act, map, afac, bview,
                              fAtnCnvFac, fMu, m_.cntrPixFac, fDens, primary, out, twf_);
        return true;
    }
    fft_strided(Pm, true, outk_, outk_, (long)B * T * ny, plane, Wc, Wc, Pm, nm, twm_);
    const long rows = (long)B * ny * nm;
    const dim3 grid((unsigned)cdiv(rows, lines_per_block(Pf_)));
#define L_(PP)                                                                                 \
    hipLaunchKernelGGL(k_fft_f_inv_src<PP>, grid, dim3(256), fft_lds(PP), stream_, outk_, rows, \
                       nm, nf, ny, Pm, Wc, (long)ny * plane, twf_, scale, act, map, afac,     \
                       bview, fAtnCnvFac, fMu, m_.cntrPixFac, fDens, primary, out)
    ESSE_P_DISPATCH(Pf_, L_);
#undef L_
    return true;
}

void EsseDev::source(const float *act, const float *map, const float *terms, int B,
                     float primary, const float *afac, const int *bview, float *out,
                     bool mapByView)
{
    const long cols = (long)m_.nz * m_.nx;
    const float fAtnCnvFac = 1.0 / (m_.mu0);
    const float fMu = m_.mu0 - m_.muavg;
    const float fDens = fAtnCnvFac / m_.pixsize;
    const dim3 grid((unsigned)cdiv(cols * B, 256));
#define L_(TT)                                                                                \
    hipLaunchKernelGGL(k_esse_source<TT>, grid, dim3(256), 0, stream_, act, map, terms, B,    \
                       m_.ny, m_.nz, m_.nx, fAtnCnvFac, fMu, m_.cntrPixFac, fDens, primary,   \
                       afac, bview, mapByView ? 1 : 0, out)
    if (m_.nterms == 1)
        L_(1);
    else if (m_.nterms == 2)
        L_(2);
    else
        L_(3);
#undef L_
    ESSE_CHECK(hipGetLastError());
}

}  // namespace irlamd

namespace irlamd {

// Handle behind pvInitESSEScatterModel / CalcEffScatterSource: the model
// and a device instance for one volume at a time.
struct EsseHandle {
    EsseModel model;
    hipStream_t stream = nullptr;
    std::unique_ptr<EsseDev> dev;
};

void *esse_handle_create(const EsseModel &m)
{
    EsseHandle *h = new EsseHandle;
    h->model = m;
    return h;
}

void esse_handle_free(void *p)
{
    EsseHandle *h = (EsseHandle *)p;
    if (!h) return;
    h->dev.reset();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

bool esse_handle_enabled(const void *p) { return p && ((const EsseHandle *)p)->model.enabled; }

const EsseModel *esse_handle_model(const void *p) { return p ? &((const EsseHandle *)p)->model : nullptr; }

// CalcEffScatterSource on host arrays [ny][nz][nx] (krnlprj3d.c:676-954).
void esse_handle_effsrc(void *p, const float *act, const float *map, int nx, int ny, int nz,
                        float *eff)
{
    EsseHandle *h = (EsseHandle *)p;
    const long n = (long)nx * ny * nz;
    if (!h || !h->model.enabled) {
        memset(eff, 0, n * sizeof(float));
        return;
    }
    const EsseModel &m = h->model;
    if (nx != m.nx || ny != m.ny || nz != m.nz)
        fail(ETYPE_ILLEGAL_VALUE, "scatter model prepared for %d x %d x %d, called with %d x %d x %d",
             m.ny, m.nz, m.nx, ny, nz, nx);
    if (!h->stream) ESSE_CHECK(hipStreamCreate(&h->stream));
    if (!h->dev) h->dev.reset(new EsseDev(m, 1, h->stream, false));
    float *da = dalloc<float>(n), *dm = dalloc<float>(n), *dt = dalloc<float>(n * m.nterms),
          *de = dalloc<float>(n);
    ESSE_CHECK(hipMemcpyAsync(da, act, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    ESSE_CHECK(hipMemcpyAsync(dm, map, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    h->dev->convolve(da, 1, dt);
    h->dev->source(da, dm, dt, 1, 0.0f, nullptr, nullptr, de);
    ESSE_CHECK(hipMemcpyAsync(eff, de, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    ESSE_CHECK(hipStreamSynchronize(h->stream));
    for (float *q : {da, dm, dt, de}) (void)hipFree(q);
}

}  // namespace irlamd
