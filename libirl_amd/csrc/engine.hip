// engine.hip -- gfx950 kernels and the device-resident OSEM engine.
//
// Data layout in HBM (one reconstruction, N = NumPixels = NumRotPixs,
// Nz = NumSlices, V views, S subsets of Vs views):
//   x, mu, U, Ninv[s]   [z][y][x] fp32 volumes, M = N*N*Nz floats each: one
//                       slice's (y,x) plane is a contiguous N*N block, which is
//                       the unit the rotations stage in LDS (the reference's
//                       internal [y][z][x] order, irlosem.c:171-195, exists
//                       only on the host side of the hooks)
//   prj, scat           [view][slice][bin] fp32
//   afac                [V][row][slice][bin] fp32 attenuation factors
//                       exp(-sum of half-pixel mu) of every view, computed
//                       once per reconstruction (mu never changes)
//   att                 [Vs][row][slice][bin] rotated estimate x afac
//   ppart               [Vs][chunk][slice][bin] partial projections
//   ratio               [Vs][slice][bin]
//   bplanes             [Vs][slice][row][bin] convolved backprojection planes
//                       (slice-major, so back rotation stages one contiguous
//                       NRot*NRot plane per slice)
//
// Kernels (reference loops they replace):
//   k_rot3dpar   rot3dpar.c:92-192 (generic, prjbck-level entry point)
//   k_afac       prjbck.c:146-163, 212-225, 247-250 (rotated mu, running
//                half-pixel line integral, exp) for every view at once
//   k_att        prjbck.c:129-140, 219-225 (rotate x, attenuate)
//   k_drf<PRJ>   prjbck.c:195-251 + conv2d_ypad.c:92-103: per-plane DRF
//                convolution accumulated into the projection
//   k_ratio      osemrecon.c:727-820
//   k_drf<BCK>   prjbck.c:346-393 + conv2d_ypad.c:92-103
//   k_bck<false> prjbck.c:408-426 + osemrecon.c:850-852: gather-form back
//                rotation, per-view EPS clamp, multiplicative update
//   k_bck<true>  osemrecon.c:282-320 with FP64 accumulation
#include <hip/hip_runtime.h>

#include <climits>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <type_traits>
#include <map>
#include <mutex>
#include <set>
#include <vector>

#include "cnn_dev.h"
#include "drf_dev.h"
#include "drf_fft.h"
#include "engine.h"
#include "esse_dev.h"

namespace irlamd {

void fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw IrlError{code, buf};
}

#define HIP_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::irlamd::fail(ETYPE_DEVICE, "HIP error '%s' in %s (%s:%d)",                 \
                           hipGetErrorString(e_), #expr, __FILE__, __LINE__);            \
    } while (0)

// ---------------------------------------------------------------------------
// Host geometry helpers (reference float arithmetic)
// ---------------------------------------------------------------------------

// Inscribed-circle row extents of rot3dpar.c:98-101.
static void circle_rows(int NewDim, std::vector<int> &head, std::vector<int> &npix)
{
    head.resize(NewDim);
    npix.resize(NewDim);
    volatile float Radius = 0.5 * NewDim;
    for (int NewY = 0; NewY < NewDim; ++NewY) {
        float rr = Radius * Radius;
        int NumPix = (int)(sqrt(rr - (Radius - NewY - .5) * (Radius - NewY - .5)) + .5);
        NumPix *= 2;
        head[NewY] = (int)(0.5 * (NewDim - NumPix));
        npix[NewY] = NumPix;
    }
}

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// Bilinear sample of rot3dpar.c:119-189 at output pixel (NewX, NewY) of one
// slice.  'plane' points at the slice's element (0, 0); source pixel
// (OldY, OldX) is plane[OldY*ystride + OldX*xstride].  Every multiply and
// add is rounded separately (the library is compiled with -ffp-contract=off),
// so sample positions and taps are those of the reference bit for bit.
__device__ __forceinline__ float rot_sample(const float *__restrict__ plane, long ystride,
                                            long xstride, int OldDim, float X, float Y,
                                            float sn, float cs, float OldAxis)
{
    const float FloatX = X * cs - Y * sn + OldAxis;
    const float FloatY = X * sn + Y * cs + OldAxis;
    const int OldX = (int)FloatX;
    const int OldY = (int)FloatY;
    const bool x0 = OldX >= 0 && OldX < OldDim;
    const bool y0 = OldY >= 0 && OldY < OldDim;
    const bool x1 = OldX + 1 >= 0 && OldX + 1 < OldDim;
    const bool y1 = OldY + 1 >= 0 && OldY + 1 < OldDim;
    const float *p0 = plane + OldY * ystride + OldX * xstride;
    const float z0 = (x0 && y0) ? p0[0] : 0.0f;
    const float z1 = (x0 && y1) ? p0[ystride] : 0.0f;
    const float z2 = (x1 && y0) ? p0[xstride] : 0.0f;
    const float z3 = (x1 && y1) ? p0[ystride + xstride] : 0.0f;
    const float wy = FloatY - (float)OldY;
    float z01 = z0 + wy * (z1 - z0);
    float z23 = z2 + wy * (z3 - z2);
    if (FloatY < 0) {
        z01 = z0;
        z23 = z2;
    }
    const float wx = FloatX - (float)OldX;
    float z0123 = z01 + wx * (z23 - z01);
    if (FloatX < 0) z0123 = z01;
    return z0123;
}

// rot_sample for the engine's slice-fastest volumes, where the 64 lanes of a
// wave are 64 slices of one pixel: the caller passes a wave-uniform position
// (X, Y, sn, cs), so the sample coordinates, the tap validity and the tap
// offsets are scalar values and a tap costs one vector load; offsets within
// a volume fit 32 bits (checked when a reconstruction is created).  The
// float operations are those of rot_sample.
struct RotPos {
    float FloatX, FloatY;
    int OldX, OldY, off;          // off: element offset of tap (OldY, OldX)
    bool t0, t1, t2, t3;          // taps (0,0), (1,0), (0,1), (1,1) (y, x) inside
};
__device__ __forceinline__ RotPos rot_pos(int ystride, int xstride, int OldDim, float X, float Y,
                                          float sn, float cs, float OldAxis)
{
    RotPos q;
    q.FloatX = X * cs - Y * sn + OldAxis;
    q.FloatY = X * sn + Y * cs + OldAxis;
    q.OldX = (int)q.FloatX;
    q.OldY = (int)q.FloatY;
    const bool x0 = q.OldX >= 0 && q.OldX < OldDim;
    const bool y0 = q.OldY >= 0 && q.OldY < OldDim;
    const bool x1 = q.OldX + 1 >= 0 && q.OldX + 1 < OldDim;
    const bool y1 = q.OldY + 1 >= 0 && q.OldY + 1 < OldDim;
    q.t0 = x0 && y0;
    q.t1 = x0 && y1;
    q.t2 = x1 && y0;
    q.t3 = x1 && y1;
    q.off = q.OldY * ystride + q.OldX * xstride;
    return q;
}
__device__ __forceinline__ float rot_lerp(const RotPos &q, float z0, float z1, float z2, float z3)
{
    const float wy = q.FloatY - (float)q.OldY;
    float z01 = z0 + wy * (z1 - z0);
    float z23 = z2 + wy * (z3 - z2);
    if (q.FloatY < 0) {
        z01 = z0;
        z23 = z2;
    }
    const float wx = q.FloatX - (float)q.OldX;
    float z0123 = z01 + wx * (z23 - z01);
    if (q.FloatX < 0) z0123 = z01;
    return z0123;
}
__device__ __forceinline__ float rot_sample_w(const float *__restrict__ plane, int ystride,
                                              int xstride, int OldDim, float X, float Y,
                                              float sn, float cs, float OldAxis)
{
    const RotPos q = rot_pos(ystride, xstride, OldDim, X, Y, sn, cs, OldAxis);
    const float *p0 = plane + q.off;
    const float z0 = q.t0 ? p0[0] : 0.0f;
    const float z1 = q.t1 ? p0[ystride] : 0.0f;
    const float z2 = q.t2 ? p0[xstride] : 0.0f;
    const float z3 = q.t3 ? p0[ystride + xstride] : 0.0f;
    return rot_lerp(q, z0, z1, z2, z3);
}

// Two bf16 values packed in 32 bits (the first in the low half), as floats.
__device__ __forceinline__ float bf_lo(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w)
{
    return __builtin_bit_cast(float, w & 0xffff0000u);
}

// rot_sample_w of two volumes held as bf16 pairs (one 4-byte tap for both):
// the samples of the low and of the high halves.
__device__ __forceinline__ float2 rot_sample_pair(const uint32_t *__restrict__ plane, int ystride,
                                                  int xstride, int OldDim, float X, float Y,
                                                  float sn, float cs, float OldAxis)
{
    const RotPos q = rot_pos(ystride, xstride, OldDim, X, Y, sn, cs, OldAxis);
    const uint32_t *p0 = plane + q.off;
    const uint32_t w0 = q.t0 ? p0[0] : 0u;
    const uint32_t w1 = q.t1 ? p0[ystride] : 0u;
    const uint32_t w2 = q.t2 ? p0[xstride] : 0u;
    const uint32_t w3 = q.t3 ? p0[ystride + xstride] : 0u;
    return make_float2(rot_lerp(q, bf_lo(w0), bf_lo(w1), bf_lo(w2), bf_lo(w3)),
                       rot_lerp(q, bf_hi(w0), bf_hi(w1), bf_hi(w2), bf_hi(w3)));
}

// Value of lane k of a wave (k wave-uniform), as a scalar.
__device__ __forceinline__ int lane_int(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ __forceinline__ float lane_float(float v, int k)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
}

__device__ __forceinline__ bool in_circle(const int *__restrict__ head,
                                          const int *__restrict__ npix, int NewY, int NewX)
{
    const int h = head[NewY];
    return NewX >= h && NewX <= h + npix[NewY] - 1;
}

// ---------------------------------------------------------------------------
// Three-shear rotation (fastrotate.c:330-563) as a gather.  The reference
// shifts whole rows and columns of a slice through two intermediate grids;
// here each output pixel recomputes the four first-pass values and the two
// second-pass values it depends on, with the same float operations, so the
// result is bit-identical.  Tables per angle slot a (host-computed in the
// reference's float arithmetic): par[a] = (iAdd, mDim, b90, b180), per row
// fl/f1/f2 [a][N], per column of the grown grid jT/jS/jE/g1/g2 [a][MD].
// ---------------------------------------------------------------------------
struct ShearDev {
    int N, MD;
    const int4 *par;
    const int *fl;
    const float *f1, *f2;
    const int *jT, *jS, *jE;
    const float *g1, *g2;
    const int *pS, *pE;           // circle row extents [start, end)
};

// First pass (x shift with the 90/180 degree flips), value at row jk,
// column ic of the grown grid.  'src' is the slice's element (0, 0); source
// pixel (row y, column x) is src[y*ys + x*xs].
__device__ __forceinline__ float shear_x1(const ShearDev &t, int a, int4 par,
                                          const float *__restrict__ src, long ys, long xs,
                                          int jk, int ic)
{
    const int N = t.N, iAdd = par.x;
    const int j = jk - iAdd;
    if (j < 0 || j >= N) return 0.0f;
    const int fl = t.fl[a * N + j];
    if (ic < t.pS[j] - fl + iAdd || ic >= t.pE[j] - fl - 1 + iAdd) return 0.0f;
    const float f1 = t.f1[a * N + j], f2 = t.f2[a * N + j];
    float v0, v1;
    if (!par.w) {
        const int c = ic + fl - iAdd;
        if (!par.z) {   // rows as they are
            v0 = src[j * ys + c * xs];
            v1 = src[j * ys + (c + 1) * xs];
        } else {        // 90 degrees: column N-j-1 read along rows
            const int jj = N - j - 1;
            v0 = src[c * ys + jj * xs];
            v1 = src[(c + 1) * ys + jj * xs];
        }
    } else {
        const int c = N - ic - (fl - iAdd + 1);
        if (!par.z) {   // 180 degrees: row N-j-1 read backwards
            const int jj = N - j - 1;
            v0 = src[jj * ys + c * xs];
            v1 = src[jj * ys + (c - 1) * xs];
        } else {        // 270 degrees: column j read backwards along rows
            v0 = src[c * ys + j * xs];
            v1 = src[(c - 1) * ys + j * xs];
        }
    }
    return f1 * v0 + f2 * v1;
}

// Second pass (y shift) at row jr, column ic of the grown grid.
__device__ __forceinline__ float shear_y(const ShearDev &t, int a, int4 par,
                                         const float *__restrict__ src, long ys, long xs, int jr,
                                         int ic)
{
    if (ic < 0 || ic >= par.y) return 0.0f;
    const int k = a * t.MD + ic;
    if (jr < t.jS[k] || jr >= t.jE[k]) return 0.0f;
    const int r = jr + t.jT[k];
    return t.g1[k] * shear_x1(t, a, par, src, ys, xs, r, ic) +
           t.g2[k] * shear_x1(t, a, par, src, ys, xs, r + 1, ic);
}

// Rotated value at output pixel (row j, column i): third pass (x shift) into
// the circle trimmed by one pixel per side.
__device__ __forceinline__ float shear_sample(const ShearDev &t, int a,
                                              const float *__restrict__ src, long ys, long xs,
                                              int i, int j)
{
    if (i < t.pS[j] + 1 || i >= t.pE[j] - 1) return 0.0f;
    const int4 par = t.par[a];
    const int fl = t.fl[a * t.N + j];
    const int jj = j + par.x, c = i + fl + par.x;
    return t.f1[a * t.N + j] * shear_y(t, a, par, src, ys, xs, jj, c) +
           t.f2[a * t.N + j] * shear_y(t, a, par, src, ys, xs, jj, c + 1);
}

// ---------------------------------------------------------------------------
// k_rot3dpar: generic rotation (rot3dpar.c:19-198), one thread per output.
// ---------------------------------------------------------------------------
__global__ void k_rot3dpar(const float *__restrict__ in, float *__restrict__ out, int Nz,
                           int OldDim, int NewDim, float sn, float cs, float NewAxis,
                           float OldAxis, const int *__restrict__ head,
                           const int *__restrict__ npix)
{
    const long total = (long)NewDim * Nz * NewDim;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const int NewX = (int)(i % NewDim);
        const long r = i / NewDim;
        const int slice = (int)(r % Nz);
        const int NewY = (int)(r / Nz);
        float v = 0.0f;
        if (in_circle(head, npix, NewY, NewX)) {
            const float X = (float)NewX - NewAxis;
            const float Y = (float)NewY - NewAxis;
            v = rot_sample(in + (long)slice * OldDim, (long)Nz * OldDim, 1, OldDim, X, Y, sn,
                           cs, OldAxis);
        }
        out[i] = v;
    }
}

// ---------------------------------------------------------------------------
// Device layout of the reconstruction engine: every volume is stored with
// the slice index fastest -- estimate, mu, update and sensitivity as
// [y][x][z], rotated volumes as [row][bin][z], projections as [bin][z].  The
// rotation is about the slice axis, so the 64 lanes of a wave (64
// consecutive slices of one pixel) share one sample position and every
// bilinear tap is one contiguous 256-byte run.
// ---------------------------------------------------------------------------

// Three-shear rotation of an N x Nz x N volume in the reference's [y][z][x]
// order (vShearsRotate), angle slot a of the tables.
__global__ void k_shear_rot(const float *__restrict__ in, float *__restrict__ out, int Nz, int N,
                            ShearDev sd, int a)
{
    const long total = (long)N * Nz * N;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int x = (int)(i % N);
    const long r = i / N;
    const int z = (int)(r % Nz);
    const int y = (int)(r / Nz);
    out[i] = shear_sample(sd, a, in + (long)z * N, (long)Nz * N, 1, x, y);
}

// ---------------------------------------------------------------------------
// k_afac: attenuation factors of every view.  One thread per (view, bin,
// slice) column walks the rows nearest-detector first: S += mu/2,
// a = exp(-S), S += mu/2 (prjbck.c:212-225, 247-250), with S rounded to
// float after each double-precision add exactly as the reference's float
// table.  mu is [y][x][z]; afac is [V][row][bin][z].
// ---------------------------------------------------------------------------
template <bool SHEAR>
__global__ void k_afac(const float *__restrict__ mu, float *__restrict__ afac, int V, int Nz,
                       int N, int NRot, const float *__restrict__ fsin,
                       const float *__restrict__ fcos, float atnFac,
                       const int *__restrict__ head, const int *__restrict__ npix, ShearDev sd)
{
    const long cols = (long)Nz * NRot;
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= cols * V) return;
    const int v = (int)(gid / cols);
    const int c = (int)(gid % cols);
    const int X = c / Nz, z = c % Nz;
    const float NewAxis = 0.5 * (NRot - 1);
    const float OldAxis = 0.5 * (N - 1);
    const float sn = fsin[v], cs = fcos[v];
    const float *plane = mu + z;
    float *o = afac + (long)v * NRot * cols + c;
    const float Xf = (float)X - NewAxis;
    float S = 0.0f;
    for (int Y = 0; Y < NRot; ++Y) {
        float m = 0.0f;
        if (SHEAR)
            m = shear_sample(sd, 2 * v, plane, (long)N * Nz, Nz, X, Y);
        else if (in_circle(head, npix, Y, X))
            m = rot_sample(plane, (long)N * Nz, Nz, N, Xf, (float)Y - NewAxis, sn, cs, OldAxis);
        S = (float)((double)S + 0.5 * (double)m);
        o[(long)Y * cols] = (float)exp((double)(-S * atnFac));
        S = (float)((double)S + 0.5 * (double)m);
    }
}

// Launch shape of the per-pixel kernels: a block is 64 consecutive slices x
// 4 pixels of one row; grid = (slice chunks x pixel quads, rows, batch).  All
// index arithmetic stays in 32 bits and per-lane divisions are avoided.
constexpr int PIX_Q = 4;

struct PixIdx {
    int z, X, Y, b;
    bool ok;
};

__device__ __forceinline__ PixIdx pix_index(int Nz, int width)
{
    const int nzb = (Nz + 63) >> 6;
    const int zb = blockIdx.x % nzb, xq = blockIdx.x / nzb;
    PixIdx p;
    p.z = zb * 64 + (threadIdx.x & 63);
    // one pixel per wave: held as a scalar
    p.X = __builtin_amdgcn_readfirstlane(xq * PIX_Q + (int)(threadIdx.x >> 6));
    p.Y = blockIdx.y;
    p.b = blockIdx.z;
    p.ok = p.z < Nz && p.X < width;
    return p;
}

static dim3 pix_grid(int Nz, int width, int rows, int B)
{
    return dim3(((Nz + 63) / 64) * ((width + PIX_Q - 1) / PIX_Q), rows, B);
}

// ---------------------------------------------------------------------------
// k_att: rotated estimate times attenuation factor for the batch's views.
// One thread per (view, bin, slice) and ATT_ROWS consecutive rows, slices
// fastest: the rows' independent gathers and factor loads are in flight
// together.  The rows' circle extents are read once per wave into lanes
// 0..ATT_ROWS-1 and taken from there as scalars.
// ---------------------------------------------------------------------------
constexpr int ATT_ROWS = 16;

// Per row Y0 + lane (lanes < ATT_ROWS): circle extents [head, end] and
// whether the row's projector input is stored in bf16 (att_store's rule);
// every lane of the wave must be active.
struct RowExt {
    int h, e, half;
};
__device__ __forceinline__ RowExt row_extents(const int *__restrict__ head,
                                              const int *__restrict__ npix, int Y0, int NRot,
                                              int v, const int *__restrict__ rk16,
                                              const int *__restrict__ dmode)
{
    const int lane = threadIdx.x & 63;
    RowExt r{0, -1, 0};
    if (lane < ATT_ROWS && Y0 + lane < NRot) {
        r.h = head[Y0 + lane];
        r.e = r.h + npix[Y0 + lane] - 1;
        r.half = rk16 && dmode[rk16[(long)v * NRot + Y0 + lane]] <= 1;
    }
    return r;
}

// Output of a plane of the BF16 projector's input: bf16 (the projector's own
// rounding of the value, stored in the first half of the plane) where the
// row's DRF entry is plain or splits only its weights, else FP32 (split
// data, rows added unconvolved); 'half' from row_extents.
__device__ __forceinline__ void att_store_h(float *att, long plane, long idx, float val, bool half)
{
    if (half)
        ((__bf16 *)(att + plane))[idx] = (__bf16)val;
    else
        att[plane + idx] = val;
}

static dim3 att_grid(int Nz, int width, int rows, int B)
{
    return dim3(((Nz + 63) / 64) * ((width + PIX_Q - 1) / PIX_Q),
                (rows + ATT_ROWS - 1) / ATT_ROWS, B);
}

template <bool SHEAR>
__global__ __launch_bounds__(256) void k_att(const float *__restrict__ x,
                                             const float *__restrict__ afac,
                                             float *__restrict__ att, int Nz, int N, int NRot,
                                             const int *__restrict__ bview,
                                             const float *__restrict__ fsin,
                                             const float *__restrict__ fcos,
                                             const int *__restrict__ head,
                                             const int *__restrict__ npix, ShearDev sd,
                                             const int *__restrict__ rk16 = nullptr,
                                             const int *__restrict__ dmode = nullptr)
{
    const PixIdx p = pix_index(Nz, NRot);
    if (p.X >= NRot) return;   // whole waves
    const long vol = (long)Nz * NRot * NRot;
    const int v = bview[p.b];
    const int Y0 = p.Y * ATT_ROWS;
    const RowExt ext = row_extents(head, npix, Y0, NRot, v, rk16, dmode);
    // lanes past the last slice sample slice Nz-1 and store nothing
    const int z = min(p.z, Nz - 1);
    const float NewAxis = 0.5f * (float)(NRot - 1);
    const float OldAxis = 0.5f * (float)(N - 1);
    const float sn = fsin[v], cs = fcos[v];
#pragma unroll
    for (int r = 0; r < ATT_ROWS; ++r) {
        const int Y = Y0 + r;
        if (Y >= NRot) break;
        const int i = (Y * NRot + p.X) * Nz + z;
        float val = 0.0f;
        // outside the circle the value is +0 whatever the factor: not read
        const bool inside = SHEAR || (p.X >= lane_int(ext.h, r) && p.X <= lane_int(ext.e, r));
        if (SHEAR) {
            val = shear_sample(sd, 2 * v, x + z, (long)N * Nz, Nz, p.X, Y);
        } else if (inside) {
            val = rot_sample_w(x + z, N * Nz, Nz, N, (float)p.X - NewAxis, (float)Y - NewAxis,
                               sn, cs, OldAxis);
        }
        if (afac && inside) val = val * afac[(long)v * vol + i];
        const long plane = (long)p.b * vol + (long)Y * NRot * Nz;
        if (p.ok) att_store_h(att, plane, i - (long)Y * NRot * Nz, val, lane_int(ext.half, r));
    }
}

// ---------------------------------------------------------------------------
// k_att_src: learned-scatter source of the batch's views, the first-order
// form of the ESSE effective source (krnlprj3d.c:836-838, 903-920, 946-947):
//   att = (rot g0 + d rot g1) afac,  d = -ln(afac) / mu0
// where -ln(afac) is the half-pixel running attenuation sum the reference's
// depth loops accumulate (afac = exp(-sum), prjbck.c:212-222).  PK: g0 and
// g1 held as bf16 pairs in gp (bilinear rotation only).
// ---------------------------------------------------------------------------
template <bool SHEAR, bool PK>
__global__ __launch_bounds__(256) void k_att_src(const float *__restrict__ g0,
                                                 const float *__restrict__ g1,
                                                 const uint32_t *__restrict__ gp,
                                                 const float *__restrict__ afac,
                                                 float *__restrict__ att, int Nz, int N, int NRot,
                                                 const int *__restrict__ bview,
                                                 const float *__restrict__ fsin,
                                                 const float *__restrict__ fcos,
                                                 const int *__restrict__ head,
                                                 const int *__restrict__ npix, ShearDev sd,
                                                 float invMu0,
                                                 const int *__restrict__ rk16 = nullptr,
                                                 const int *__restrict__ dmode = nullptr)
{
    const PixIdx p = pix_index(Nz, NRot);
    if (p.X >= NRot) return;   // whole waves
    const long vol = (long)Nz * NRot * NRot;
    const int v = bview[p.b];
    const int Y0 = p.Y * ATT_ROWS;
    const RowExt ext = row_extents(head, npix, Y0, NRot, v, rk16, dmode);
    const int z = min(p.z, Nz - 1);   // lanes past the last slice store nothing
    const float NewAxis = 0.5f * (float)(NRot - 1);
    const float OldAxis = 0.5f * (float)(N - 1);
    const float sn = fsin[v], cs = fcos[v];
#pragma unroll
    for (int r = 0; r < ATT_ROWS; ++r) {
        const int Y = Y0 + r;
        if (Y >= NRot) break;
        const int i = (Y * NRot + p.X) * Nz + z;
        float v0 = 0.0f, v1 = 0.0f;
        // outside the circle the source is +0 whatever the factor: not read
        const bool inside = SHEAR || (p.X >= lane_int(ext.h, r) && p.X <= lane_int(ext.e, r));
        if (SHEAR) {
            v0 = shear_sample(sd, 2 * v, g0 + z, (long)N * Nz, Nz, p.X, Y);
            v1 = shear_sample(sd, 2 * v, g1 + z, (long)N * Nz, Nz, p.X, Y);
        } else if (inside && PK) {
            const float2 v = rot_sample_pair(gp + z, N * Nz, Nz, N, (float)p.X - NewAxis,
                                             (float)Y - NewAxis, sn, cs, OldAxis);
            v0 = v.x;
            v1 = v.y;
        } else if (inside) {
            v0 = rot_sample_w(g0 + z, N * Nz, Nz, N, (float)p.X - NewAxis, (float)Y - NewAxis,
                              sn, cs, OldAxis);
            v1 = rot_sample_w(g1 + z, N * Nz, Nz, N, (float)p.X - NewAxis, (float)Y - NewAxis,
                              sn, cs, OldAxis);
        }
        float out = 0.0f;
        if (inside) {
            const float af = afac[(long)v * vol + i];
            const float d = af > 0.0f ? -logf(af) * invMu0 : 0.0f;
            out = fmaf(d, v1, v0) * af;
        }
        const long plane = (long)p.b * vol + (long)Y * NRot * Nz;
        if (p.ok) att_store_h(att, plane, i - (long)Y * NRot * Nz, out, lane_int(ext.half, r));
    }
}

// Source check: per (view, bin, slice) column the depth sums of the model's
// source attE (its projection before the collimator blur), of its scatter
// part (attE - rotx afac) and of the difference to the learned source
// (attE - attC); per block the sums over its columns of (sum diff)^2,
// (sum scatter)^2 and (sum attE)^2 into part3[(b * blocks + block) * 3 ..].
// attC holds the projector's input as stored: with rk16, the rows
// att_store_h wrote in bf16 are read as bf16.
constexpr int CHK_MAXROWS = 2048;   // rows of a checked plane (NumRotPixs)
constexpr int CHK_RG = 8;           // row groups (waves) per block
__global__ __launch_bounds__(64 * CHK_RG) void k_scatter_check(const float *__restrict__ attE,
                                                       const float *__restrict__ attC,
                                                       const int *__restrict__ cplane,
                                                       const float *__restrict__ rotx,
                                                       const float *__restrict__ afac,
                                                       const int *__restrict__ bview, int Nz,
                                                       int NRot, double *__restrict__ part3,
                                                       const int *__restrict__ rk16 = nullptr,
                                                       const int *__restrict__ dmode = nullptr)
{
    // grid (blocks of 64 columns, checked views): view b's model source
    // attE plane b, its learned source attC plane cplane[b], sums into
    // part3.  The block's CHK_RG waves take every CHK_RG-th row of
    // the same 64 columns; their partial depth sums are added in row-group
    // order.
    __shared__ float part[3][CHK_RG][64];
    __shared__ double red[3][64];
    __shared__ unsigned char half[CHK_MAXROWS];   // per row: attC stored in bf16
    const int plane = NRot * Nz;
    const long vol = (long)plane * NRot;
    const int col = threadIdx.x & 63, rg = threadIdx.x >> 6;
    const int q = blockIdx.x * 64 + col;
    const int b = blockIdx.y;
    const int v = bview[b];
    for (int r = threadIdx.x; r < NRot; r += 64 * CHK_RG)
        half[r] = rk16 && dmode[rk16[(long)v * NRot + r]] <= 1;
    __syncthreads();
    float dsum = 0.0f, ssum = 0.0f, tsum = 0.0f;
    if (q < plane) {
        const long ce = (long)b * vol, cc = (long)cplane[b] * vol, cf = (long)v * vol;
        // CB rows of the thread's row group at a time: every load of the
        // batch is issued before the sums take them, in row order
        constexpr int CB = 8;
        for (int r0 = rg; r0 < NRot; r0 += CHK_RG * CB) {
            float e[CB], c[CB], m[CB];
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                const int r = r0 + CHK_RG * k;
                e[k] = c[k] = m[k] = 0.0f;
                if (r < NRot) {
                    const long o = (long)r * plane + q;
                    e[k] = attE[ce + o];
                    c[k] = half[r] ? (float)((const __bf16 *)(attC + cc + (long)r * plane))[q]
                                   : attC[cc + o];
                    m[k] = rotx[ce + o] * afac[cf + o];
                }
            }
#pragma unroll
            for (int k = 0; k < CB; ++k)
                if (r0 + CHK_RG * k < NRot) {
                    dsum += e[k] - c[k];
                    ssum += e[k] - m[k];
                    tsum += e[k];
                }
        }
    }
    part[0][rg][col] = dsum;
    part[1][rg][col] = ssum;
    part[2][rg][col] = tsum;
    __syncthreads();
    if (rg == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float t = part[k][0][col];
#pragma unroll
            for (int j = 1; j < CHK_RG; ++j) t += part[k][j][col];
            red[k][col] = (double)t * t;
        }
    }
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 3)
        part3[((long)b * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// Outcome of a check: the block sums of k_scatter_check added into
// acc[3j..3j+2] for each of the nchk views (nb blocks per view; thread t
// adds blocks t, t + 256, ... in order, then a fixed tree over the threads),
// and fail = 1 when the larger relative L2 error of the learned projection
// (sqrt of acc[3j] / acc[3j+2]) exceeds maxNrmse or is not a number.
__global__ __launch_bounds__(256) void k_check_outcome(const double *__restrict__ part3, int nb,
                                                       int nchk, float maxNrmse,
                                                       double *__restrict__ acc,
                                                       int *__restrict__ fail)
{
    __shared__ double red[6][256];
    const int t = threadIdx.x;
    for (int q = 0; q < 3 * nchk; ++q) {
        const int j = q / 3, k = q % 3;
        double s = 0.0;
        for (int i = t; i < nb; i += 256) s += part3[((long)j * nb + i) * 3 + k];
        red[q][t] = s;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 3 * nchk; ++q) red[q][t] += red[q][t + w];
        __syncthreads();
    }
    if (t < 3 * nchk) acc[t] = red[t][0];
    if (t != 0) return;
    float e = 0.0f;
    for (int j = 0; j < nchk; ++j) {
        const double q0 = red[3 * j][0], q2 = red[3 * j + 2][0];
        const float r = q2 > 0 ? (float)sqrt(q0 / q2) : 0.0f;
        if (r > e || r != r) e = r;
    }
    fail[0] = !(e <= maxNrmse);
}

// ---------------------------------------------------------------------------
// k_drf: per-plane DRF convolution (conv2d_ypad.c:39-108).
//   PRJ: partial projection of a chunk of rows, sum over rows of
//        DRF_row (*) att[row]                       (prjbck.c:195-251)
//   BCK: planes[row] = DRF_row (*) (ratio x afac[row]) (prjbck.c:346-393)
// Block = 64 bins x ZG slice groups (ZG = 4 or 8 waves); each thread owns
// DRF_ZT consecutive slices of one bin.  The input tile with its halo is
// staged in LDS column-major (slices contiguous per bin, column stride = 4
// mod 32 dwords so 128-bit reads of neighbouring lanes hit disjoint banks)
// and the padded PSF transposed (its columns contiguous).  For each PSF
// column a thread reads the column and its input window with 128-bit LDS
// loads and performs DRF_ZT x KY FMAs from registers.  Two blocks of eight
// waves share a CU, so one block's staging overlaps the other's FMAs.
// Rows whose plane index is -1 (distance <= 0) are added without
// convolution, as the reference skips ConvolveDrfYpad there.
// Element addressing is by strides, so the same kernel serves the engine's
// slice-fastest planes and the projector-level entry points' bin-fastest
// planes:  in (b,row,z,x) = in[b*inB + row*inRow + z*zs + x*xs],
// afac (v,row,z,x) = afac[v*afB + row*afRow + z*zs + x*xs],
// out (b,row|chunk,z,x) = out[b*outB + row*outRow + z*outZS + x*outXS].
// ---------------------------------------------------------------------------
constexpr int DRF_TX = 64;
constexpr int DRF_ZT = 16;

__host__ __device__ constexpr int drf_col_stride(int TH)
{
    // >= TH, multiple of 4 (aligned float4), and 4 mod 32 dwords
    return ((TH + 27) / 32) * 32 + 4;
}

template <bool V4>
__device__ __forceinline__ void drf_store(const DrfArgs &a, float *o, int z, const float *acc,
                                          float scale)
{
    if (V4) {
#pragma unroll
        for (int t = 0; t < DRF_ZT; t += 4)
            if (z + t < a.Nz)
                *(float4 *)(o + z + t) = make_float4(acc[t] * scale, acc[t + 1] * scale,
                                                     acc[t + 2] * scale, acc[t + 3] * scale);
    } else {
#pragma unroll
        for (int t = 0; t < DRF_ZT; ++t)
            if (z + t < a.Nz) o[(long)(z + t) * a.outZS] = acc[t] * scale;
    }
}

template <int KY, int ZG, bool BCK>
__global__ __launch_bounds__(64 * ZG) void k_drf(DrfArgs a)
{
    extern __shared__ float lds[];
    constexpr int NT = DRF_TX * ZG;
    constexpr int ZT = DRF_ZT;
    constexpr int CY = (KY - 1) / 2;
    constexpr int TZ = ZG * ZT;
    constexpr int TH = TZ + KY - 1;           // tile rows (slices)
    constexpr int CS = drf_col_stride(TH);    // LDS column stride
    constexpr int KYP = (KY + 3) & ~3;        // padded PSF column length
    constexpr int WIN = ZT + KY - 1;          // input window per thread
    constexpr int WIN4 = (WIN + 3) / 4;
    const int KX = a.KX;
    const int CX = (KX - 1) / 2;
    const int TW = DRF_TX + KX - 1;           // tile columns (bins)
    float *tile = lds;                        // [TW][CS]
    float *psfT = lds + TW * CS;              // [KX][KYP]
    const bool zfast = a.zs == 1;

    int bid = blockIdx.x;
    const int xt = bid % a.xtiles;
    bid /= a.xtiles;
    const int zt = bid % a.ztiles;
    bid /= a.ztiles;
    int b, r0, r1;
    if (BCK) {
        b = bid / a.nrows;
        r0 = bid % a.nrows;
        r1 = r0 + 1;
    } else {
        b = bid / a.nchunk;
        const int ch = bid % a.nchunk;
        r0 = ch * a.rchunk;
        r1 = min(r0 + a.rchunk, a.nrows);
    }
    const int v = a.bview[b];
    const int x0 = xt * DRF_TX, z0 = zt * TZ;
    const int tx = threadIdx.x % DRF_TX;
    const int zl0 = (threadIdx.x / DRF_TX) * ZT;
    const int x = x0 + tx;
    const int zb = z0 + zl0;

    float acc[ZT];
#pragma unroll
    for (int t = 0; t < ZT; ++t) acc[t] = 0.0f;

    for (int row = r0; row < r1; ++row) {
        const int kidx = a.rowKernel[(long)v * a.nrows + row];
        if (BCK) {
#pragma unroll
            for (int t = 0; t < ZT; ++t) acc[t] = 0.0f;
        }
        if (kidx < 0) {
            if (x < a.NRot) {
#pragma unroll
                for (int t = 0; t < ZT; ++t) {
                    const int z = zb + t;
                    if (z < a.Nz) acc[t] += drf_in(a, b, v, row, z, x);
                }
            }
        } else {
            const int pad = a.rowPad[(long)v * a.nrows + row];
            // PSF columns that are zero throughout add exact zeros: skipped,
            // with the tile columns only they would read
            const int2 cr = a.colRange ? a.colRange[kidx] : make_int2(0, KX - 1);
            const int cb = KX - 1 - cr.y, twc = DRF_TX + cr.y - cr.x;
            // stage: consecutive threads take consecutive elements of the
            // input's fastest axis (coalesced loads)
            if (zfast) {
                for (int e = threadIdx.x; e < TH * twc; e += NT) {
                    const int c = cb + e / TH, q = e % TH;
                    tile[c * CS + q] = drf_stage_val(a, b, v, row, z0 - CY + q, x0 - CX + c, pad);
                }
            } else {
                for (int e = threadIdx.x; e < TH * twc; e += NT) {
                    const int q = e / twc, c = cb + e % twc;
                    tile[c * CS + q] = drf_stage_val(a, b, v, row, z0 - CY + q, x0 - CX + c, pad);
                }
            }
            const float *kp = a.psf + (long)kidx * KX * KYP;
            for (int e = threadIdx.x; e < KX * KYP; e += NT) psfT[e] = kp[e];
            __syncthreads();
            for (int i = cr.x; i <= cr.y; ++i) {
                float pc[KYP];
                const float4 *pp = (const float4 *)(psfT + i * KYP);
#pragma unroll
                for (int j4 = 0; j4 < KYP / 4; ++j4) {
                    const float4 w = pp[j4];
                    pc[4 * j4] = w.x;
                    pc[4 * j4 + 1] = w.y;
                    pc[4 * j4 + 2] = w.z;
                    pc[4 * j4 + 3] = w.w;
                }
                const float4 *tc = (const float4 *)(tile + (tx + KX - 1 - i) * CS + zl0);
#pragma unroll
                for (int q4 = 0; q4 < WIN4; ++q4) {
                    const float4 w = tc[q4];
                    const float wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int q = 4 * q4 + e;
                        if (q >= WIN) continue;
#pragma unroll
                        for (int t = 0; t < ZT; ++t) {
                            const int j = t + KY - 1 - q;
                            if (j >= 0 && j < KY) acc[t] = fmaf(pc[j], wv[e], acc[t]);
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (BCK && x < a.NRot) {
            float *o = a.out + (long)b * a.outB + (long)row * a.outRow + (long)x * a.outXS;
            if (a.vec4)
                drf_store<true>(a, o, zb, acc, a.scale);
            else
                drf_store<false>(a, o, zb, acc, a.scale);
        }
    }
    if (!BCK && x < a.NRot) {
        const int ch = r0 / a.rchunk;
        float *o = a.out + (long)b * a.outB + (long)ch * a.outRow + (long)x * a.outXS;
        if (a.vec4)
            drf_store<true>(a, o, zb, acc, 1.0f);
        else
            drf_store<false>(a, o, zb, acc, 1.0f);
    }
}

// ---------------------------------------------------------------------------
// k_drf_bf16: the DRF convolution of k_drf as BF16 MFMA contractions with
// FP32 accumulation (the precision variant).  For PSF column i the
// convolution along the slices is a banded Toeplitz product:
//   out[z0+16t+m][x] += sum_k A_i[m][k] * tile[16t + k][x + CX - i],
//   A_i[m][k] = psf[m + KY - k][i]  (zero outside the band),
// tile row q holding slice z0 - CY - 1 + q,
// one v_mfma_f32_16x16x32_bf16 per 32 rows of the window (M = 16 output
// slices, N = 16 bins, K = window slices).  The A fragments of every table
// entry are built on the host; B fragments are 128-bit reads of the bf16
// tile (slices contiguous per bin).  Per table entry (distMode): plain
// BF16; the weights as hi + lo bf16 (2 MFMAs per step); or weights and data
// as hi + lo (3 MFMAs per step: hi*hi + lo*hi weights on the values' hi
// parts, then a second staging pass of the lo parts times the hi weights).
// Rows that the reference adds unconvolved (distance <= 0) use an extra
// table entry (mode 3) and are added exactly in FP32.
// Block = 4 waves, wave w owns bins [16w, 16w+16) of the 64-bin tile and
// NZT slice tiles of 16.  Engine layout only (slices fastest).
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

// K steps of 32 window rows: a tile of 16 output slices needs 16 + KY - 1
// input rows, staged from one row above the PSF's reach (z0 - CY - 1, a
// multiple of 4 for every supported KY: CY = 3 mod 4) so that groups of four
// slices are 16-byte aligned loads
__host__ __device__ constexpr int drf16_nks(int KY) { return (16 + KY + 31) / 32; }
__host__ __device__ constexpr int drf16_rows(int KY, int NZT) { return 16 * (NZT - 1) + 32 * drf16_nks(KY); }
__host__ __device__ constexpr int drf16_cs(int rows) { return ((rows + 55) / 64) * 64 + 8; }

__device__ __forceinline__ uint32_t pack_bf16(__bf16 lo, __bf16 hi)
{
    return (uint32_t)__builtin_bit_cast(uint16_t, lo) |
           ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
}

// Four consecutive slices' values as two dwords of packed bf16: the values
// rounded to bf16 (pass 0) or the bf16 rounding of their remainders v - hi
// (pass 1, the lo parts of a split).
__device__ __forceinline__ uint2 drf16_pack4(const float (&v)[4], int pass)
{
    __bf16 h[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        h[j] = (__bf16)v[j];
        if (pass) h[j] = (__bf16)(v[j] - (float)h[j]);
    }
    return make_uint2(pack_bf16(h[0], h[1]), pack_bf16(h[2], h[3]));
}

// PSF columns i0..i1 of one staged plane: Toeplitz products of the entry's A
// fragments with the B fragments of the bf16 tile, added into acc; with AS
// the lo parts of the weights add a second product (weights to ~16 bits).
// colBase = 16w + n + KX - 1 (the lane's tile column for PSF column 0),
// rowOff = ks + 8g (the lane's first window row).
template <int NZT, int NKS, bool AS>
__device__ __forceinline__ void drf16_columns(floatx4_t (&acc)[NZT], const uint4 *ap,
                                              const uint4 *apl, int i0, int i1, int nks,
                                              const uint32_t *tileH, int colBase, int CSB,
                                              int rowOff)
{
    if (nks == 1) {
        // one K step: the NZT B fragments of a column are read once
        bf16x8_t A = __builtin_bit_cast(bf16x8_t, ap[i0 * NKS * 64]), An = A;
        bf16x8_t AL = A, ALn = A;
        if (AS) AL = ALn = __builtin_bit_cast(bf16x8_t, apl[i0 * NKS * 64]);
        for (int i = i0; i <= i1; ++i) {
            if (i < i1) {
                An = __builtin_bit_cast(bf16x8_t, ap[(i + 1) * NKS * 64]);
                if (AS) ALn = __builtin_bit_cast(bf16x8_t, apl[(i + 1) * NKS * 64]);
            }
            const uint32_t *ch = tileH + ((colBase - i) * CSB + rowOff) / 2;
            bf16x8_t B[NZT];
#pragma unroll
            for (int u = 0; u < NZT; ++u)
                B[u] = __builtin_bit_cast(bf16x8_t, *(const uint4 *)(ch + 8 * u));
#pragma unroll
            for (int t = 0; t < NZT; ++t) {
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[t], acc[t], 0, 0, 0);
                if (AS) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AL, B[t], acc[t], 0, 0, 0);
            }
            A = An;
            if (AS) AL = ALn;
        }
    } else {
        // the window of tile t, step s starts at row ks + 16(t + 2s): the
        // NZT + 2(NKS-1) distinct B fragments of a column are read once
        constexpr int NB = NZT + 2 * (NKS - 1);
        bf16x8_t A[NKS], An[NKS], AL[NKS], ALn[NKS];
#pragma unroll
        for (int s = 0; s < NKS; ++s) {
            A[s] = __builtin_bit_cast(bf16x8_t, ap[(i0 * NKS + s) * 64]);
            AL[s] = AS ? __builtin_bit_cast(bf16x8_t, apl[(i0 * NKS + s) * 64]) : A[s];
        }
        for (int i = i0; i <= i1; ++i) {
            if (i < i1) {
#pragma unroll
                for (int s = 0; s < NKS; ++s) {
                    An[s] = __builtin_bit_cast(bf16x8_t, ap[((i + 1) * NKS + s) * 64]);
                    if (AS) ALn[s] = __builtin_bit_cast(bf16x8_t, apl[((i + 1) * NKS + s) * 64]);
                }
            }
            const uint32_t *ch = tileH + ((colBase - i) * CSB + rowOff) / 2;
            bf16x8_t B[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u)
                B[u] = __builtin_bit_cast(bf16x8_t, *(const uint4 *)(ch + 8 * u));
#pragma unroll
            for (int t = 0; t < NZT; ++t)
#pragma unroll
                for (int s = 0; s < NKS; ++s) {
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[s], B[t + 2 * s], acc[t], 0,
                                                                     0, 0);
                    if (AS)
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AL[s], B[t + 2 * s],
                                                                         acc[t], 0, 0, 0);
                }
#pragma unroll
            for (int s = 0; s < NKS; ++s) {
                A[s] = An[s];
                if (AS) AL[s] = ALn[s];
            }
        }
    }
}

// Launch bounds: two blocks per CU at least.  The backprojector (one plane
// per block) compiles to about 125 registers (4 waves per SIMD); the
// projector's loop over its chunk's rows to about 230 (2 waves per SIMD),
// which tighter bounds only turn into spills.
template <int KY, int NZT, bool BCK>
__global__ __launch_bounds__(256, 2) void k_drf_bf16(DrfArgs a, const uint4 *__restrict__ afr,
                                                  const uint4 *__restrict__ afrLo,
                                                  const int *__restrict__ distMode,
                                                  const int4 *__restrict__ band)
{
    extern __shared__ uint32_t ldsw[];
    constexpr int CY = (KY - 1) / 2;
    constexpr int NKS = drf16_nks(KY);
    constexpr int TZ = 16 * NZT;
    constexpr int THP = drf16_rows(KY, NZT);   // staged rows (window of the last tile)
    constexpr int CSB = drf16_cs(THP);          // bf16 column stride: 4 mod 32 dwords
    constexpr int NBAT = 6;                     // staged groups whose loads issue together
    const int KX = a.KX;
    const int CX = (KX - 1) / 2;
    const int TW = DRF_TX + KX - 1;
    uint32_t *tileH = ldsw;                     // [TW][CSB] bf16, as dword pairs

    int bid = blockIdx.x;
    const int xt = bid % a.xtiles;
    bid /= a.xtiles;
    const int zt = bid % a.ztiles;
    bid /= a.ztiles;
    int b, r0, r1;
    if (BCK) {
        b = bid / a.nrows;
        r0 = bid % a.nrows;
        r1 = r0 + 1;
    } else {
        b = bid / a.nchunk;
        const int ch = bid % a.nchunk;
        r0 = ch * a.rchunk;
        r1 = min(r0 + a.rchunk, a.nrows);
    }
    const int v = a.bview[b];
    const int x0 = xt * DRF_TX, z0 = zt * TZ;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const int x = x0 + 16 * w + n;

    floatx4_t acc[NZT];
#pragma unroll
    for (int t = 0; t < NZT; ++t) acc[t] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
    int taken = 0;

    for (int row = r0; row < r1; ++row) {
        const int kidx = a.rowKernel[(long)v * a.nrows + row];
        // per table entry: 0 plain BF16, 1 weights as hi+lo, 2 weights and
        // data as hi+lo (two passes), 3 the delta of rows added unconvolved:
        // added exactly in FP32
        const int md = distMode[kidx];
        if (a.rowSel >= 0 && (md == 2) != (a.rowSel == 1)) continue;
        ++taken;
        if (BCK) {
#pragma unroll
            for (int t = 0; t < NZT; ++t) acc[t] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
        }
        if (md == 3) {
            // the lane's outputs (slices z0+16t+4g..+3 of bin x), unconvolved
            if (x < a.NRot) {
                const long ib = (long)b * a.inB + (long)row * a.inRow + (long)x * a.xs;
                const long fb = (long)v * a.afB + (long)row * a.afRow + (long)x * a.xs;
#pragma unroll
                for (int t = 0; t < NZT; ++t)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int z = z0 + 16 * t + 4 * g + j;
                        if (z < a.Nz) {
                            float f = a.in[ib + (long)z * a.zs];
                            if (a.afac) f = f * a.afac[fb + (long)z * a.zs];
                            acc[t][j] += f;
                        }
                    }
            }
        } else {
            const int pad = a.rowPad[(long)v * a.nrows + row];
            const bool spa = md >= 1;
            // the entry's band: PSF columns [i0, i1] and window rows from ks
            // (multiple of 8) over nks K steps; outside it the A fragments are zero
            const int4 bd = band[kidx];
            const int i0 = bd.x, i1 = bd.y, ks = bd.z, nks = bd.w;
            const int srows = 16 * (NZT - 1) + 32 * nks;   // staged rows from ks
            const int c0 = KX - 1 - i1, nc = DRF_TX + i1 - i0;   // staged columns
            const int sp4 = srows / 4, total = sp4 * nc;
            const int dcc = 256 / sp4, dqi = 256 - dcc * sp4;
            const int zb = z0 - CY - 1;
            // without an axial pad (and in-range bins) a value is the input
            // times the factor, or zero outside the slices
            const bool plain = pad <= 0 || a.avg <= 0;
            // the projector's input is +0 outside the row's circle: not read
            int gx0 = 0, gx1 = a.NRot;
            if (!BCK && a.circHead) {
                gx0 = a.circHead[row];
                gx1 = gx0 + a.circNpix[row];
            }
            const long ibase = (long)b * a.inB + (long)row * a.inRow;
            const long fbase = (long)v * a.afB + (long)row * a.afRow;
            const uint4 *ap = afr + (long)kidx * KX * NKS * 64 + lane;
            const uint4 *apl = afrLo + (long)kidx * KX * NKS * 64 + lane;
            const int colBase = 16 * w + n + KX - 1, rowOff = ks + 8 * g;
            // data split (mode 2): a second pass stages the lo parts of the
            // values in the same tile and adds their product with the hi
            // weights (hi*hi + lo*hi weights in the first pass): one tile of
            // LDS per block in every mode
            const int passes = md == 2 ? 2 : 1;
            for (int ph = 0; ph < passes; ++ph) {
                // stage groups of four consecutive slices as packed bf16,
                // window row q <-> slice zb + q (a multiple of 4: one aligned
                // float4 per group, wholly inside or outside [0, Nz) as
                // Nz % 4 == 0); element e = (column cc, group qi) advances by
                // 256 per step with (cc, qi) updated incrementally
                int cc = (int)threadIdx.x / sp4, qi = (int)threadIdx.x - cc * sp4;
                if (!BCK && a.inBf16 && md <= 1 && plain) {
                    // input already in bf16 (the producer's rounding of the
                    // same value): four slices per 8-byte load
                    const uint16_t *inh = (const uint16_t *)(a.in + ibase);
                    for (int e0 = threadIdx.x; e0 < total; e0 += 256 * NBAT) {
                        uint2 hv[NBAT];
                        int li[NBAT];
#pragma unroll
                        for (int u = 0; u < NBAT; ++u) {
                            const int c = c0 + cc, q = ks + 4 * qi;
                            const int gx = x0 - CX + c, gz = zb + q;
                            li[u] = (c * CSB + q) / 2;
                            const bool ok = e0 + 256 * u < total && gx >= gx0 && gx < gx1 &&
                                            gz >= 0 && gz < a.Nz;
                            hv[u] = ok ? *(const uint2 *)(inh + (long)gx * a.xs + gz)
                                       : make_uint2(0, 0);
                            cc += dcc;
                            qi += dqi;
                            if (qi >= sp4) {
                                qi -= sp4;
                                ++cc;
                            }
                        }
#pragma unroll
                        for (int u = 0; u < NBAT; ++u)
                            if (e0 + 256 * u < total) *(uint2 *)(tileH + li[u]) = hv[u];
                    }
                } else if (plain) {
                    // NBAT groups per thread at a time: every load of the batch
                    // is issued before the first value is used, so a thread
                    // waits for memory once per batch instead of once per group
                    for (int e0 = threadIdx.x; e0 < total; e0 += 256 * NBAT) {
                        float4 f[NBAT], m[NBAT];
                        int li[NBAT];
#pragma unroll
                        for (int u = 0; u < NBAT; ++u) {
                            const int c = c0 + cc, q = ks + 4 * qi;
                            const int gx = x0 - CX + c, gz = zb + q;
                            li[u] = (c * CSB + q) / 2;
                            const bool ok = e0 + 256 * u < total && gx >= gx0 && gx < gx1 &&
                                            gz >= 0 && gz < a.Nz;
                            const long o = (long)gx * a.xs + gz;
                            f[u] = ok ? *(const float4 *)(a.in + ibase + o)
                                      : make_float4(0, 0, 0, 0);
                            m[u] = (ok && a.afac) ? *(const float4 *)(a.afac + fbase + o)
                                                  : make_float4(1, 1, 1, 1);
                            cc += dcc;
                            qi += dqi;
                            if (qi >= sp4) {
                                qi -= sp4;
                                ++cc;
                            }
                        }
#pragma unroll
                        for (int u = 0; u < NBAT; ++u) {
                            if (e0 + 256 * u < total) {
                                const float vv[4] = {f[u].x * m[u].x, f[u].y * m[u].y,
                                                     f[u].z * m[u].z, f[u].w * m[u].w};
                                *(uint2 *)(tileH + li[u]) = drf16_pack4(vv, ph);
                            }
                        }
                    }
                } else {
                    // axial padding: element by element through the pad rules
                    for (int e = threadIdx.x; e < total; e += 256) {
                        const int c = c0 + cc, q = ks + 4 * qi;
                        const int gx = x0 - CX + c, gz = zb + q;
                        float fv[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            fv[j] = drf_stage_val(a, b, v, row, gz + j, gx, pad);
                        *(uint2 *)(tileH + (c * CSB + q) / 2) = drf16_pack4(fv, ph);
                        cc += dcc;
                        qi += dqi;
                        if (qi >= sp4) {
                            qi -= sp4;
                            ++cc;
                        }
                    }
                }
                __syncthreads();
                if (spa && ph == 0)
                    drf16_columns<NZT, NKS, true>(acc, ap, apl, i0, i1, nks, tileH, colBase, CSB,
                                                  rowOff);
                else
                    drf16_columns<NZT, NKS, false>(acc, ap, apl, i0, i1, nks, tileH, colBase,
                                                   CSB, rowOff);
                __syncthreads();
            }
        }
        if (BCK && x < a.NRot) {
            float *o = a.out + (long)b * a.outB + (long)row * a.outRow + (long)x * a.outXS;
#pragma unroll
            for (int t = 0; t < NZT; ++t) {
                const int z = z0 + 16 * t + 4 * g;
                if (z < a.Nz)   // Nz % 4 == 0 on this path
                    *(float4 *)(o + z) = make_float4(acc[t][0] * a.scale, acc[t][1] * a.scale,
                                                     acc[t][2] * a.scale, acc[t][3] * a.scale);
            }
        }
    }
    if (!BCK && x < a.NRot && (!a.accumulate || taken > 0)) {
        const int ch = r0 / a.rchunk;
        float *o = a.out + (long)b * a.outB + (long)ch * a.outRow + (long)x * a.outXS;
#pragma unroll
        for (int t = 0; t < NZT; ++t) {
            const int z = z0 + 16 * t + 4 * g;
            if (z < a.Nz) {
                float4 r = make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
                if (a.accumulate) {
                    const float4 p = *(const float4 *)(o + z);
                    r = make_float4(p.x + r.x, p.y + r.y, p.z + r.z, p.w + r.w);
                }
                *(float4 *)(o + z) = r;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_ratio: projection -> measured/calculated ratio (osemrecon.c:727-820).
// ---------------------------------------------------------------------------
struct RatioArgs {
    int B, Nz, NRot, nchunk;
    const int *bview;
    const float *ppart;           // [B][nchunk][NRot][Nz]
    const float *prj;             // [V][NRot][Nz] measured
    const float *scat;            // [V][NRot][Nz] or null
    const float *vtsens;          // [V] ViewTime*Sensitivity (float product)
    const float *vt;              // [V]
    const uint8_t *colKeep;       // [V][NRot]
    float scale;                  // fScaleFac
    float *ratio;                 // [B][NRot][Nz]
    // Scatter-model bookkeeping (osemrecon.c:734-788)
    const float *ppart2;          // partials of the projection without scatter model
    float *prjscat;               // [V][NRot][Nz] stored scatter projection estimate
    int srfMode;                  // 0 none, 2 store estimate, 3 add stored estimate
};

// grid (plane blocks, batch): 32-bit indices within a plane
// Sum of a column's chunk partials in chunk order; the loads of eight chunks
// are issued before their adds.
__device__ __forceinline__ float chunk_sum(const float *__restrict__ pp, int nchunk, int pl)
{
    float p = 0.0f;
    int c = 0;
    for (; c + 8 <= nchunk; c += 8) {
        float t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = pp[(c + u) * pl];
#pragma unroll
        for (int u = 0; u < 8; ++u) p += t[u];
    }
    for (; c < nchunk; ++c) p += pp[c * pl];
    return p;
}

__global__ void k_ratio(RatioArgs a)
{
    const int pl = a.Nz * a.NRot;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pl) return;
    const int b = blockIdx.y;
    const long gid = (long)b * pl + i;
    const int xb = i / a.Nz;   // projections are [bin][slice]
    const int v = a.bview[b];
    const float p = chunk_sum(a.ppart + (long)b * a.nchunk * pl + i, a.nchunk, pl);
    float calc = p * a.scale;
    calc = calc * a.vtsens[v];
    if (a.srfMode == 3) {
        calc = calc + a.prjscat[(long)v * pl + i];
    } else if (a.srfMode == 2) {
        float p2 = chunk_sum(a.ppart2 + (long)b * a.nchunk * pl + i, a.nchunk, pl);
        p2 = (p2 * a.scale) * a.vt[v];   // scaled by ViewTime only (osemrecon.c:772-773)
        a.prjscat[(long)v * pl + i] = calc - p2;
    }
    if (a.scat) calc = calc + a.scat[(long)v * pl + i];
    // osemrecon.c:802-803 hands view 0's measured data to the mask
    if (!a.colKeep[(long)v * a.NRot + xb] || a.prj[i] < 0.0f) calc = 0.0f;
    float r = ((double)calc > 1e-20) ? a.prj[(long)v * pl + i] / calc : 0.0f;
    a.ratio[gid] = r * a.vt[v];
}

// k_prj_finish: IrlGenprj's per-view output (irlgenprj.c:229-258): projection
// scaled by the primary factor, plus the scatter estimate, times
// ViewTime*Sensitivity.
__global__ void k_prj_finish(RatioArgs a, float *__restrict__ out)
{
    const long pl = (long)a.Nz * a.NRot;
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= pl * a.B) return;
    const int b = (int)(gid / pl);
    const long i = gid % pl;
    const int v = a.bview[b];
    float p = 0.0f;
    for (int c = 0; c < a.nchunk; ++c) p += a.ppart[((long)b * a.nchunk + c) * pl + i];
    float calc = p * a.scale;
    if (a.scat) calc = calc + a.scat[(long)v * pl + i];
    out[(long)v * pl + i] = calc * a.vtsens[v];
}

// ---------------------------------------------------------------------------
// k_bck: gather-form back rotation of the batch's convolved planes, in
// subset order, with the EPS clamp after every view (prjbck.c:408-426).
// One thread per voxel keeps the running sum in a register, so the
// order-dependent clamp is reproduced exactly and no atomics are needed.
// Voxels are [y][x][z] and planes [row][bin][z]: a wave's 64 slices share
// one sample position and each tap is a contiguous run.
//   NORM = false: x *= U * Ninv (osemrecon.c:850-852), U optionally stored.
//   NORM = true:  FP64 sums, Ninv = U < 1e-6 ? 0 : 1/U (osemrecon.c:317-320).
// ---------------------------------------------------------------------------
struct BckArgs {
    int B, N, Nz, NRot;
    const int *bview;
    const float *bsin, *bcos;
    const float *planes;          // [B][NRot][NRot][Nz]
    const int *head, *npix;
    float *x;                     // estimate (updated in place) or null
    const float *ninv;            // sensitivity inverse
    float *u;                     // accumulator; read when accumulate != 0
    int accumulate;               // start from u instead of 0
    int writeU;
    float *ninvOut;               // NORM: output
    ShearDev sd;                  // shear tables (back rotation in slot 2v+1)
    const int *skip;              // when *skip != 0 nothing is written (a failed check)
};

template <bool NORM, bool SHEAR>
__global__ __launch_bounds__(256) void k_bck(BckArgs a)
{
    typedef typename std::conditional<NORM, double, float>::type acc_t;
    if (a.skip && *a.skip) return;
    const PixIdx p = pix_index(a.Nz, a.N);
    if (p.X >= a.N) return;   // whole waves
    // lanes past the last slice sample slice Nz-1 and store nothing, so
    // every lane of the wave takes part in the per-view lane broadcasts
    const int z = min(p.z, a.Nz - 1);
    const long gid = ((long)p.Y * a.N + p.X) * a.Nz + z;
    const int skip = (a.NRot - a.N) / 2;
    const int Y = p.Y + skip, X = p.X + skip;
    const float Axis = 0.5f * (float)(a.NRot - 1);
    const bool inside = in_circle(a.head, a.npix, Y, X);
    const float Xf = (float)X - Axis, Yf = (float)Y - Axis;
    const long vol = (long)a.NRot * a.NRot * a.Nz;
    const int lane = threadIdx.x & 63;
    acc_t U = (!NORM && a.accumulate) ? (acc_t)a.u[gid] : (acc_t)0;
    // views in groups of 64: lane k holds view b0+k's sine and cosine
    for (int b0 = 0; b0 < a.B; b0 += 64) {
        float lsn = 0.0f, lcs = 0.0f;
        if (!SHEAR && b0 + lane < a.B) {
            const int v = a.bview[b0 + lane];
            lsn = a.bsin[v];
            lcs = a.bcos[v];
        }
        const int nb = min(64, a.B - b0);
        for (int k = 0; k < nb; ++k) {
            const int b = b0 + k;
            float val = 0.0f;
            if (SHEAR)
                val = shear_sample(a.sd, 2 * a.bview[b] + 1, a.planes + (long)b * vol + z,
                                   (long)a.NRot * a.Nz, a.Nz, X, Y);
            else if (inside)
                val = rot_sample_w(a.planes + (long)b * vol + z, a.NRot * a.Nz, a.Nz, a.NRot, Xf,
                                   Yf, lane_float(lsn, k), lane_float(lcs, k), Axis);
            U = U + (acc_t)val;
            if ((double)U < 1.0e-5) U = 0;
        }
    }
    if (!p.ok) return;
    if (NORM) {
        a.ninvOut[gid] = U < 1e-6 ? 0.0f : (float)(1.0 / (double)U);
    } else {
        const float Uf = (float)U;
        if (a.writeU) a.u[gid] = Uf;
        if (a.x) a.x[gid] = a.x[gid] * (Uf * a.ninv[gid]);
    }
}

// x *= U * Ninv (osemrecon.c:850-852), for the per-angle path.
__global__ void k_update(float *__restrict__ x, const float *__restrict__ u,
                         const float *__restrict__ ninv, long M)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < M) x[i] = x[i] * (u[i] * ninv[i]);
}

// Sensitivity input: measured-bin mask times view time (osemrecon.c:290-295).
__global__ void k_norm_input(int B, int Nz, int NRot, const int *__restrict__ bview,
                             const float *__restrict__ prj, const uint8_t *__restrict__ colMeasOne,
                             const float *__restrict__ vt, float *__restrict__ out)
{
    const long pl = (long)Nz * NRot;
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= pl * B) return;
    const int b = (int)(gid / pl);
    const long i = gid % pl;
    const int v = bview[b];
    const int xb = (int)(i / Nz);
    float m = (colMeasOne[(long)v * NRot + xb] && prj[(long)v * pl + i] >= 0.0f) ? 1.0f : 0.0f;
    out[gid] = m * vt[v];
}

// ---------------------------------------------------------------------------
// Engine implementation
// ---------------------------------------------------------------------------
static int div_up(long a, long b) { return (int)((a + b - 1) / b); }

static void launch_drf(hipStream_t stream, bool bck, DrfArgs args, int B, int KYb, int tile = 0);
static void launch_drf_bf16(hipStream_t stream, bool bck, DrfArgs args, int B, int KYb,
                            bool split, const uint4 *afr, const uint4 *afrLo,
                            const int *distSplit, const int4 *band, int tile = 0);

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        free();
        if (count == 0) return;
        HIP_CHECK(hipMalloc((void **)&p, count * sizeof(T)));
        n = count;
    }
    void free()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s)
    {
        if (count == 0) return;
        HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t count, hipStream_t s) const
    {
        if (count == 0) return;
        HIP_CHECK(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    ~DBuf() { free(); }
};

// Device copy of the shear tables of a set of angle slots.
struct ShearSet {
    DBuf<int> par, fl, jT, jS, jE, pS, pE;
    DBuf<float> f1, f2, g1, g2;
    int N = 0, MD = 0;

    void set(const std::vector<ShearGeom> &sg, int n, hipStream_t st)
    {
        N = n;
        const int A = (int)sg.size();
        MD = 0;
        for (auto &q : sg) MD = std::max(MD, q.mDim);
        std::vector<int> hpar(4 * A), hfl((size_t)A * N), hjT((size_t)A * MD, 0),
            hjS((size_t)A * MD, 0), hjE((size_t)A * MD, 0);
        std::vector<float> hf1((size_t)A * N), hf2((size_t)A * N), hg1((size_t)A * MD, 0.0f),
            hg2((size_t)A * MD, 0.0f);
        for (int a = 0; a < A; ++a) {
            const ShearGeom &q = sg[a];
            hpar[4 * a] = q.iAdd;
            hpar[4 * a + 1] = q.mDim;
            hpar[4 * a + 2] = q.b90;
            hpar[4 * a + 3] = q.b180;
            std::copy(q.fl.begin(), q.fl.end(), hfl.begin() + (size_t)a * N);
            std::copy(q.f1.begin(), q.f1.end(), hf1.begin() + (size_t)a * N);
            std::copy(q.f2.begin(), q.f2.end(), hf2.begin() + (size_t)a * N);
            std::copy(q.jT.begin(), q.jT.end(), hjT.begin() + (size_t)a * MD);
            std::copy(q.jS.begin(), q.jS.end(), hjS.begin() + (size_t)a * MD);
            std::copy(q.jE.begin(), q.jE.end(), hjE.begin() + (size_t)a * MD);
            std::copy(q.g1.begin(), q.g1.end(), hg1.begin() + (size_t)a * MD);
            std::copy(q.g2.begin(), q.g2.end(), hg2.begin() + (size_t)a * MD);
        }
        std::vector<int> hpS, hpE;
        shear_extents(N, hpS, hpE);
        auto upi = [&](DBuf<int> &d, const std::vector<int> &h) {
            if (d.n < h.size()) d.alloc(h.size());
            d.upload(h.data(), h.size(), st);
        };
        auto upf = [&](DBuf<float> &d, const std::vector<float> &h) {
            if (d.n < h.size()) d.alloc(h.size());
            d.upload(h.data(), h.size(), st);
        };
        upi(par, hpar);
        upi(fl, hfl);
        upi(jT, hjT);
        upi(jS, hjS);
        upi(jE, hjE);
        upi(pS, hpS);
        upi(pE, hpE);
        upf(f1, hf1);
        upf(f2, hf2);
        upf(g1, hg1);
        upf(g2, hg2);
        // the host vectors die with this call
        HIP_CHECK(hipStreamSynchronize(st));
    }

    ShearDev dev() const
    {
        ShearDev d{};
        d.N = N;
        d.MD = MD;
        d.par = (const int4 *)par.p;
        d.fl = fl.p;
        d.f1 = f1.p;
        d.f2 = f2.p;
        d.jT = jT.p;
        d.jS = jS.p;
        d.jE = jE.p;
        d.g1 = g1.p;
        d.g2 = g2.p;
        d.pS = pS.p;
        d.pE = pE.p;
        return d;
    }
};

struct GraphKey {
    int s, srf, keep, learned;
    bool operator<(const GraphKey &o) const
    {
        if (s != o.s) return s < o.s;
        if (srf != o.srf) return srf < o.srf;
        if (keep != o.keep) return keep < o.keep;
        return learned < o.learned;
    }
};

struct Recon::Impl {
    Config c;
    hipStream_t stream = nullptr;
    long M = 0, P = 0, vol = 0;
    int nchunk = 1, rchunk = 1;
    DBuf<float> x, mu, prj, scat, u, ninv, afac, att, ppart, ratio, planes;
    DBuf<float> rotx, rotmu, terms, ppart2, prjscat;  // scatter model (rotmu: every view)
    DBuf<int> allViews;                                // 0..V-1
    std::unique_ptr<EsseDev> esse;
    // learned scatter
    DBuf<float> g0, g1, att2;
    // BF16 precision modes other than the always-split one (bilinear
    // rotation): g0 and g1 as bf16 pairs instead, one 4-byte gather tap for
    // both (the projector takes bf16 values of the same magnitude)
    DBuf<uint32_t> gpk;
    std::unique_ptr<CnnWork> cnnWork;   // this reconstruction's inference buffers
    DBuf<double> chk;
    // A check's scatter-model pass runs on the side stream beside the
    // projection of the learned source; its outcome (chkFail, 1 = failed)
    // gates the update on the device, and the sums and outcome come back
    // through pinned host memory (chkDone marks their arrival).
    DBuf<int> chkFail;
    DBuf<double> chkPart;         // per checked view and block of k_scatter_check: 3 sums
    hipStream_t side = nullptr;
    hipEvent_t chkFork = nullptr, chkDone = nullptr;
    struct CheckHost {
        double q[6];
        int fail;
    };
    CheckHost *chkHost = nullptr;
    CnnPhys phys;
    Recon::CnnStatus cst;
    int cnnSubsets = 0;           // modelled subsets seen
    bool cnnCheckNow = false;     // this subset is checked against the scatter model
    DBuf<int> chkViews, chkPlanes;   // [S][2] checked views, their planes in the batch
    std::vector<int> chkViewsHost;
    // captured sub-iterations, by (subset, scatter mode, keep U, learned
    // scatter in use); null until the key's second occurrence
    bool useGraphs = false;
    std::map<GraphKey, hipGraphExec_t> graphs;
    void clearGraphs()
    {
        for (auto &kv : graphs)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    DBuf<float> fsin, fcos, bsin, bcos, vt, vtsens, psf;
    DBuf<int> rowKernel, rowPad, subsets, head, npix, rowNone;
    // BF16 precision variant: Toeplitz A fragments of every table entry
    DBuf<uint4> afrHi, afrLo;
    DBuf<int> distSplit;          // per table entry: hi+lo split
    DBuf<int4> drfBand;           // per table entry: kept PSF columns, window shift, K steps
    DBuf<int2> colRange;          // per table entry: nonzero PSF columns (FP32 kernel)
    DBuf<int> rowKernel16;        // rowKernel with -1 replaced by the delta entry nDist
    DBuf<int> rowKernel16B;       // the same for the backprojector's rows
    DBuf<int> rowKernelB, rowPadB;   // backprojector's table rows (its own table, or copies)
    bool bf16 = false, bf16Split = false;   // bf16Split: some entry splits its data
    int nSplit = 0, nWsplit = 0, nUsed = 0; // entries with data / weights split, used entries
    std::vector<int> bf16Modes;             // per table entry (the delta last): 0, 1, 2
    std::vector<float> bf16Err;             // measured errors: plain, weights split, per entry
    // three-shear rotation tables, slot 2v (forward) and 2v+1 (back)
    ShearSet shset;
    ShearDev shear() const { return shset.dev(); }
    DBuf<uint8_t> colMeasOne, colKeep;
    std::vector<int> hhead, hnpix;
    const int *rowKernelNone() const { return rowNone.p; }

    bool timing = false;
    struct Stage {
        std::vector<hipEvent_t> ev;  // start/stop pairs awaiting readout
        double ms = 0;
        int launches = 0;
    };
    std::map<std::string, Stage> stages;
    std::vector<std::string> order;

    template <typename F>
    void launch(const std::string &name, F &&f)
    {
        if (!timing) {
            f();
            HIP_CHECK(hipGetLastError());
            return;
        }
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipEventRecord(e0, stream));
        f();
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipEventRecord(e1, stream));
        auto it = stages.find(name);
        if (it == stages.end()) {
            order.push_back(name);
            it = stages.emplace(name, Stage{}).first;
        }
        it->second.ev.push_back(e0);
        it->second.ev.push_back(e1);
    }

    void collect()
    {
        HIP_CHECK(hipStreamSynchronize(stream));
        for (auto &kv : stages) {
            auto &st = kv.second;
            for (size_t k = 0; k + 1 < st.ev.size(); k += 2) {
                float ms = 0;
                HIP_CHECK(hipEventElapsedTime(&ms, st.ev[k], st.ev[k + 1]));
                st.ms += ms;
                st.launches++;
                (void)hipEventDestroy(st.ev[k]);
                (void)hipEventDestroy(st.ev[k + 1]);
            }
            st.ev.clear();
        }
    }

    int ky_bucket() const
    {
        static const int sizes[] = {7, 15, 23, 31, 47, 63};
        for (int s : sizes)
            if (c.drf.KY <= s) return s;
        fail(ETYPE_ILLEGAL_VALUE, "DRF axial size %d exceeds the supported 63", c.drf.KY);
    }

    void run_att(const float *afac, int B, const int *bview, const float *src = nullptr,
                 float *dst = nullptr, bool half = false)
    {
        const Geometry &g = c.g;
        const int *rk = half ? rowKernel16.p : nullptr;
        const int *dm = half ? distSplit.p : nullptr;
        if (c.shears)
            hipLaunchKernelGGL(k_att<true>, att_grid(g.Nz, g.NRot, g.NRot, B), dim3(256), 0,
                               stream, src ? src : x.p, afac, dst ? dst : att.p, g.Nz, g.N,
                               g.NRot, bview, fsin.p, fcos.p, head.p, npix.p, shear(), rk, dm);
        else
            hipLaunchKernelGGL(k_att<false>, att_grid(g.Nz, g.NRot, g.NRot, B), dim3(256), 0,
                               stream, src ? src : x.p, afac, dst ? dst : att.p, g.Nz, g.N,
                               g.NRot, bview, fsin.p, fcos.p, head.p, npix.p, shear(), rk, dm);
    }

    void run_att_src(const float *afacp, int B, const int *bview, float *dst, bool half = false)
    {
        const Geometry &g = c.g;
        const float inv = 1.0f / phys.mu0;
        const int *rk = half ? rowKernel16.p : nullptr;
        const int *dm = half ? distSplit.p : nullptr;
        const dim3 grid = att_grid(g.Nz, g.NRot, g.NRot, B);
        if (c.shears)
            hipLaunchKernelGGL((k_att_src<true, false>), grid, dim3(256), 0, stream, g0.p, g1.p,
                               nullptr, afacp, dst, g.Nz, g.N, g.NRot, bview, fsin.p, fcos.p,
                               head.p, npix.p, shear(), inv, rk, dm);
        else if (gpk.p)
            hipLaunchKernelGGL((k_att_src<false, true>), grid, dim3(256), 0, stream, nullptr,
                               nullptr, gpk.p, afacp, dst, g.Nz, g.N, g.NRot, bview, fsin.p,
                               fcos.p, head.p, npix.p, shear(), inv, rk, dm);
        else
            hipLaunchKernelGGL((k_att_src<false, false>), grid, dim3(256), 0, stream, g0.p, g1.p,
                               nullptr, afacp, dst, g.Nz, g.N, g.NRot, bview, fsin.p, fcos.p,
                               head.p, npix.p, shear(), inv, rk, dm);
    }

    // Learned-scatter source volumes g0, g1 from the current estimate: the
    // network, or (cnnExact) the scatter model's FFT convolution of the
    // unrotated estimate.
    void run_cnn_sources()
    {
        const Geometry &g = c.g;
        if (c.cnnExact) {
            esse->convolve(x.p, 1, terms.p);
            cnn_exact_sources(terms.p, terms.p + M, x.p, mu.p, g.N, g.Nz, phys, 1.0f, g0.p, g1.p,
                              stream, gpk.p);
        } else {
            c.cnn->sources(x.p, mu.p, g.N, g.Nz, phys, 1.0f, g0.p, g1.p, *cnnWork, stream, gpk.p);
        }
    }

    // Attenuated scatter-plus-primary source of the batch's views into att
    // (prjbck.c:129-176, 219-222): rotate x and mu, ESSE, combine.
    void run_srf_att(const float *afacp, int B, const int *bview, float *dst = nullptr)
    {
        launch("rot_x", [&] { run_att(nullptr, B, bview, x.p, rotx.p); });
        if (esse->canConvolveSource(rotmu.p, afacp, bview)) {
            // the source formed from the convolution's last pass (no terms stored)
            launch("esse_conv_source", [&] {
                esse->convolveSource(rotx.p, rotmu.p, B, 1.0f, afacp, bview, dst ? dst : att.p);
            });
            return;
        }
        launch("esse_conv", [&] { esse->convolve(rotx.p, B, terms.p); });
        launch("esse_source", [&] {
            esse->source(rotx.p, rotmu.p, terms.p, B, 1.0f, afacp, bview, dst ? dst : att.p, true);
        });
    }

    // Check of the learned source against the scatter model on two of the
    // subset's views (its first and middle one; one view where the batch has
    // one), on the side stream (the engine's while kernels are timed): the
    // model's source (att2 planes 0, 1), the depth-sum errors against the
    // learned source as the projector takes it (att planes chkPlanes, bf16
    // rows with 'half'), the outcome chkFail, the copies to chkHost and
    // chkDone.  The caller makes the engine's stream wait for chkDone before
    // the update gated by chkFail and before att is written again.
    void start_check(const float *afacp, int s, int Vs, const int *bview, float maxNrmse,
                     bool half)
    {
        const Geometry &g = c.g;
        const int nchk = Vs >= 2 ? 2 : 1;
        const int *cv = nchk == 2 ? chkViews.p + 2 * s : bview;
        if (g.NRot > CHK_MAXROWS)
            fail(ETYPE_ILLEGAL_VALUE, "learned-scatter check: NumRotPixs %d above %d", g.NRot,
                 CHK_MAXROWS);
        const hipStream_t own = stream;
        const bool fork = !timing;
        if (fork) {
            HIP_CHECK(hipEventRecord(chkFork, own));
            HIP_CHECK(hipStreamWaitEvent(side, chkFork, 0));
            stream = side;
            esse->setStream(side);
        }
        auto restore = [&] {
            stream = own;
            esse->setStream(own);
        };
        try {
            run_srf_att(afacp, nchk, cv, att2.p);
            const int nb = (int)div_up(P, 64);
            launch("cnn_check", [&] {
                hipLaunchKernelGGL(k_scatter_check, dim3(nb, nchk), dim3(64 * CHK_RG), 0, stream,
                                   att2.p, att.p, chkPlanes.p, rotx.p, afacp, cv, g.Nz, g.NRot,
                                   chkPart.p, half ? rowKernel16.p : nullptr,
                                   half ? distSplit.p : nullptr);
                hipLaunchKernelGGL(k_check_outcome, dim3(1), dim3(256), 0, stream, chkPart.p, nb,
                                   nchk, maxNrmse, chk.p, chkFail.p);
            });
            HIP_CHECK(hipMemcpyAsync(chkHost->q, chk.p, 3 * nchk * sizeof(double),
                                     hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipMemcpyAsync(&chkHost->fail, chkFail.p, sizeof(int),
                                     hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipEventRecord(chkDone, stream));
        } catch (...) {
            restore();
            throw;
        }
        restore();
    }

    BckArgs bck_args(const int *bview)
    {
        const Geometry &g = c.g;
        BckArgs ba{};
        ba.B = c.Vs;
        ba.N = g.N;
        ba.Nz = g.Nz;
        ba.NRot = g.NRot;
        ba.bview = bview;
        ba.bsin = bsin.p;
        ba.bcos = bcos.p;
        ba.planes = planes.p;
        ba.head = head.p;
        ba.npix = npix.p;
        return ba;
    }

    void run_bck(bool norm, const BckArgs &ba)
    {
        const dim3 grid = pix_grid(c.g.Nz, c.g.N, c.g.N, 1), block(256);
        BckArgs b = ba;
        b.sd = shear();
        if (norm && c.shears)
            hipLaunchKernelGGL((k_bck<true, true>), grid, block, 0, stream, b);
        else if (norm)
            hipLaunchKernelGGL((k_bck<true, false>), grid, block, 0, stream, b);
        else if (c.shears)
            hipLaunchKernelGGL((k_bck<false, true>), grid, block, 0, stream, b);
        else
            hipLaunchKernelGGL((k_bck<false, false>), grid, block, 0, stream, b);
    }

    // DRF arguments of the engine's slice-fastest planes.  PRJ: input att
    // [B][row][bin][z], output ppart [B][chunk][bin][z]; BCK: input ratio
    // [B][bin][z] times afac [V][row][bin][z], output planes [B][row][bin][z].
    // k_ratio's arguments for a batch of B views (no scatter-model mode)
    RatioArgs ratio_args(const int *bview, int B) const
    {
        RatioArgs ra{};
        ra.ppart2 = ppart2.p;
        ra.prjscat = prjscat.p;
        ra.B = B;
        ra.Nz = c.g.Nz;
        ra.NRot = c.g.NRot;
        ra.nchunk = nchunk;
        ra.bview = bview;
        ra.ppart = ppart.p;
        ra.prj = prj.p;
        ra.scat = scat.p;
        ra.vtsens = vtsens.p;
        ra.vt = vt.p;
        ra.colKeep = colKeep.p;
        ra.scale = 1.0f;
        ra.ratio = ratio.p;
        return ra;
    }

    DrfArgs drf_args(bool bck, const int *bview, int model) const
    {
        const Geometry &g = c.g;
        const long pl = (long)g.Nz * g.NRot;
        DrfArgs d{};
        d.NRot = g.NRot;
        d.Nz = g.Nz;
        d.nrows = g.NRot;
        d.KX = c.drf.nDist > 0 ? c.drf.KX : 1;
        d.psf = psf.p;
        d.colRange = colRange.p;
        if (bck)
            d.rowKernel = (model & 2) ? (bf16 ? rowKernel16B.p : rowKernelB.p) : rowNone.p;
        else
            d.rowKernel = (model & 2) ? (bf16 ? rowKernel16.p : rowKernel.p) : rowNone.p;
        d.rowPad = bck ? rowPadB.p : rowPad.p;
        d.avg = c.axialAvg;
        d.bview = bview;
        d.zs = 1;
        d.xs = g.Nz;
        d.scale = 1.0f;
        d.outZS = 1;
        d.outXS = g.Nz;
        d.vec4 = g.Nz % 4 == 0;
        d.xtiles = div_up(g.NRot, DRF_TX);
        if (bck) {
            d.in = ratio.p;
            d.inB = pl;
            d.inRow = 0;
            d.afac = (model & 1) ? afac.p : nullptr;
            d.afB = vol;
            d.afRow = pl;
            d.out = planes.p;
            d.outB = vol;
            d.outRow = pl;
        } else {
            d.in = att.p;
            d.inB = vol;
            d.inRow = pl;
            d.circHead = head.p;
            d.circNpix = npix.p;
            d.out = ppart.p;
            d.outB = pl * nchunk;
            d.outRow = pl;
            d.rchunk = rchunk;
            d.nchunk = nchunk;
        }
        return d;
    }

    // Stage name of a DRF launch: the precision variant is suffixed _bf16
    // (plain BF16 operands) or _bf16x3 (hi+lo split).
    std::string drf_name(const char *base, bool bck) const
    {
        if (!bf16) return (bck ? fftBck : fftPrj) ? std::string(base) + "_fft" : base;
        const char *suf = nSplit == nUsed && nUsed > 0 ? "_bf16x3"
                          : nSplit > 0                  ? "_bf16mix"
                          : nWsplit > 0                 ? "_bf16w"
                                                        : "_bf16";
        return std::string(base) + suf;
    }

    // DRF tile shapes (0: the size heuristic): slice groups per block of the
    // FP32 kernel, or slice tiles per wave of the BF16 kernel
    int tilePrj = 0, tileBck = 0;

    // FP32 DRF by 2-D FFT (drf_fft.hip): tables of the PSF spectra, the
    // directions using it, and the projector's row chunking for each
    // algorithm (the FFT kernel keeps one CU per block, so fewer, larger
    // chunks)
    DrfFftTables fft;
    bool fftPrj = false, fftBck = false;
    int rchunkD = 1, nchunkD = 1, rchunkF = 1, nchunkF = 1;

    void set_fft(bool prjOn, bool bckOn)
    {
        fftPrj = prjOn;
        fftBck = bckOn;
        rchunk = prjOn ? rchunkF : rchunkD;
        nchunk = prjOn ? nchunkF : nchunkD;
    }

    // Device table of Vs view indices for the set-up timings (tile autotuning,
    // DRF algorithm choice): subset 0's views, or the first Vs views when the
    // reconstruction has no subset table (IrlGenprj's forward projection).
    const int *timing_views() const
    {
        return c.subsets.size() >= (size_t)c.Vs ? subsets.p : allViews.p;
    }

    // Times the spatial and the FFT kernel per direction on the first
    // subset's views (one launch to warm up, three timed) and keeps the
    // faster.  Inputs are whatever the buffers hold; only the time counts.
    void choose_drf(int Vs, int prjModel, int bckModel)
    {
        const int *tv = timing_views();
        const int KYb = ky_bucket();
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        for (int bck = 0; bck < 2; ++bck) {
            float ms[2] = {0.0f, 0.0f};
            for (int f = 0; f < 2; ++f) {
                if (bck)
                    set_fft(fftPrj, f != 0);
                else
                    set_fft(f != 0, fftBck);
                const DrfArgs da = drf_args(bck != 0, tv, bck ? bckModel : prjModel);
                run_drf(bck != 0, da, Vs, KYb);
                HIP_CHECK(hipEventRecord(e0, stream));
                for (int r = 0; r < 3; ++r) run_drf(bck != 0, da, Vs, KYb);
                HIP_CHECK(hipEventRecord(e1, stream));
                HIP_CHECK(hipEventSynchronize(e1));
                HIP_CHECK(hipEventElapsedTime(&ms[f], e0, e1));
            }
            if (bck)
                set_fft(fftPrj, ms[1] < ms[0]);
            else
                set_fft(ms[1] < ms[0], fftBck);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }

    // Per-plane BF16 error of every DRF table entry (the last one the delta of
    // the unconvolved rows), measured on the GPU: one probe plane (values in
    // [0.5, 1.5) from a fixed hash, so the mantissas spread as in measured
    // data) convolved with each entry in plain BF16 and with the weights
    // split, against the FP32 kernel (the probe itself for the delta).  The
    // cheapest mode whose relative L2 error is at most 1e-3 is kept: 0
    // plain, 1 weights hi+lo, 2 weights and data hi+lo (not measured: its
    // error is the FP32 one's).
    void measure_bf16_modes(int nd, int KYb, std::vector<int> &modes)
    {
        const Geometry &g = c.g;
        const long pl = (long)g.Nz * g.NRot;
        const int rows = nd + 1;
        std::vector<float> probe(pl);
        for (long i = 0; i < pl; ++i) {
            uint32_t h = (uint32_t)i * 2654435761u + 0x9e3779b9u;
            h ^= h >> 15;
            h *= 2246822519u;
            h ^= h >> 13;
            probe[i] = 0.5f + (float)(h >> 8) * (1.0f / 16777216.0f);
        }
        std::vector<int> rkB(rows), rkF(rows), zero(rows + 1, 0), m0(rows, 0), m1(rows, 1);
        for (int k = 0; k < rows; ++k) {
            rkB[k] = k;
            rkF[k] = k < nd ? k : -1;   // the FP32 kernel adds a -1 row unconvolved
        }
        DBuf<float> dprobe, out;
        DBuf<int> drkB, drkF, dzero, dm0, dm1;
        dprobe.alloc(pl);
        dprobe.upload(probe.data(), pl, stream);
        out.alloc(pl * rows * 3);
        drkB.alloc(rows);
        drkB.upload(rkB.data(), rows, stream);
        drkF.alloc(rows);
        drkF.upload(rkF.data(), rows, stream);
        dzero.alloc(rows + 1);
        dzero.upload(zero.data(), rows + 1, stream);
        dm0.alloc(rows);
        dm0.upload(m0.data(), rows, stream);
        dm1.alloc(rows);
        dm1.upload(m1.data(), rows, stream);
        DrfArgs d{};
        d.NRot = g.NRot;
        d.Nz = g.Nz;
        d.nrows = rows;
        d.KX = c.drf.KX;
        d.psf = psf.p;
        d.colRange = colRange.p;
        d.rowPad = dzero.p;
        d.avg = 0;
        d.bview = dzero.p;
        d.in = dprobe.p;
        d.zs = 1;
        d.xs = g.Nz;
        d.scale = 1.0f;
        d.outRow = pl;
        d.outZS = 1;
        d.outXS = g.Nz;
        d.vec4 = g.Nz % 4 == 0;
        d.xtiles = div_up(g.NRot, DRF_TX);
        d.rowKernel = drkF.p;
        d.out = out.p;
        launch_drf(stream, true, d, 1, KYb, 0);
        d.rowKernel = drkB.p;
        d.rowSel = -1;
        d.out = out.p + pl * rows;
        launch_drf_bf16(stream, true, d, 1, KYb, false, afrHi.p, afrLo.p, dm0.p, drfBand.p, 0);
        d.out = out.p + 2 * pl * rows;
        launch_drf_bf16(stream, true, d, 1, KYb, false, afrHi.p, afrLo.p, dm1.p, drfBand.p, 0);
        std::vector<float> h(pl * rows * 3);
        out.download(h.data(), h.size(), stream);
        HIP_CHECK(hipStreamSynchronize(stream));
        modes.assign(rows, 2);
        for (int k = 0; k < rows; ++k) {
            const float *ref = &h[pl * k];
            double rr = 0, e0 = 0, e1 = 0;
            for (long i = 0; i < pl; ++i) {
                const double r = ref[i];
                rr += r * r;
                e0 += (h[pl * (rows + k) + i] - r) * (h[pl * (rows + k) + i] - r);
                e1 += (h[pl * (2 * rows + k) + i] - r) * (h[pl * (2 * rows + k) + i] - r);
            }
            const float err0 = rr > 0 ? (float)std::sqrt(e0 / rr) : 0.0f;
            const float err1 = rr > 0 ? (float)std::sqrt(e1 / rr) : 0.0f;
            bf16Err[2 * k] = err0;
            bf16Err[2 * k + 1] = err1;
            modes[k] = err0 <= 1e-3f ? 0 : err1 <= 1e-3f ? 1 : 2;
        }
        modes[nd] = 3;   // the delta: rows added unconvolved, exactly in FP32
    }

    void run_drf(bool bck, const DrfArgs &args, int B, int KYb, int tile = -1)
    {
        if (tile < 0) tile = bck ? tileBck : tilePrj;
        if (!bf16 && (bck ? fftBck : fftPrj)) {
            launch_drf_fft(stream, bck, args, B, fft);
            return;
        }
        if (bf16) {
            // one launch over all rows (data-split rows run two staging
            // passes in the same tile; separate launches for them measured
            // slower, the second one's few blocks running their rows back to
            // back)
            DrfArgs a1 = args;
            a1.rowSel = -1;
            a1.accumulate = 0;
            launch_drf_bf16(stream, bck, a1, B, KYb, bf16Split, afrHi.p, afrLo.p, distSplit.p,
                            drfBand.p, tile);
        } else
            launch_drf(stream, bck, args, B, KYb, tile);
    }
};


// Round-to-nearest-even float -> bf16 bits, and back.
static uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFF + ((u >> 16) & 1);
    return (uint16_t)(u >> 16);
}

static float bf2f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// A fragments of k_drf_bf16 for every table entry, column i and K step s:
// lane l holds A_i[m = l%16][k = ks + 32s + 8(l/16) + e], e < 8, with
// A_i[m][k] = psf[m + KYb - k][i] inside the band (PSF centred in the
// KYb bucket), as bf16 hi parts and lo = bf16(value - hi).
// Each entry is first cut to the smallest rectangle of its PSF outside which
// the taps' absolute sum is at most 'cut' of the whole (edges with the
// smallest sums are dropped first; cut = 0 drops only all-zero edges); band[k] = (first, last PSF column, window
// shift ks, K steps): when the rectangle's rows fit one 32-row K step from a
// shift ks (a multiple of 8, so 128-bit LDS reads stay aligned) the kernel
// runs that single step, else all NKS steps from ks = 0.
// est[k] is the estimated relative error of a plane convolved with entry k
// in plain BF16: 1.6e-3 (the std of the relative rounding of a product of
// two bf16-rounded operands) times |psf|_2 / |psf|_1 of the kept taps, plus
// the dropped fraction.
static void build_bf16_fragments(const PaddedDrf &d, int KYb, double cut, std::vector<uint4> &hi,
                                 std::vector<uint4> &lo, std::vector<double> &est,
                                 std::vector<int4> &band)
{
    const int NKS = drf16_nks(KYb), oy = (KYb - d.KY) / 2;
    hi.assign((size_t)d.nDist * d.KX * NKS * 64, uint4{0, 0, 0, 0});
    lo = hi;
    est.assign(d.nDist, 0.0);
    band.assign(d.nDist, int4{0, d.KX - 1, 0, NKS});
    for (int k = 0; k < d.nDist; ++k) {
        const float *P = &d.psf[(size_t)k * d.KY * d.KX];
        double total = 0;
        for (int t = 0; t < d.KX * d.KY; ++t) total += fabs(P[t]);
        int r0 = 0, r1 = d.KY - 1, c0 = 0, c1 = d.KX - 1;
        double dropped = 0;
        for (;;) {
            auto rsum = [&](int j) {
                double s = 0;
                for (int i = c0; i <= c1; ++i) s += fabs(P[j * d.KX + i]);
                return s;
            };
            auto csum = [&](int i) {
                double s = 0;
                for (int j = r0; j <= r1; ++j) s += fabs(P[j * d.KX + i]);
                return s;
            };
            double best = 1e300;
            int which = -1;
            if (r1 > r0) {
                const double a = rsum(r0), b = rsum(r1);
                if (a < best) best = a, which = 0;
                if (b < best) best = b, which = 1;
            }
            if (c1 > c0) {
                const double a = csum(c0), b = csum(c1);
                if (a < best) best = a, which = 2;
                if (b < best) best = b, which = 3;
            }
            if (which < 0 || dropped + best > cut * total) break;
            dropped += best;
            if (which == 0) ++r0;
            else if (which == 1) --r1;
            else if (which == 2) ++c0;
            else --c1;
        }
        // window rows (staged from one row above the PSF's reach) of the band
        const int klo = KYb - oy - r1, khi = 15 + KYb - oy - r0;
        int ks = (klo / 8) * 8, nks = 1;
        if (khi - ks >= 32 || ks + 32 > 32 * NKS) {
            ks = 0;
            nks = NKS;
        }
        band[k] = int4{c0, c1, ks, nks};
        double l1 = 0, l2 = 0;
        for (int j = r0; j <= r1; ++j)
            for (int i = c0; i <= c1; ++i) {
                const double p = P[j * d.KX + i];
                l1 += fabs(p);
                l2 += p * p;
            }
        est[k] = (l1 > 0 ? 1.6e-3 * sqrt(l2) / l1 : 0.0) + (total > 0 ? dropped / total : 0.0);
        for (int i = 0; i < d.KX; ++i)
            for (int st = 0; st < NKS; ++st)
                for (int l = 0; l < 64; ++l) {
                    uint16_t h[8], w[8];
                    for (int e = 0; e < 8; ++e) {
                        const int m = l % 16, kk = ks + 32 * st + 8 * (l / 16) + e;
                        const int j = m + KYb - kk - oy;
                        const float val =
                            (j >= r0 && j <= r1 && i >= c0 && i <= c1) ? P[j * d.KX + i] : 0.0f;
                        h[e] = f2bf(val);
                        w[e] = f2bf(val - bf2f(h[e]));
                    }
                    const size_t o = (((size_t)k * d.KX + i) * NKS + st) * 64 + l;
                    memcpy(&hi[o], h, 16);
                    memcpy(&lo[o], w, 16);
                }
    }
}

Recon::Recon(const Config &cfg) : impl_(new Impl), cfg_(cfg)
{
    Impl &I = *impl_;
    I.c = cfg;
    const Geometry &g = cfg.g;
    if (g.N <= 0 || g.Nz <= 0 || g.NRot < g.N || g.V <= 0)
        fail(ETYPE_ILLEGAL_VALUE, "bad geometry N=%d Nz=%d NRot=%d V=%d", g.N, g.Nz, g.NRot, g.V);
    if (g.N % 2)
        fail(ETYPE_ILLEGAL_VALUE,
             "odd NumPixels (%d) is not supported: the reference rotation writes past the "
             "row end for odd sizes",
             g.N);
    if ((long)g.NRot * g.NRot * g.Nz >= INT_MAX)
        fail(ETYPE_ILLEGAL_VALUE, "volume %d x %d x %d too large (2^31 elements per view)",
             g.NRot, g.NRot, g.Nz);
    HIP_CHECK(hipStreamCreate(&I.stream));
    I.M = (long)g.N * g.Nz * g.N;
    I.P = (long)g.NRot * g.Nz;
    I.vol = (long)g.NRot * g.Nz * g.NRot;
    // Rows per projection chunk: enough chunks that a subset's projector
    // launch fills the CUs several times over.  The BF16 kernel takes twice
    // the rows (half the partial planes for the ratio step to add) where its
    // launch at the largest tile (128 slices per block) still has a block
    // for every CU.  A rule of the geometry, not a timing: the chunking
    // groups the projector's float sums, so results stay reproducible.
    I.rchunk = I.rchunkD = std::max(1, g.NRot / 32);
    if (cfg.precision > 0 && 2 * I.rchunkD < g.NRot) {
        int dev = 0, cus = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const long blocks = (long)div_up(g.NRot, DRF_TX) * div_up(g.Nz, 128) * cfg.Vs *
                            div_up(g.NRot, 2 * I.rchunkD);
        if (blocks >= cus) I.rchunk = I.rchunkD = 2 * I.rchunkD;
    }
    I.nchunk = I.nchunkD = div_up(g.NRot, I.rchunk);
    I.rchunkF = std::max(I.rchunkD, g.NRot / 16);
    I.nchunkF = div_up(g.NRot, I.rchunkF);
    const int Vs = cfg.Vs;
    I.x.alloc(I.M);
    I.u.alloc(I.M);
    I.ninv.alloc(I.M * cfg.S);
    I.prj.alloc(I.P * g.V);
    if (cfg.hasScatterEstimate) I.scat.alloc(I.P * g.V);
    if (cfg.hasAtn) {
        I.mu.alloc(I.M);
        I.afac.alloc(I.vol * g.V);
    }
    I.att.alloc(I.vol * Vs);
    if (cfg.prjModel & 4) {
        if (!cfg.hasAtn)
            fail(ETYPE_ILLEGAL_VALUE, "scatter modelling needs an attenuation map");
        I.rotx.alloc(I.vol * Vs);
        I.rotmu.alloc(I.vol * g.V);   // rotated mu of every view, computed once
        if (cfg.esse && cfg.esse->enabled) {
            if (cfg.esse->nx != g.NRot || cfg.esse->ny != g.NRot || cfg.esse->nz != g.Nz)
                fail(ETYPE_ILLEGAL_VALUE, "scatter model prepared for another volume size");
            I.esse.reset(new EsseDev(*cfg.esse, Vs, I.stream, true));
            I.esse->setFused(cfg.esseFused);
            I.terms.alloc(I.vol * Vs * cfg.esse->nterms);
            const EsseModel &em = *cfg.esse;
            I.cst.configured = cfg.cnn != nullptr || cfg.cnnExact;
            // the learned source needs the unrotated frame of the model's
            // grid, two terms, no collapse and the attenuation factors
            I.cst.active = I.cst.configured && g.N == g.NRot && g.N % 2 == 0 && g.Nz % 2 == 0 &&
                           em.nterms >= 2 && em.c == 1 && (cfg.prjModel & 1);
            if (I.cst.active) {
                I.phys.mu0 = em.mu0;
                I.phys.fMu = em.mu0 - em.muavg;
                I.phys.fDens = (1.0f / em.mu0) / em.pixsize;
                I.phys.cntr = em.cntrPixFac;
                if (cfg.precision > 0 && cfg.precision != 2 && !cfg.shears) {
                    I.gpk.alloc(I.M);
                } else {
                    I.g0.alloc(I.M);
                    I.g1.alloc(I.M);
                }
                if (cfg.cnn) I.cnnWork.reset(new CnnWork(g.N, g.Nz));
                I.att2.alloc(I.vol * Vs);   // the model's source of the checked views
                I.chk.alloc(6);
                I.chkFail.alloc(1);
                I.chkPart.alloc(2 * 3 * div_up(I.P, 64));
                HIP_CHECK(hipStreamCreate(&I.side));
                HIP_CHECK(hipEventCreateWithFlags(&I.chkFork, hipEventDisableTiming));
                HIP_CHECK(hipEventCreateWithFlags(&I.chkDone, hipEventDisableTiming));
                HIP_CHECK(hipHostMalloc((void **)&I.chkHost, sizeof(Impl::CheckHost),
                                        hipHostMallocDefault));
            }
        }
        I.ppart2.alloc(I.P * I.nchunk * Vs);
        I.prjscat.alloc(I.P * g.V);
        HIP_CHECK(hipMemsetAsync(I.prjscat.p, 0, I.P * g.V * sizeof(float), I.stream));
    }
    I.ppart.alloc(I.P * I.nchunk * Vs);
    I.ratio.alloc(I.P * Vs);
    I.planes.alloc(I.vol * Vs);

    std::vector<float> fs(g.V), fc(g.V), bs(g.V), bc(g.V), vt(g.V), vts(g.V);
    for (int v = 0; v < g.V; ++v) {
        fs[v] = cfg.views[v].fsin;
        fc[v] = cfg.views[v].fcos;
        bs[v] = cfg.views[v].bsin;
        bc[v] = cfg.views[v].bcos;
        vt[v] = cfg.views[v].viewtime;
        vts[v] = cfg.views[v].viewtime * cfg.views[v].sensitivity;
    }
    auto up = [&](DBuf<float> &d, const std::vector<float> &h) {
        d.alloc(h.size());
        d.upload(h.data(), h.size(), I.stream);
    };
    up(I.fsin, fs);
    up(I.fcos, fc);
    up(I.bsin, bs);
    up(I.bcos, bc);
    up(I.vt, vt);
    up(I.vtsens, vts);
    if (cfg.drf.nDist > 0) {
        // Device copy centred in the kernel's axial bucket size and stored
        // column-major ([dist][x][y], y padded to a multiple of 4).
        const int KY = cfg.drf.KY, KX = cfg.drf.KX, KYb = I.ky_bucket(), oy = (KYb - KY) / 2;
        const int KYP = (KYb + 3) & ~3;
        std::vector<float> padded((size_t)cfg.drf.nDist * KX * KYP, 0.0f);
        for (int d = 0; d < cfg.drf.nDist; ++d)
            for (int j = 0; j < KY; ++j)
                for (int i = 0; i < KX; ++i)
                    padded[((size_t)d * KX + i) * KYP + j + oy] =
                        cfg.drf.psf[((size_t)d * KY + j) * KX + i];
        up(I.psf, padded);
        // first and last column of each entry holding a nonzero tap
        std::vector<int2> cr(cfg.drf.nDist);
        for (int d = 0; d < cfg.drf.nDist; ++d) {
            int lo = KX, hi = -1;
            for (int i = 0; i < KX; ++i)
                for (int j = 0; j < KY; ++j)
                    if (cfg.drf.psf[((size_t)d * KY + j) * KX + i] != 0.0f) {
                        lo = std::min(lo, i);
                        hi = std::max(hi, i);
                    }
            cr[d] = hi < 0 ? make_int2(0, -1) : make_int2(lo, hi);
        }
        I.colRange.alloc(cr.size());
        I.colRange.upload(cr.data(), cr.size(), I.stream);
        HIP_CHECK(hipStreamSynchronize(I.stream));   // cr dies with this block
        if (cfg.precision == 0 && cfg.drfFft > 0) {
            const int P = drf_fft_size(g.NRot, g.Nz, KX, KYb);
            if (P > 0) {
                // real, even spectra when every PSF is its own mirror image in
                // both axes (exact comparison: the products are then equal)
                bool sym = true;
                for (int d = 0; d < cfg.drf.nDist && sym; ++d)
                    for (int j = 0; j < KY && sym; ++j)
                        for (int i = 0; i < KX && sym; ++i) {
                            const float *q = &cfg.drf.psf[(size_t)d * KY * KX];
                            const float v0 = q[(size_t)j * KX + i];
                            sym = v0 == q[(size_t)j * KX + (KX - 1 - i)] &&
                                  v0 == q[(size_t)(KY - 1 - j) * KX + i];
                        }
                I.fft.build(I.stream, P, I.psf.p, cfg.drf.nDist, KX, KYb, KYP, sym);
                if (cfg.drfFft == 1) I.set_fft(true, true);
            }
        }
        if (cfg.precision > 0) {
            if (g.Nz % 4)
                fail(ETYPE_ILLEGAL_VALUE, "the BF16 DRF path needs NumSlices %% 4 == 0 (%d)", g.Nz);
            // the table plus a centred delta (entry nDist) for unconvolved rows
            PaddedDrf d = cfg.drf;
            const int nd = d.nDist;
            d.psf.resize((size_t)(nd + 1) * d.KX * d.KY, 0.0f);
            d.psf[((size_t)nd * d.KY + (d.KY - 1) / 2) * d.KX + (d.KX - 1) / 2] = 1.0f;
            d.nDist = nd + 1;
            std::vector<uint4> hi, lo;
            std::vector<double> est;
            std::vector<int4> band;
            // the always-split mode keeps every nonzero tap; the others drop
            // PSF edges holding at most 1e-4 of the plane's weight
            build_bf16_fragments(d, KYb, cfg.precision == 2 ? 0.0 : 1e-4, hi, lo, est, band);
            I.drfBand.alloc(band.size());
            I.drfBand.upload(band.data(), band.size(), I.stream);
            std::vector<int> split(nd + 1, 0), used(nd + 1, 0), rk16 = cfg.rowKernel,
                                               rk16B = cfg.rowKernelBck;
            if (rk16.empty()) rk16.assign((size_t)g.V * g.NRot, -1);
            if (rk16B.empty()) rk16B = rk16;
            for (auto *v : {&rk16, &rk16B})
                for (auto &r : *v) {
                    if (r < 0) r = nd;
                    used[r] = 1;
                }
            I.bf16 = true;
            I.rowKernel16.alloc(rk16.size());
            I.rowKernel16.upload(rk16.data(), rk16.size(), I.stream);
            I.rowKernel16B.alloc(rk16B.size());
            I.rowKernel16B.upload(rk16B.data(), rk16B.size(), I.stream);
            I.afrHi.alloc(hi.size());
            I.afrHi.upload(hi.data(), hi.size(), I.stream);
            I.afrLo.alloc(lo.size());
            I.afrLo.upload(lo.data(), lo.size(), I.stream);
            // per entry: 0 plain BF16, 1 weights as hi+lo, 2 weights and data
            // as hi+lo.  Mode 1 (the per-plane choice) measures each entry's
            // error on a probe plane and takes the cheapest mode within 1e-3.
            I.bf16Err.assign(2 * (nd + 1), 0.0f);
            if (cfg.precision == 1)
                I.measure_bf16_modes(nd, KYb, split);
            else
                split.assign(nd + 1, cfg.precision == 2 ? 2 : 0);
            split[nd] = 3;   // unconvolved rows: exact FP32 in every mode
            I.nSplit = I.nWsplit = I.nUsed = 0;
            for (int k = 0; k < nd; ++k) {
                I.nUsed += used[k];
                I.nSplit += used[k] && split[k] == 2;
                I.nWsplit += used[k] && split[k] >= 1;
            }
            I.bf16Split = I.nSplit > 0;
            I.bf16Modes = split;
            I.distSplit.alloc(split.size());
            I.distSplit.upload(split.data(), split.size(), I.stream);
            HIP_CHECK(hipStreamSynchronize(I.stream));   // host vectors die here
        }
    }
    std::vector<int> rk = cfg.rowKernel, rp = cfg.rowPad;
    if (rk.empty()) rk.assign((size_t)g.V * g.NRot, -1);
    if (rp.empty()) rp.assign((size_t)g.V * g.NRot, 0);
    I.rowKernel.alloc(rk.size());
    I.rowKernel.upload(rk.data(), rk.size(), I.stream);
    I.rowPad.alloc(rp.size());
    I.rowPad.upload(rp.data(), rp.size(), I.stream);
    {
        const std::vector<int> &rkB = cfg.rowKernelBck.empty() ? rk : cfg.rowKernelBck;
        const std::vector<int> &rpB = cfg.rowPadBck.empty() ? rp : cfg.rowPadBck;
        I.rowKernelB.alloc(rkB.size());
        I.rowKernelB.upload(rkB.data(), rkB.size(), I.stream);
        I.rowPadB.alloc(rpB.size());
        I.rowPadB.upload(rpB.data(), rpB.size(), I.stream);
    }
    std::vector<int> none((size_t)g.V * g.NRot, -1);
    I.rowNone.alloc(none.size());
    I.rowNone.upload(none.data(), none.size(), I.stream);
    I.subsets.alloc(cfg.subsets.size());
    I.subsets.upload(cfg.subsets.data(), cfg.subsets.size(), I.stream);
    {
        // learned-scatter check views of every subset (its view 0 and Vs/2)
        // and their planes in the subset's batch
        const int S = cfg.Vs > 0 ? (int)(cfg.subsets.size() / cfg.Vs) : 0;
        I.chkViewsHost.resize(2 * (size_t)std::max(S, 1), 0);
        for (int k = 0; k < S; ++k) {
            I.chkViewsHost[2 * k] = cfg.subsets[(size_t)k * cfg.Vs];
            I.chkViewsHost[2 * k + 1] = cfg.subsets[(size_t)k * cfg.Vs + cfg.Vs / 2];
        }
        I.chkViews.alloc(I.chkViewsHost.size());
        I.chkViews.upload(I.chkViewsHost.data(), I.chkViewsHost.size(), I.stream);
        const int planes[2] = {0, cfg.Vs / 2};
        I.chkPlanes.alloc(2);
        I.chkPlanes.upload(planes, 2, I.stream);
        HIP_CHECK(hipStreamSynchronize(I.stream));   // 'planes' dies here
    }
    {
        std::vector<int> iota(g.V);
        for (int v = 0; v < g.V; ++v) iota[v] = v;
        I.allViews.alloc(g.V);
        I.allViews.upload(iota.data(), g.V, I.stream);
        HIP_CHECK(hipStreamSynchronize(I.stream));
    }
    if (cfg.shears) {
        if (g.N != g.NRot)
            fail(ETYPE_ILLEGAL_VALUE, "the shear rotation needs NumRotPixs == NumPixels");
        std::vector<ShearGeom> sg;
        for (int v = 0; v < g.V; ++v) {
            sg.push_back(shear_geometry(cfg.views[v].fang, g.N));
            sg.push_back(shear_geometry(cfg.views[v].bang, g.N));
        }
        I.shset.set(sg, g.N, I.stream);
    }
    circle_rows(g.NRot, I.hhead, I.hnpix);
    I.head.alloc(g.NRot);
    I.head.upload(I.hhead.data(), g.NRot, I.stream);
    I.npix.alloc(g.NRot);
    I.npix.upload(I.hnpix.data(), g.NRot, I.stream);
    I.colMeasOne.alloc(cfg.colMeasOne.size());
    I.colMeasOne.upload(cfg.colMeasOne.data(), cfg.colMeasOne.size(), I.stream);
    I.colKeep.alloc(cfg.colKeep.size());
    I.colKeep.upload(cfg.colKeep.data(), cfg.colKeep.size(), I.stream);
    HIP_CHECK(hipStreamSynchronize(I.stream));
}

Recon::~Recon()
{
    if (impl_) {
        impl_->clearGraphs();
        if (impl_->stream) {
            (void)hipStreamSynchronize(impl_->stream);
            for (auto &kv : impl_->stages)
                for (auto e : kv.second.ev) (void)hipEventDestroy(e);
            (void)hipStreamDestroy(impl_->stream);
        }
        if (impl_->side) {
            (void)hipStreamSynchronize(impl_->side);
            (void)hipStreamDestroy(impl_->side);
        }
        if (impl_->chkFork) (void)hipEventDestroy(impl_->chkFork);
        if (impl_->chkDone) (void)hipEventDestroy(impl_->chkDone);
        if (impl_->chkHost) (void)hipHostFree(impl_->chkHost);
        delete impl_;
    }
}

// Host-side layout conversions of the engine boundary: volumes [z][y][x] <->
// device [y][x][z], projections [view][slice][bin] <-> device [view][bin][slice].
static std::vector<float> vol_to_dev(const float *h, int Nz, int N)
{
    std::vector<float> d((size_t)N * N * Nz);
    for (int z = 0; z < Nz; ++z)
        for (int y = 0; y < N; ++y) {
            const float *src = h + ((size_t)z * N + y) * N;
            float *dst = d.data() + (size_t)y * N * Nz + z;
            for (int x = 0; x < N; ++x) dst[(size_t)x * Nz] = src[x];
        }
    return d;
}

static void vol_from_dev(const float *d, float *h, int Nz, int N)
{
    for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
            const float *src = d + ((size_t)y * N + x) * Nz;
            for (int z = 0; z < Nz; ++z) h[((size_t)z * N + y) * N + x] = src[z];
        }
}

static std::vector<float> prj_to_dev(const float *h, int V, int Nz, int NRot)
{
    std::vector<float> d((size_t)V * Nz * NRot);
    for (int v = 0; v < V; ++v)
        for (int z = 0; z < Nz; ++z)
            for (int x = 0; x < NRot; ++x)
                d[((size_t)v * NRot + x) * Nz + z] = h[((size_t)v * Nz + z) * NRot + x];
    return d;
}

static void prj_from_dev(const float *d, float *h, int V, int Nz, int NRot)
{
    for (int v = 0; v < V; ++v)
        for (int x = 0; x < NRot; ++x)
            for (int z = 0; z < Nz; ++z)
                h[((size_t)v * Nz + z) * NRot + x] = d[((size_t)v * NRot + x) * Nz + z];
}

void Recon::uploadEstimate(const float *h)
{
    const auto d = vol_to_dev(h, cfg_.g.Nz, cfg_.g.N);
    impl_->x.upload(d.data(), impl_->M, impl_->stream);
    synchronize();
}
void Recon::uploadAtn(const float *h)
{
    if (!impl_->mu.p) return;
    const auto d = vol_to_dev(h, cfg_.g.Nz, cfg_.g.N);
    impl_->mu.upload(d.data(), impl_->M, impl_->stream);
    synchronize();
}
void Recon::uploadProjections(const float *h)
{
    const auto d = prj_to_dev(h, cfg_.g.V, cfg_.g.Nz, cfg_.g.NRot);
    impl_->prj.upload(d.data(), impl_->P * cfg_.g.V, impl_->stream);
    synchronize();
}
void Recon::uploadScatterEstimate(const float *h)
{
    if (!impl_->scat.p) return;
    const auto d = prj_to_dev(h, cfg_.g.V, cfg_.g.Nz, cfg_.g.NRot);
    impl_->scat.upload(d.data(), impl_->P * cfg_.g.V, impl_->stream);
    synchronize();
}
void Recon::uploadUpdate(const float *h)
{
    const auto d = vol_to_dev(h, cfg_.g.Nz, cfg_.g.N);
    impl_->u.upload(d.data(), impl_->M, impl_->stream);
    synchronize();
}
void Recon::uploadNorm(int s, const float *h)
{
    const auto d = vol_to_dev(h, cfg_.g.Nz, cfg_.g.N);
    HIP_CHECK(hipMemcpyAsync(impl_->ninv.p + impl_->M * s, d.data(), impl_->M * sizeof(float),
                             hipMemcpyHostToDevice, impl_->stream));
    synchronize();
}
void Recon::downloadEstimate(float *h) const
{
    std::vector<float> d(impl_->M);
    impl_->x.download(d.data(), impl_->M, impl_->stream);
    synchronize();
    vol_from_dev(d.data(), h, cfg_.g.Nz, cfg_.g.N);
}
void Recon::downloadUpdate(float *h) const
{
    std::vector<float> d(impl_->M);
    impl_->u.download(d.data(), impl_->M, impl_->stream);
    synchronize();
    vol_from_dev(d.data(), h, cfg_.g.Nz, cfg_.g.N);
}
void Recon::downloadNorm(int s, float *h) const
{
    std::vector<float> d(impl_->M);
    HIP_CHECK(hipMemcpyAsync(d.data(), impl_->ninv.p + impl_->M * s, impl_->M * sizeof(float),
                             hipMemcpyDeviceToHost, impl_->stream));
    synchronize();
    vol_from_dev(d.data(), h, cfg_.g.Nz, cfg_.g.N);
}
void Recon::downloadScatterProjections(float *h) const
{
    if (!impl_->prjscat.p) return;
    std::vector<float> d(impl_->P * cfg_.g.V);
    impl_->prjscat.download(d.data(), impl_->P * cfg_.g.V, impl_->stream);
    synchronize();
    prj_from_dev(d.data(), h, cfg_.g.V, cfg_.g.Nz, cfg_.g.NRot);
}
void Recon::synchronize() const { HIP_CHECK(hipStreamSynchronize(impl_->stream)); }

Recon::CnnStatus Recon::cnnStatus() const { return impl_->cst; }
void Recon::setTiming(bool on) { impl_->timing = on; }
void Recon::resetTimes()
{
    impl_->collect();
    impl_->stages.clear();
    impl_->order.clear();
}
std::vector<Recon::StageTime> Recon::stageTimes() const
{
    impl_->collect();
    std::vector<StageTime> out;
    for (auto &n : impl_->order) {
        auto &st = impl_->stages[n];
        out.push_back(StageTime{n, st.ms, st.launches});
    }
    return out;
}

void Recon::precompute(bool norms)
{
    Impl &I = *impl_;
    const Geometry &g = cfg_.g;
    const int threads = 256;
    if (cfg_.autotune && !I.timing) autotuneTiles();
    if (cfg_.drfFft == 2 && I.fft.P > 0) I.choose_drf(cfg_.Vs, cfg_.prjModel, cfg_.bckModel);
    if (cfg_.hasAtn) {
        const long n = (long)g.V * g.Nz * g.NRot;
        I.launch("afac", [&] {
            if (cfg_.shears)
                hipLaunchKernelGGL(k_afac<true>, dim3(div_up(n, threads)), dim3(threads), 0,
                                   I.stream, I.mu.p, I.afac.p, g.V, g.Nz, g.N, g.NRot, I.fsin.p,
                                   I.fcos.p, 1.0f, I.head.p, I.npix.p, I.shear());
            else
                hipLaunchKernelGGL(k_afac<false>, dim3(div_up(n, threads)), dim3(threads), 0,
                                   I.stream, I.mu.p, I.afac.p, g.V, g.Nz, g.N, g.NRot, I.fsin.p,
                                   I.fcos.p, 1.0f, I.head.p, I.npix.p, I.shear());
        });
    }
    if (I.rotmu.p)   // the scatter model's rotated mu, fixed for a view like afac
        I.launch("rot_mu", [&] { I.run_att(nullptr, g.V, I.allViews.p, I.mu.p, I.rotmu.p); });
    // Sensitivity images: backproject mask*ViewTime for each subset.
    const int Vs = cfg_.Vs;
    const int KYb = cfg_.drf.nDist > 0 ? I.ky_bucket() : 7;
    for (int s = 0; norms && s < cfg_.S; ++s) {
        const int *bview = I.subsets.p + (long)s * Vs;
        I.launch("norm_input", [&] {
            hipLaunchKernelGGL(k_norm_input, dim3(div_up(I.P * Vs, threads)), dim3(threads), 0,
                               I.stream, Vs, g.Nz, g.NRot, bview, I.prj.p, I.colMeasOne.p,
                               I.vt.p, I.ratio.p);
        });
        const DrfArgs da = I.drf_args(true, bview, cfg_.bckModel);
        I.launch(I.drf_name("drf_bck_norm", true), [&] { I.run_drf(true, da, Vs, KYb); });
        BckArgs ba = I.bck_args(bview);
        ba.ninvOut = I.ninv.p + I.M * s;
        I.launch("norm", [&] { I.run_bck(true, ba); });
    }
    synchronize();
}

// Host bookkeeping of the learned scatter at the start of a subset: counts
// the modelled subsets and decides whether this one is checked against the
// scatter model.  Returns true when the subset must run eagerly (a check
// reads its result on the host).
bool Recon::beginSubset(SrfMode srf)
{
    Impl &I = *impl_;
    I.cnnCheckNow = false;
    const bool model = srf == SRF_MODEL || srf == SRF_MODEL_STORE;
    if (!(model && I.esse && I.cst.active)) return false;
    const int k = I.cnnSubsets++;
    I.cnnCheckNow = k == 0 || (cfg_.cnnCheckEvery > 0 && k % cfg_.cnnCheckEvery == 0);
    return I.cnnCheckNow;
}

void Recon::endSubset(SrfMode srf)
{
    Impl &I = *impl_;
    const bool model = srf == SRF_MODEL || srf == SRF_MODEL_STORE;
    if (model && I.esse && I.cst.active) I.cst.used++;
}

void Recon::subIteration(int s, bool keepUpdate, SrfMode srf)
{
    Impl &I = *impl_;
    const bool eager = beginSubset(srf);
    const bool learned = I.cst.active;
    // A captured graph of the sub-iteration replays exactly the launches of
    // the eager path (same kernels, arguments and order); it is captured on
    // the second occurrence of its key, when every lazily allocated buffer
    // exists, and never for checked subsets or while kernels are timed.
    const GraphKey key{s, (int)srf, keepUpdate ? 1 : 0, learned ? 1 : 0};
    if (I.useGraphs && !eager && !I.timing) {
        auto it = I.graphs.find(key);
        if (it != I.graphs.end()) {
            if (!it->second) {
                hipGraph_t gr = (hipGraph_t)captureSubIteration(s, keepUpdate, srf);
                hipGraphExec_t ex = nullptr;
                HIP_CHECK(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
                HIP_CHECK(hipGraphDestroy(gr));
                it->second = ex;
            }
            HIP_CHECK(hipGraphLaunch(it->second, I.stream));
            endSubset(srf);
            return;
        }
        I.graphs.emplace(key, nullptr);   // first occurrence: eager
    }
    runViews(s, 0, cfg_.Vs, srf, false, true, keepUpdate);
    endSubset(srf);
}

void Recon::prepareGraph(int s, bool keepUpdate, SrfMode srf)
{
    Impl &I = *impl_;
    if (!I.useGraphs || I.timing) return;
    const GraphKey key{s, (int)srf, keepUpdate ? 1 : 0, I.cst.active ? 1 : 0};
    auto it = I.graphs.find(key);
    if (it != I.graphs.end() && it->second) return;
    hipGraph_t gr = (hipGraph_t)captureSubIteration(s, keepUpdate, srf);
    hipGraphExec_t ex = nullptr;
    HIP_CHECK(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(gr));
    I.graphs[key] = ex;
}

// The launches of one sub-iteration captured into a graph (not executed).
// Learned-scatter bookkeeping is the caller's (beginSubset / endSubset).
void *Recon::captureSubIteration(int s, bool keepUpdate, SrfMode srf)
{
    Impl &I = *impl_;
    if (I.timing) fail(ETYPE_ILLEGAL_VALUE, "sub-iteration capture with kernel timing on");
    const bool check = I.cnnCheckNow;
    I.cnnCheckNow = false;
    HIP_CHECK(hipStreamBeginCapture(I.stream, hipStreamCaptureModeThreadLocal));
    try {
        runViews(s, 0, cfg_.Vs, srf, false, true, keepUpdate);
    } catch (...) {
        hipGraph_t dummy = nullptr;
        (void)hipStreamEndCapture(I.stream, &dummy);
        if (dummy) (void)hipGraphDestroy(dummy);
        I.cnnCheckNow = check;
        throw;
    }
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamEndCapture(I.stream, &g));
    I.cnnCheckNow = check;
    return g;
}

void Recon::recordSubIteration(int s, SrfMode srf)
{
    Impl &I = *impl_;
    if (I.timing) fail(ETYPE_ILLEGAL_VALUE, "sub-iteration capture with kernel timing on");
    const bool check = I.cnnCheckNow;
    I.cnnCheckNow = false;
    try {
        runViews(s, 0, cfg_.Vs, srf, false, true, false);
    } catch (...) {
        I.cnnCheckNow = check;
        throw;
    }
    I.cnnCheckNow = check;
}

// Tile autotuner: each candidate shape of the DRF kernels is timed with HIP
// events on the views of subset 0 (three launches after one untimed one, on
// whatever the work buffers hold: timing does not depend on the values) and
// the fastest is kept, separately for the projector and the backprojector.
void Recon::autotuneTiles()
{
    Impl &I = *impl_;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    // the scatter network's row length (slices per wave of its MFMA layers):
    // the source build timed for each, on the current estimate
    if (I.cnnWork && !cfg_.cnnExact && cfg_.cnn) {
        int best = 64;
        float bestMs = 1e30f;
        for (int rz : {64, 32, 16}) {
            I.cnnWork->rowz = rz;
            I.run_cnn_sources();
            HIP_CHECK(hipEventRecord(e0, I.stream));
            for (int r = 0; r < 3; ++r) I.run_cnn_sources();
            HIP_CHECK(hipEventRecord(e1, I.stream));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < bestMs) {
                bestMs = ms;
                best = rz;
            }
        }
        I.cnnWork->rowz = best;
    }
    const int KYb = cfg_.drf.nDist > 0 ? I.ky_bucket() : 0;
    const int *bview = I.timing_views();
    for (int bck = 0; bck < 2 && cfg_.drf.nDist > 0; ++bck) {
        DrfArgs da = I.drf_args(bck != 0, bview, bck ? cfg_.bckModel : cfg_.prjModel);
        // the projector's input as the sub-iterations store it (the values
        // are not used)
        if (!bck) da.inBf16 = I.bf16 ? 1 : 0;
        int best = 0;
        float bestMs = 1e30f;
        // BF16: 2, 4 or 8 slice tiles of 16 per wave; FP32: 4 or 8 slice
        // groups per block
        const std::vector<int> cand = I.bf16 ? std::vector<int>{2, 4, 8} : std::vector<int>{4, 8};
        for (int tile : cand) {
            I.run_drf(bck != 0, da, cfg_.Vs, KYb, tile);
            HIP_CHECK(hipEventRecord(e0, I.stream));
            for (int r = 0; r < 3; ++r) I.run_drf(bck != 0, da, cfg_.Vs, KYb, tile);
            HIP_CHECK(hipEventRecord(e1, I.stream));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < bestMs) {
                bestMs = ms;
                best = tile;
            }
        }
        (bck ? I.tileBck : I.tilePrj) = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void Recon::tiles(int *prj, int *bck) const
{
    *prj = impl_->tilePrj;
    *bck = impl_->tileBck;
}

void Recon::setTiles(int prj, int bck)
{
    Impl &I = *impl_;
    auto ok = [&](int t) { return t == 0 || t == 4 || t == 8 || (I.bf16 && t == 2); };
    if (!ok(prj) || !ok(bck))
        fail(ETYPE_ILLEGAL_VALUE, "DRF tiles %d, %d (0, 4, 8%s)", prj, bck, I.bf16 ? ", 2" : "");
    I.tilePrj = prj;
    I.tileBck = bck;
    I.clearGraphs();
}

int Recon::cnnRowLength() const { return impl_->cnnWork ? impl_->cnnWork->rowz : 0; }

int Recon::prjRowsPerChunk() const { return impl_->rchunkD; }

void Recon::drfAlgorithms(int *prj, int *bck, int *P) const
{
    *prj = impl_->fftPrj ? 1 : 0;
    *bck = impl_->fftBck ? 1 : 0;
    *P = impl_->fft.P;
}

void *Recon::streamHandle() const { return (void *)impl_->stream; }

void Recon::bf16Planes(std::vector<int> &modes, std::vector<float> &err) const
{
    modes = impl_->bf16Modes;
    err = impl_->bf16Err;
}

bool Recon::learnedScatterActive() const { return impl_->cst.active; }

void Recon::setGraphs(bool on)
{
    Impl &I = *impl_;
    I.useGraphs = on;
    if (!on) I.clearGraphs();
}

// ---------------------------------------------------------------------------
// ReconBatch
// ---------------------------------------------------------------------------
struct ReconBatch::Impl {
    std::vector<Recon *> r;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;                      // one per reconstruction
    hipEvent_t done = nullptr;
    // captured sub-iterations of reconstruction i by key (owned graphs)
    std::vector<std::map<GraphKey, hipGraph_t>> child;
    std::vector<std::set<GraphKey>> seen;            // keys run eagerly once
    // parent graphs by topology: the members and their (mode, learned) keys
    struct Parent {
        hipGraphExec_t exec = nullptr;
        std::vector<hipGraphNode_t> node;           // per member
    };
    std::map<std::vector<int>, Parent> parents;
    // one-graph mode: executables by topology (members and their keys)
    std::map<std::vector<int>, hipGraphExec_t> flat;
    hipEvent_t fork = nullptr;
    // own-stream mode: each member's instantiated graphs by key
    std::vector<std::map<GraphKey, hipGraphExec_t>> exec;
    int parent = 2;
    int launches = 0, eager = 0, replays = 0;
};

ReconBatch::ReconBatch(const std::vector<Recon *> &recons) : impl_(new Impl)
{
    Impl &B = *impl_;
    B.r = recons;
    HIP_CHECK(hipStreamCreate(&B.stream));
    B.ev.resize(recons.size());
    for (auto &e : B.ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&B.done, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
    B.child.resize(recons.size());
    B.seen.resize(recons.size());
    B.exec.resize(recons.size());
}

ReconBatch::~ReconBatch()
{
    Impl &B = *impl_;
    (void)hipStreamSynchronize(B.stream);
    for (auto &kv : B.parents)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto &kv : B.flat)
        if (kv.second) (void)hipGraphExecDestroy(kv.second);
    for (auto *r : B.r) r->synchronize();
    for (auto &m : B.child)
        for (auto &kv : m) (void)hipGraphDestroy(kv.second);
    for (auto &m : B.exec)
        for (auto &kv : m) (void)hipGraphExecDestroy(kv.second);
    for (auto e : B.ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(B.done);
    (void)hipEventDestroy(B.fork);
    (void)hipStreamDestroy(B.stream);
    delete impl_;
}

void ReconBatch::step(const std::vector<int> &subset, const std::vector<SrfMode> &srf)
{
    Impl &B = *impl_;
    const int n = (int)B.r.size();
    if ((int)subset.size() != n || (int)srf.size() != n)
        fail(ETYPE_ILLEGAL_VALUE, "batch step: %d reconstructions, %d subsets, %d modes", n,
             (int)subset.size(), (int)srf.size());
    std::vector<int> members, eager, topo;
    std::vector<hipGraph_t> graphs;
    std::vector<char> check(n, 0);
    for (int i = 0; i < n; ++i) {
        Recon &R = *B.r[i];
        check[i] = R.beginSubset(srf[i]);
        const GraphKey key{subset[i], (int)srf[i], 0, R.learnedScatterActive() ? 1 : 0};
        if (check[i] || !B.seen[i].count(key)) {
            if (!check[i]) B.seen[i].insert(key);
            eager.push_back(i);
            continue;
        }
        if (!B.parent) {
            // own stream: the member's cached graph of this key
            auto it = B.exec[i].find(key);
            if (it == B.exec[i].end()) {
                hipGraph_t g = (hipGraph_t)R.captureSubIteration(subset[i], false, srf[i]);
                hipGraphExec_t ex = nullptr;
                HIP_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
                HIP_CHECK(hipGraphDestroy(g));
                it = B.exec[i].emplace(key, ex).first;
            }
            HIP_CHECK(hipGraphLaunch(it->second, (hipStream_t)R.streamHandle()));
            R.endSubset(srf[i]);
            B.replays++;
            continue;
        }
        if (B.parent == 1) {
            auto it = B.child[i].find(key);
            if (it == B.child[i].end())
                it = B.child[i]
                         .emplace(key, (hipGraph_t)R.captureSubIteration(subset[i], false, srf[i]))
                         .first;
            graphs.push_back(it->second);
        }
        members.push_back(i);
        topo.push_back(i);
        topo.push_back(key.s);
        topo.push_back(key.srf);
        topo.push_back(key.learned);
    }
    if (!members.empty() && B.parent == 2) {
        // one graph of this combination of members and subsets: the members'
        // launches captured on their own streams as branches forked from and
        // joined to the batch's stream (the combinations recur with the
        // subset cycles; past 128 of them the cache starts again)
        auto fit = B.flat.find(topo);
        if (fit == B.flat.end()) {
            if (B.flat.size() >= 128) {
                HIP_CHECK(hipStreamSynchronize(B.stream));
                for (auto &kv : B.flat) (void)hipGraphExecDestroy(kv.second);
                B.flat.clear();
            }
            HIP_CHECK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                HIP_CHECK(hipEventRecord(B.fork, B.stream));
                for (int i : members) {
                    hipStream_t st = (hipStream_t)B.r[i]->streamHandle();
                    HIP_CHECK(hipStreamWaitEvent(st, B.fork, 0));
                    B.r[i]->recordSubIteration(subset[i], srf[i]);
                    HIP_CHECK(hipEventRecord(B.ev[i], st));
                    HIP_CHECK(hipStreamWaitEvent(B.stream, B.ev[i], 0));
                }
            } catch (...) {
                (void)hipStreamEndCapture(B.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(B.stream, &g));
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_CHECK(ie);
            fit = B.flat.emplace(topo, ex).first;
        }
        // the members' earlier work precedes the launch; their later work follows it
        for (int i : members) {
            HIP_CHECK(hipEventRecord(B.ev[i], (hipStream_t)B.r[i]->streamHandle()));
            HIP_CHECK(hipStreamWaitEvent(B.stream, B.ev[i], 0));
        }
        HIP_CHECK(hipGraphLaunch(fit->second, B.stream));
        HIP_CHECK(hipEventRecord(B.done, B.stream));
        for (int i : members) {
            HIP_CHECK(hipStreamWaitEvent((hipStream_t)B.r[i]->streamHandle(), B.done, 0));
            B.r[i]->endSubset(srf[i]);
        }
        B.launches++;
        B.replays += (int)members.size();
    } else if (!members.empty()) {
        // the parent graph of this combination of members and subsets (the
        // combinations recur with the subset cycles): one child node per
        // member, no edges.  Past a bound on the cache, an executable of the
        // same topology gets its child graphs swapped instead.
        auto pit = B.parents.find(topo);
        bool swap = false;
        if (pit == B.parents.end() && B.parents.size() >= 64) {
            for (auto q = B.parents.begin(); q != B.parents.end(); ++q) {
                const std::vector<int> &k = q->first;
                bool same = k.size() == topo.size();
                for (size_t j = 0; same && j < k.size(); j += 4)
                    same = k[j] == topo[j] && k[j + 2] == topo[j + 2] && k[j + 3] == topo[j + 3];
                if (same) {
                    Impl::Parent P = q->second;
                    B.parents.erase(q);
                    pit = B.parents.emplace(topo, P).first;
                    swap = true;
                    break;
                }
            }
        }
        bool fresh = pit == B.parents.end();
        if (swap) {
            for (size_t k = 0; k < members.size(); ++k)
                if (hipGraphExecChildGraphNodeSetParams(pit->second.exec, pit->second.node[k],
                                                        graphs[k]) != hipSuccess) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(pit->second.exec);
                    B.parents.erase(pit);
                    fresh = true;
                    break;
                }
        }
        if (fresh) {
            hipGraph_t parent;
            HIP_CHECK(hipGraphCreate(&parent, 0));
            Impl::Parent P;
            for (auto g : graphs) {
                hipGraphNode_t node;
                HIP_CHECK(hipGraphAddChildGraphNode(&node, parent, nullptr, 0, g));
                P.node.push_back(node);
            }
            HIP_CHECK(hipGraphInstantiate(&P.exec, parent, nullptr, nullptr, 0));
            HIP_CHECK(hipGraphDestroy(parent));
            pit = B.parents.emplace(topo, P).first;
        }
        // the members' earlier work precedes the launch; their later work follows it
        for (int i : members) {
            HIP_CHECK(hipEventRecord(B.ev[i], (hipStream_t)B.r[i]->streamHandle()));
            HIP_CHECK(hipStreamWaitEvent(B.stream, B.ev[i], 0));
        }
        HIP_CHECK(hipGraphLaunch(pit->second.exec, B.stream));
        HIP_CHECK(hipEventRecord(B.done, B.stream));
        for (int i : members) {
            HIP_CHECK(hipStreamWaitEvent((hipStream_t)B.r[i]->streamHandle(), B.done, 0));
            B.r[i]->endSubset(srf[i]);
        }
        B.launches++;
        B.replays += (int)members.size();
    }
    // eager members last: a checked subset reads its check on the host, which
    // waits only for its own stream
    for (int i : eager) {
        Recon &R = *B.r[i];
        R.runSubIteration(subset[i], srf[i]);
        R.endSubset(srf[i]);
        B.eager++;
    }
}

void ReconBatch::setParentGraph(int mode)
{
    if (mode < 0 || mode > 2) fail(ETYPE_ILLEGAL_VALUE, "batch launch mode %d (0, 1 or 2)", mode);
    impl_->parent = mode;
}
int ReconBatch::replays() const { return impl_->replays; }

void ReconBatch::synchronize()
{
    Impl &B = *impl_;
    HIP_CHECK(hipStreamSynchronize(B.stream));
    for (auto *r : B.r) r->synchronize();
}

int ReconBatch::graphLaunches() const { return impl_->launches; }
int ReconBatch::eagerSubIterations() const { return impl_->eager; }

void Recon::angleStep(int s, int m, SrfMode srf)
{
    // per-angle hooks: the learned-scatter bookkeeping runs with the first view
    if (m == 0) beginSubset(srf);
    runViews(s, m, 1, srf, true, false, true);
    if (m == cfg_.Vs - 1) endSubset(srf);
}

void Recon::zeroUpdate()
{
    HIP_CHECK(hipMemsetAsync(impl_->u.p, 0, impl_->M * sizeof(float), impl_->stream));
}

void Recon::finishSubset(int s)
{
    Impl &I = *impl_;
    hipLaunchKernelGGL(k_update, dim3(div_up(I.M, 256)), dim3(256), 0, I.stream, I.x.p, I.u.p,
                       I.ninv.p + I.M * s, I.M);
    HIP_CHECK(hipGetLastError());
}

// Views [first, first+Vs) of subset s: projection, ratio, backprojection into
// U (starting from the stored U when 'accumulate'), and the multiplicative
// update of x when 'update'.
void Recon::runViews(int s, int first, int Vs, SrfMode srf, bool accumulate, bool update,
                     bool keepUpdate)
{
    Impl &I = *impl_;
    const Geometry &g = cfg_.g;
    const int threads = 256;
    if (s < 0 || first < 0 || first + Vs > cfg_.Vs ||
        cfg_.subsets.size() < (size_t)(s + 1) * cfg_.Vs)
        fail(ETYPE_ILLEGAL_VALUE, "sub-iteration %d (views %d..%d) outside the subset table", s,
             first, first + Vs);
    const int *bview = I.subsets.p + (long)s * cfg_.Vs + first;
    const int KYb = cfg_.drf.nDist > 0 ? I.ky_bucket() : 7;
    const float *afac = (cfg_.prjModel & 1) ? I.afac.p : nullptr;
    const bool model = srf == SRF_MODEL || srf == SRF_MODEL_STORE;
    if (srf != SRF_NONE && !I.prjscat.p)
        fail(ETYPE_ILLEGAL_VALUE, "scatter model requested without a scatter kernel");

    // The BF16 projector's input planes may be written in bf16 (rows whose
    // DRF entry does not split its data; lossless: the projector rounds the
    // same values to bf16) by the rotation kernels; the scatter model's
    // source kernels stay FP32.
    bool attHalf = false, checking = false;
    if (model && I.esse && I.cst.active) {
        // once per subset: the source volumes of the current estimate (whether
        // the subset is also checked was decided by beginSubset)
        if (first == 0) I.launch("cnn_src", [&] { I.run_cnn_sources(); });
        attHalf = I.bf16;
        I.launch("cnn_att", [&] { I.run_att_src(afac, Vs, bview, I.att.p, attHalf); });
        if (I.cnnCheckNow && first == 0) {
            // the check runs beside the projection of the learned source and
            // gates the update: a failed check leaves x and U as they were,
            // and the subset runs again on the scatter model below
            checking = true;
            I.start_check(afac, s, Vs, bview, cfg_.cnnMaxNrmse, attHalf);
        }
    } else if (model && I.esse)
        I.run_srf_att(afac, Vs, bview);
    else {   // no scatter model (or scatFac == 0: the projected image is x)
        attHalf = I.bf16;
        I.launch("att", [&] { I.run_att(afac, Vs, bview, nullptr, nullptr, attHalf); });
    }
    DrfArgs da = I.drf_args(false, bview, cfg_.prjModel);
    da.inBf16 = attHalf ? 1 : 0;
    I.launch(I.drf_name("drf_prj", false), [&] { I.run_drf(false, da, Vs, KYb); });
    if (srf == SRF_MODEL_STORE) {
        // projection without the scatter model for the stored estimate
        // (osemrecon.c:757-786), reusing the rotated estimate; a check reads
        // att until chkDone
        if (checking) HIP_CHECK(hipStreamWaitEvent(I.stream, I.chkDone, 0));
        I.launch("att", [&] { I.run_att(afac, Vs, bview, nullptr, nullptr, I.bf16); });
        DrfArgs d2 = da;
        d2.inBf16 = I.bf16 ? 1 : 0;
        d2.out = I.ppart2.p;
        I.launch(I.drf_name("drf_prj", false), [&] { I.run_drf(false, d2, Vs, KYb); });
    }
    RatioArgs ra = I.ratio_args(bview, Vs);
    ra.srfMode = srf == SRF_MODEL_STORE ? 2 : srf == SRF_ADD_STORED ? 3 : 0;
    I.launch("ratio", [&] {
        hipLaunchKernelGGL(k_ratio, dim3(div_up(I.P, threads), Vs), dim3(threads), 0,
                           I.stream, ra);
    });
    const DrfArgs db = I.drf_args(true, bview, cfg_.bckModel);
    I.launch(I.drf_name("drf_bck", true), [&] { I.run_drf(true, db, Vs, KYb); });
    BckArgs ba = I.bck_args(bview);
    ba.B = Vs;
    ba.x = update ? I.x.p : nullptr;
    ba.ninv = I.ninv.p + I.M * s;
    ba.u = I.u.p;
    ba.accumulate = accumulate ? 1 : 0;
    ba.writeU = keepUpdate ? 1 : 0;
    if (checking) {
        HIP_CHECK(hipStreamWaitEvent(I.stream, I.chkDone, 0));
        ba.skip = I.chkFail.p;
    }
    I.launch("bck_update", [&] { I.run_bck(false, ba); });
    if (!checking) return;
    // the check's outcome (its copies arrive while the backprojection runs)
    HIP_CHECK(hipEventSynchronize(I.chkDone));
    const int nchk = Vs >= 2 ? 2 : 1;
    float e = 0.0f, es = 0.0f;   // the larger of the views; a NaN stays
    for (int j = 0; j < nchk; ++j) {
        const double *q = I.chkHost->q + 3 * j;
        const float t = q[2] > 0 ? (float)std::sqrt(q[0] / q[2]) : 0.0f;
        const float ts = q[1] > 0 ? (float)std::sqrt(q[0] / q[1]) : 0.0f;
        if (t > e || t != t) e = t;
        if (ts > es || ts != ts) es = ts;
    }
    I.cst.checks++;
    I.cst.lastNrmse = e;
    I.cst.lastScatterNrmse = es;
    if (e > I.cst.worstNrmse || e != e) I.cst.worstNrmse = e;
    if (es > I.cst.worstScatterNrmse || es != es) I.cst.worstScatterNrmse = es;
    if (I.chkHost->fail) {
        // back to the scatter model, from this subset on
        I.cst.active = false;
        runViews(s, first, Vs, srf, accumulate, update, keepUpdate);
    }
}

// Forward projection of the current estimate for every view (IrlGenprj).
void Recon::forwardProject(float primary, float *out_host, bool srf)
{
    Impl &I = *impl_;
    const Geometry &g = cfg_.g;
    const int threads = 256;
    const int Vs = cfg_.Vs;
    const int KYb = cfg_.drf.nDist > 0 ? I.ky_bucket() : 7;
    const float *afac = (cfg_.prjModel & 1) ? I.afac.p : nullptr;
    DBuf<int> order;
    std::vector<int> all(g.V);
    for (int v = 0; v < g.V; ++v) all[v] = v;
    order.alloc(g.V);
    order.upload(all.data(), g.V, I.stream);
    DBuf<float> out;
    out.alloc(I.P * g.V);
    for (int v0 = 0; v0 < g.V; v0 += Vs) {
        const int B = std::min(Vs, g.V - v0);
        const int *bview = order.p + v0;
        if (srf && I.esse) {
            // the reconstruction's scatter path (run_srf_att) with fPrimaryFac
            I.launch("rot_x", [&] { I.run_att(nullptr, B, bview, I.x.p, I.rotx.p); });
            if (I.esse->canConvolveSource(I.rotmu.p, afac, bview)) {
                I.launch("esse_conv_source", [&] {
                    I.esse->convolveSource(I.rotx.p, I.rotmu.p, B, primary, afac, bview, I.att.p);
                });
            } else {
                I.launch("esse_conv", [&] { I.esse->convolve(I.rotx.p, B, I.terms.p); });
                I.launch("esse_source", [&] {
                    I.esse->source(I.rotx.p, I.rotmu.p, I.terms.p, B, primary, afac, bview,
                                   I.att.p, true);
                });
            }
        } else {
            I.launch("att", [&] { I.run_att(afac, B, bview); });
        }
        const DrfArgs da = I.drf_args(false, bview, cfg_.prjModel);
        I.launch(I.drf_name("drf_prj", false), [&] { I.run_drf(false, da, B, KYb); });
        RatioArgs ra{};
        ra.B = B;
        ra.Nz = g.Nz;
        ra.NRot = g.NRot;
        ra.nchunk = I.nchunk;
        ra.bview = bview;
        ra.ppart = I.ppart.p;
        ra.scat = I.scat.p;
        ra.vtsens = I.vtsens.p;
        ra.scale = (srf && I.esse) ? 1.0f : primary;
        I.launch("prj_finish", [&] {
            hipLaunchKernelGGL(k_prj_finish, dim3(div_up(I.P * B, threads)), dim3(threads), 0,
                               I.stream, ra, out.p);
        });
    }
    std::vector<float> hd(I.P * g.V);
    out.download(hd.data(), I.P * g.V, I.stream);
    synchronize();
    prj_from_dev(hd.data(), out_host, g.V, g.Nz, g.NRot);
}


// ---------------------------------------------------------------------------
// Projector-level entry points (prjbck.c semantics, one view per call).
// Volumes are the reference's internal [y][z][x] order here, because the
// callers hand over host arrays in that order and the workspace contents
// (rotated estimate, rotated attenuation map) are part of the contract.
// ---------------------------------------------------------------------------

// Attenuation factors of one view from an already rotated mu:
// afac[row][z][x] = exp(-(sum_{r<row} mu + mu_row/2) * atnFac).
__global__ void k_afac_rotated(const float *__restrict__ rmu, float *__restrict__ afac, int Nz,
                               int NRot, float atnFac)
{
    const long cols = (long)Nz * NRot;
    const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float S = 0.0f;
    for (int Y = 0; Y < NRot; ++Y) {
        const float m = rmu[(long)Y * cols + c];
        S = (float)((double)S + 0.5 * (double)m);
        afac[(long)Y * cols + c] = (float)exp((double)(-S * atnFac));
        S = (float)((double)S + 0.5 * (double)m);
    }
}

__global__ void k_scale(const float *__restrict__ in, float *__restrict__ out, long n, float f)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] * f;
}

__global__ void k_sum_chunks(const float *__restrict__ ppart, float *__restrict__ out,
                             long pl, int nchunk, float scale)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= pl) return;
    float p = 0.0f;
    for (int c = 0; c < nchunk; ++c) p += ppart[(long)c * pl + i];
    out[i] = p * scale;
}

// imagematrix += crop(unrot), then the EPS clamp over the whole image
// (prjbck.c:418-426).
__global__ void k_acc_clamp(float *__restrict__ img, const float *__restrict__ unrot, int N,
                            int Nz, int NRot)
{
    const long M = (long)N * Nz * N;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int c = (int)(i % N);
    const long r = i / N;
    const int sl = (int)(r % Nz);
    const int y = (int)(r / Nz);
    const int skip = (NRot - N) / 2;
    float v = img[i] + unrot[((long)(y + skip) * Nz + sl) * NRot + c + skip];
    if ((double)v < 1.0e-5) v = 0.0f;
    img[i] = v;
}

struct PrjWs {
    int Nz = 0, N = 0, NRot = 0, model = 0;
    hipStream_t stream = nullptr;
    DBuf<float> rotRecon, rotAtn, unrot, vol, afac, ppart, plane, img, psf;
    DBuf<int> rowKernel, rowPad, zero, head, npix;
    const void *drfKey = nullptr;
    int drfKX = 1, drfKYb = 7;
    std::vector<int> hhead, hnpix;
    const void *esseKey = nullptr;
    std::unique_ptr<EsseDev> esse;
    DBuf<float> scatter, terms;
    ShearSet shears;              // slot 0: forward angle, slot 1: back angle of the call

    // Rotation of a [y][z][x] volume: the three-shear rotation (slot a of
    // the call's tables) when 'fast', else rot3dpar.
    void rotate_any(bool fast, int a, const float *src, float *dst, float sn, float cs,
                    int oldDim)
    {
        if (!fast) {
            rotate(src, dst, sn, cs, oldDim);
            return;
        }
        const long n = (long)NRot * Nz * NRot;
        hipLaunchKernelGGL(k_shear_rot, dim3(div_up(n, 256)), dim3(256), 0, stream, src, dst, Nz,
                           NRot, shears.dev(), a);
        HIP_CHECK(hipGetLastError());
    }

    void set_shears(const PrjCall &c)
    {
        if (!c.fast) return;
        if (c.N != c.NRot)
            fail(ETYPE_ILLEGAL_VALUE, "the shear rotation needs NumRotPixs == NumPixels");
        if (c.N % 2)
            fail(ETYPE_ILLEGAL_VALUE, "odd NumPixels (%d) is not supported: the reference "
                 "shear rotation reads past the row end for odd sizes", c.N);
        shears.set({shear_geometry(c.fang, NRot), shear_geometry(c.bang, NRot)}, NRot, stream);
    }

    // ESSE device instance of the handle (kept until another handle comes).
    EsseDev *esse_for(const void *h)
    {
        const EsseModel *m = esse_handle_model(h);
        if (!m || !m->enabled) return nullptr;
        if (h != esseKey) {
            if (m->nx != NRot || m->ny != NRot || m->nz != Nz)
                fail(ETYPE_ILLEGAL_VALUE, "scatter model prepared for another volume size");
            esse.reset(new EsseDev(*m, 1, stream, false));
            terms.alloc((long)NRot * Nz * NRot * m->nterms);
            esseKey = h;
        }
        return esse.get();
    }

    // q = act * primaryFac + effective scatter source (prjbck.c:164-176,
    // 395-406), written to dst (may alias act).
    void scatter_source(const void *h, const float *act, float primaryFac, float *dst)
    {
        EsseDev *e = esse_for(h);
        const long n = (long)NRot * Nz * NRot;
        if (e) {
            e->convolve(act, 1, terms.p);
            e->source(act, rotAtn.p, terms.p, 1, primaryFac, nullptr, nullptr, dst);
        } else {
            hipLaunchKernelGGL(k_scale, dim3(div_up(n, 256)), dim3(256), 0, stream, act, dst, n,
                               primaryFac);
            HIP_CHECK(hipGetLastError());
        }
    }

    PrjWs(int nz, int n, int nrot, int m) : Nz(nz), N(n), NRot(nrot), model(m)
    {
        HIP_CHECK(hipStreamCreate(&stream));
        const long v = (long)NRot * Nz * NRot;
        rotRecon.alloc(v);
        rotAtn.alloc(v);
        unrot.alloc(v);
        vol.alloc(v);
        afac.alloc(v);
        ppart.alloc((long)Nz * NRot * NRot);
        plane.alloc((long)Nz * NRot);
        img.alloc((long)N * Nz * N);
        if (m & 4) scatter.alloc(v);
        rowKernel.alloc(NRot);
        rowPad.alloc(NRot);
        int z0 = 0;
        zero.alloc(1);
        zero.upload(&z0, 1, stream);
        circle_rows(NRot, hhead, hnpix);
        head.alloc(NRot);
        head.upload(hhead.data(), NRot, stream);
        npix.alloc(NRot);
        npix.upload(hnpix.data(), NRot, stream);
        HIP_CHECK(hipMemsetAsync(rotRecon.p, 0, v * sizeof(float), stream));
        HIP_CHECK(hipMemsetAsync(rotAtn.p, 0, v * sizeof(float), stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }
    ~PrjWs()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    static int bucket(int KY)
    {
        static const int sizes[] = {7, 15, 23, 31, 47, 63};
        for (int s : sizes)
            if (KY <= s) return s;
        fail(ETYPE_ILLEGAL_VALUE, "DRF axial size %d exceeds the supported 63", KY);
    }

    void set_drf(const void *key, const PaddedDrf &d)
    {
        if (key == drfKey && key != nullptr) return;
        const int KYb = bucket(d.KY), KYP = (KYb + 3) & ~3, oy = (KYb - d.KY) / 2;
        std::vector<float> h((size_t)d.nDist * d.KX * KYP, 0.0f);
        for (int k = 0; k < d.nDist; ++k)
            for (int j = 0; j < d.KY; ++j)
                for (int i = 0; i < d.KX; ++i)
                    h[((size_t)k * d.KX + i) * KYP + j + oy] = d.psf[((size_t)k * d.KY + j) * d.KX + i];
        psf.alloc(h.size());
        psf.upload(h.data(), h.size(), stream);
        drfKey = key;
        drfKX = d.KX;
        drfKYb = KYb;
    }

    void rotate(const float *src, float *dst, float sn, float cs, int oldDim)
    {
        const long nout = (long)NRot * Nz * NRot;
        const float NewAxis = 0.5 * (NRot - 1), OldAxis = 0.5 * (oldDim - 1);
        hipLaunchKernelGGL(k_rot3dpar, dim3(div_up(nout, 256)), dim3(256), 0, stream, src, dst,
                           Nz, oldDim, NRot, sn, cs, NewAxis, OldAxis, head.p, npix.p);
        HIP_CHECK(hipGetLastError());
    }

    // DRF arguments of the workspace's bin-fastest planes ([row][z][x], as
    // the reference's RotReconImage), one view per call.
    DrfArgs drf_args(const int *rk, const int *rp, int avg)
    {
        const long pl = (long)Nz * NRot;
        DrfArgs da{};
        da.NRot = NRot;
        da.nrows = NRot;
        da.Nz = Nz;
        da.KX = drfKX;
        da.psf = psf.p;
        da.rowKernel = rk;
        da.rowPad = rp;
        da.avg = std::max(0, std::min(avg, Nz));
        da.bview = zero.p;
        da.scale = 1.0f;
        da.zs = NRot;
        da.xs = 1;
        da.inRow = pl;
        da.afRow = pl;
        da.outRow = pl;
        da.outZS = NRot;
        da.outXS = 1;
        da.xtiles = div_up(NRot, DRF_TX);
        return da;
    }
};

static PrjWs *g_prj = nullptr;

void prjws_init(int Nz, int N, int NRot, int model)
{
    delete g_prj;
    g_prj = nullptr;
    if (N <= 0 || Nz <= 0 || NRot < N)
        fail(ETYPE_ILLEGAL_VALUE, "InitPrj: bad sizes %d %d %d", Nz, N, NRot);
    g_prj = new PrjWs(Nz, N, NRot, model);
}

void prjws_free()
{
    delete g_prj;
    g_prj = nullptr;
}

static PrjWs &ws(int Nz, int N, int NRot)
{
    if (!g_prj) fail(ETYPE_ILLEGAL_VALUE, "InitPrj must be called before Project/BackProject");
    if (g_prj->Nz != Nz || g_prj->N != N || g_prj->NRot != NRot)
        fail(ETYPE_ILLEGAL_VALUE, "sizes differ from those passed to InitPrj");
    return *g_prj;
}

void prjws_project(const PrjCall &c, const float *atn, const float *recon, float *calcprj)
{
    PrjWs &W = ws(c.Nz, c.N, c.NRot);
    const long nin = (long)c.N * c.Nz * c.N;
    const bool useAtn = (c.model & 1) || (c.model & 4);
    W.set_shears(c);
    if (!c.actRotated) {
        W.vol.upload(recon, nin, W.stream);
        W.rotate_any(c.fast, 0, W.vol.p, W.rotRecon.p, c.fsin, c.fcos, c.N);
    }
    if (useAtn && !c.atnRotated) {
        W.vol.upload(atn, nin, W.stream);
        W.rotate_any(c.fast, 0, W.vol.p, W.rotAtn.p, c.fsin, c.fcos, c.N);
    }
    const long cols = (long)c.Nz * c.NRot;
    if (c.model & 1)
        hipLaunchKernelGGL(k_afac_rotated, dim3(div_up(cols, 256)), dim3(256), 0, W.stream,
                           W.rotAtn.p, W.afac.p, c.Nz, c.NRot, c.atnFac);
    if (c.drf) W.set_drf(c.drfKey, *c.drf);
    W.rowKernel.upload(c.rowKernel, c.NRot, W.stream);
    W.rowPad.upload(c.rowPad, c.NRot, W.stream);
    DrfArgs da = W.drf_args(W.rowKernel.p, W.rowPad.p, c.avg);
    da.in = W.rotRecon.p;
    if (c.model & 4) {
        if (!W.scatter.p) fail(ETYPE_ILLEGAL_VALUE, "InitPrj was called without MODEL_SRF");
        W.scatter_source(c.esse, W.rotRecon.p, c.primaryFac, W.scatter.p);
        da.in = W.scatter.p;
    }
    da.afac = (c.model & 1) ? W.afac.p : nullptr;
    da.out = W.ppart.p;
    da.rchunk = 1;
    da.nchunk = c.NRot;
    launch_drf(W.stream, false, da, 1, W.drfKYb);
    hipLaunchKernelGGL(k_sum_chunks, dim3(div_up(cols, 256)), dim3(256), 0, W.stream, W.ppart.p,
                       W.plane.p, cols, c.NRot, c.primary);
    HIP_CHECK(hipGetLastError());
    W.plane.download(calcprj, cols, W.stream);
    HIP_CHECK(hipStreamSynchronize(W.stream));
}

void prjws_backproject(const PrjCall &c, const float *atn, const float *ratio, float *image)
{
    PrjWs &W = ws(c.Nz, c.N, c.NRot);
    const long nin = (long)c.N * c.Nz * c.N;
    const bool useAtn = (c.model & 1) || (c.model & 4);
    W.set_shears(c);
    if (useAtn && !c.atnRotated) {
        W.vol.upload(atn, nin, W.stream);
        W.rotate_any(c.fast, 0, W.vol.p, W.rotAtn.p, c.fsin, c.fcos, c.N);
    }
    const long cols = (long)c.Nz * c.NRot;
    if (c.model & 1)
        hipLaunchKernelGGL(k_afac_rotated, dim3(div_up(cols, 256)), dim3(256), 0, W.stream,
                           W.rotAtn.p, W.afac.p, c.Nz, c.NRot, c.atnFac);
    if (c.drf) W.set_drf(c.drfKey, *c.drf);
    W.rowKernel.upload(c.rowKernel, c.NRot, W.stream);
    W.rowPad.upload(c.rowPad, c.NRot, W.stream);
    W.plane.upload(ratio, cols, W.stream);
    DrfArgs da = W.drf_args(W.rowKernel.p, W.rowPad.p, c.avg);
    da.in = W.plane.p;
    da.inRow = 0;
    da.afac = (c.model & 1) ? W.afac.p : nullptr;
    da.out = W.rotRecon.p;   // the reference backprojects into RotReconImage
    da.scale = c.primary;
    launch_drf(W.stream, true, da, 1, W.drfKYb);
    if (c.model & 4) W.scatter_source(c.esse, W.rotRecon.p, c.primaryFac, W.rotRecon.p);
    W.rotate_any(c.fast, 1, W.rotRecon.p, W.unrot.p, c.bsin, c.bcos, c.NRot);
    W.img.upload(image, nin, W.stream);
    hipLaunchKernelGGL(k_acc_clamp, dim3(div_up(nin, 256)), dim3(256), 0, W.stream, W.img.p,
                       W.unrot.p, c.N, c.Nz, c.NRot);
    HIP_CHECK(hipGetLastError());
    W.img.download(image, nin, W.stream);
    HIP_CHECK(hipStreamSynchronize(W.stream));
}

// ConvolveDrfYpad on one plane (convolve_ypad.c:22-96).
void dev_convolve_plane(const float *in, int inX, int inY, const PaddedDrf &d, int pad, int avg,
                        float *out)
{
    hipStream_t st;
    HIP_CHECK(hipStreamCreate(&st));
    const int KYb = PrjWs::bucket(d.KY), KYP = (KYb + 3) & ~3, oy = (KYb - d.KY) / 2;
    std::vector<float> h((size_t)d.KX * KYP, 0.0f);
    for (int j = 0; j < d.KY; ++j)
        for (int i = 0; i < d.KX; ++i) h[(size_t)i * KYP + j + oy] = d.psf[(size_t)j * d.KX + i];
    DBuf<float> dpsf, din, dout;
    DBuf<int> rk, rp, zero;
    dpsf.alloc(h.size());
    dpsf.upload(h.data(), h.size(), st);
    const long pl = (long)inX * inY;
    din.alloc(pl);
    din.upload(in, pl, st);
    dout.alloc(pl);
    int k0 = 0, p0 = pad;
    rk.alloc(1);
    rk.upload(&k0, 1, st);
    rp.alloc(1);
    rp.upload(&p0, 1, st);
    zero.alloc(1);
    zero.upload(&k0, 1, st);
    DrfArgs da{};
    da.NRot = inX;
    da.nrows = 1;
    da.Nz = inY;
    da.KX = d.KX;
    da.psf = dpsf.p;
    da.rowKernel = rk.p;
    da.rowPad = rp.p;
    da.avg = std::max(0, std::min(avg, inY));
    da.bview = zero.p;
    da.in = din.p;
    da.zs = inX;
    da.xs = 1;
    da.out = dout.p;
    da.outZS = inX;
    da.outXS = 1;
    da.scale = 1.0f;
    da.xtiles = div_up(inX, DRF_TX);
    launch_drf(st, true, da, 1, KYb);
    dout.download(out, pl, st);
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipStreamDestroy(st));
}

// ---------------------------------------------------------------------------
// Allocation-free plane convolution (ConvolveDrfYpadNoAlloc).  The PSF (in
// the kernel's padded column-major form), the input and output planes and
// the row tables are staged in one file-scope device array; each GPU holds
// its own copy until the process ends.  Host code never uses the array's
// name as an address: the device address is looked up once per device with
// hipGetSymbolAddress and kept.  Every call on a device shares that device's
// array, so calls that overlap on one device race; nothing in it has to
// start at zero (the kernel writes every output value).
// ---------------------------------------------------------------------------
constexpr int NOALLOC_MAX_PLANE = 256 * 256;   // inX * inY
constexpr int NOALLOC_PSF = 63 * 64;           // KX <= 63 columns of KYP <= 64 taps
constexpr int NOALLOC_TAB = 64;                // row tables (ints)
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) float scratch[NOALLOC_PSF + 2 * NOALLOC_MAX_PLANE + NOALLOC_TAB];

static float *noalloc_scratch()
{
    static float *ptr[NOALLOC_MAX_DEVICES] = {};
    static std::mutex m;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        fail(ETYPE_ILLEGAL_VALUE, "device %d beyond the %d the allocation-free path serves", dev,
             NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(m);
    if (!ptr[dev]) {
        void *p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = (float *)p;
    }
    return ptr[dev];
}

void dev_convolve_plane_noalloc(const float *in, int inX, int inY, const float *psf, int KX, int KY,
                                int pad, int avg, float *out)
{
    if (inX <= 0 || inY <= 0 || (long)inX * inY > NOALLOC_MAX_PLANE)
        fail(ETYPE_ILLEGAL_VALUE,
             "plane of %d x %d values: the allocation-free convolution takes at most %d", inX,
             inY, NOALLOC_MAX_PLANE);
    if (KX < 1 || KY < 1 || KX > 63 || KY > 63)
        fail(ETYPE_ILLEGAL_VALUE, "PSF of %d x %d: the allocation-free convolution takes at most "
                                  "63 x 63", KX, KY);
    const int KYb = PrjWs::bucket(KY), KYP = (KYb + 3) & ~3, oy = (KYb - KY) / 2;
    float h[NOALLOC_PSF];
    std::fill(h, h + (size_t)KX * KYP, 0.0f);
    for (int j = 0; j < KY; ++j)
        for (int i = 0; i < KX; ++i) h[(size_t)i * KYP + j + oy] = psf[(size_t)j * KX + i];
    float *dpsf = noalloc_scratch();
    float *din = dpsf + NOALLOC_PSF, *dout = din + NOALLOC_MAX_PLANE;
    int *tab = (int *)(dout + NOALLOC_MAX_PLANE);
    const int htab[3] = {0, pad, 0};   // DRF index of the row, its axial pad, view 0
    const long pl = (long)inX * inY;
    hipStream_t st = nullptr;          // the current device's null stream
    HIP_CHECK(hipMemcpyAsync(dpsf, h, sizeof(float) * KX * KYP, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(din, in, sizeof(float) * pl, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(tab, htab, sizeof htab, hipMemcpyHostToDevice, st));
    DrfArgs da{};
    da.NRot = inX;
    da.nrows = 1;
    da.Nz = inY;
    da.KX = KX;
    da.psf = dpsf;
    da.rowKernel = tab;
    da.rowPad = tab + 1;
    da.avg = std::max(0, std::min(avg, inY));
    da.bview = tab + 2;
    da.in = din;
    da.zs = inX;
    da.xs = 1;
    da.out = dout;
    da.outZS = inX;
    da.outXS = 1;
    da.scale = 1.0f;
    da.xtiles = div_up(inX, DRF_TX);
    launch_drf(st, true, da, 1, KYb);
    HIP_CHECK(hipMemcpyAsync(out, dout, sizeof(float) * pl, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

static size_t drf_lds_bytes(int ZG, int KYb, int KX)
{
    const int TW = DRF_TX + KX - 1;
    const int TH = ZG * DRF_ZT + KYb - 1;
    const int KYP = (KYb + 3) & ~3;
    return sizeof(float) * ((size_t)TW * drf_col_stride(TH) + (size_t)KX * KYP);
}

// Launches k_drf over B batch items.  Eight slice groups (128 slices per
// block) when the volume has more than 64 slices and two such blocks fit in
// a CU's 160 KiB of LDS, else four.
static void launch_drf(hipStream_t stream, bool bck, DrfArgs args, int B, int KYb, int tile)
{
    const size_t ldsMax = 160 * 1024;
    int ZG = (args.Nz > 64 && 2 * drf_lds_bytes(8, KYb, args.KX) <= ldsMax) ? 8 : 4;
    if ((tile == 4 || tile == 8) && drf_lds_bytes(tile, KYb, args.KX) <= ldsMax) ZG = tile;
    const size_t lds = drf_lds_bytes(ZG, KYb, args.KX);
    if (lds > ldsMax) fail(ETYPE_ILLEGAL_VALUE, "DRF PSF too wide (KX=%d)", args.KX);
    args.ztiles = div_up(args.Nz, ZG * DRF_ZT);
    const long nblocks = (long)args.xtiles * args.ztiles * B * (bck ? args.nrows : args.nchunk);
    if (nblocks <= 0 || nblocks > INT_MAX) fail(ETYPE_ILLEGAL_VALUE, "bad DRF launch size");
    dim3 grid((unsigned)nblocks), block(64 * ZG);
#define IRL_DRF_CASE(K)                                                                   \
    case K:                                                                               \
        if (ZG == 8) {                                                                    \
            if (bck)                                                                      \
                hipLaunchKernelGGL((k_drf<K, 8, true>), grid, block, lds, stream, args);  \
            else                                                                          \
                hipLaunchKernelGGL((k_drf<K, 8, false>), grid, block, lds, stream, args); \
        } else {                                                                          \
            if (bck)                                                                      \
                hipLaunchKernelGGL((k_drf<K, 4, true>), grid, block, lds, stream, args);  \
            else                                                                          \
                hipLaunchKernelGGL((k_drf<K, 4, false>), grid, block, lds, stream, args); \
        }                                                                                 \
        break;
    switch (KYb) {
        IRL_DRF_CASE(7)
        IRL_DRF_CASE(15)
        IRL_DRF_CASE(23)
        IRL_DRF_CASE(31)
        IRL_DRF_CASE(47)
        IRL_DRF_CASE(63)
    default:
        fail(ETYPE_ILLEGAL_VALUE, "unsupported DRF bucket %d", KYb);
    }
#undef IRL_DRF_CASE
    HIP_CHECK(hipGetLastError());
}

static void launch_drf_bf16(hipStream_t stream, bool bck, DrfArgs args, int B, int KYb,
                            bool split, const uint4 *afr, const uint4 *afrLo,
                            const int *distSplit, const int4 *band, int tile)
{
    if (args.zs != 1 || args.outZS != 1 || args.Nz % 4)
        fail(ETYPE_ILLEGAL_VALUE, "the BF16 DRF path needs slice-fastest planes, Nz %% 4 == 0");
    const int NZT = tile == 2 || tile == 4 || tile == 8 ? tile : args.Nz > 64 ? 8 : 4;
    const int TW = DRF_TX + args.KX - 1;
    const int THP = 16 * (NZT - 1) + 32 * drf16_nks(KYb);
    const size_t lds = (size_t)TW * drf16_cs(THP) * 2;   // one bf16 tile (splits run two passes)
    if (lds > 160 * 1024) fail(ETYPE_ILLEGAL_VALUE, "DRF PSF too wide for the BF16 path (KX=%d)", args.KX);
    args.ztiles = div_up(args.Nz, 16 * NZT);
    const long nblocks = (long)args.xtiles * args.ztiles * B * (bck ? args.nrows : args.nchunk);
    if (nblocks <= 0 || nblocks > INT_MAX) fail(ETYPE_ILLEGAL_VALUE, "bad DRF launch size");
    dim3 grid((unsigned)nblocks), block(256);
#define IRL_DRF16_LAUNCH(K, Z, BK) \
    hipLaunchKernelGGL((k_drf_bf16<K, Z, BK>), grid, block, lds, stream, args, afr, afrLo, \
                       distSplit, band)
#define IRL_DRF16_CASE(K)                                                            \
    case K:                                                                          \
        if (NZT == 8) {                                                              \
            if (bck) IRL_DRF16_LAUNCH(K, 8, true); else IRL_DRF16_LAUNCH(K, 8, false); \
        } else if (NZT == 4) {                                                       \
            if (bck) IRL_DRF16_LAUNCH(K, 4, true); else IRL_DRF16_LAUNCH(K, 4, false); \
        } else {                                                                     \
            if (bck) IRL_DRF16_LAUNCH(K, 2, true); else IRL_DRF16_LAUNCH(K, 2, false); \
        }                                                                            \
        break;
    switch (KYb) {
        IRL_DRF16_CASE(7)
        IRL_DRF16_CASE(15)
        IRL_DRF16_CASE(23)
        IRL_DRF16_CASE(31)
        IRL_DRF16_CASE(47)
        IRL_DRF16_CASE(63)
    default:
        fail(ETYPE_ILLEGAL_VALUE, "unsupported DRF bucket %d", KYb);
    }
#undef IRL_DRF16_CASE
#undef IRL_DRF16_LAUNCH
    HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Stateless helper for the rot3dpar entry point.
// ---------------------------------------------------------------------------
void dev_rot3dpar(float angle, const float *in_host, float *out_host, int pbFlag, int Nz,
                  int N, int NRot)
{
    const int OldDim = pbFlag ? NRot : N, NewDim = NRot;
    const float sn = (float)sin(angle), cs = (float)cos(angle);
    const float NewAxis = 0.5 * (NewDim - 1), OldAxis = 0.5 * (OldDim - 1);
    std::vector<int> head, npix;
    circle_rows(NewDim, head, npix);
    hipStream_t st;
    HIP_CHECK(hipStreamCreate(&st));
    DBuf<float> din, dout;
    DBuf<int> dh, dn;
    const long nin = (long)OldDim * Nz * OldDim, nout = (long)NewDim * Nz * NewDim;
    din.alloc(nin);
    dout.alloc(nout);
    dh.alloc(NewDim);
    dn.alloc(NewDim);
    din.upload(in_host, nin, st);
    dh.upload(head.data(), NewDim, st);
    dn.upload(npix.data(), NewDim, st);
    hipLaunchKernelGGL(k_rot3dpar, dim3(div_up(nout, 256)), dim3(256), 0, st, din.p, dout.p, Nz,
                       OldDim, NewDim, sn, cs, NewAxis, OldAxis, dh.p, dn.p);
    HIP_CHECK(hipGetLastError());
    dout.download(out_host, nout, st);
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipStreamDestroy(st));
}

// rot3dpar of one N x N image (PB_FLAG 0, one slice) into NRot x NRot for
// each of V angles (ComputeOrbit, computeorbit.c:97-110); out is [V][NRot][NRot].
void dev_rot2d_views(const float *img_host, int N, int NRot, const float *angles, int V,
                     float *out_host)
{
    std::vector<int> head, npix;
    circle_rows(NRot, head, npix);
    hipStream_t st;
    HIP_CHECK(hipStreamCreate(&st));
    {
        DBuf<float> din, dout;
        DBuf<int> dh, dn;
        const long nin = (long)N * N, nout = (long)NRot * NRot;
        din.alloc(nin);
        dout.alloc(nout * V);
        dh.alloc(NRot);
        dn.alloc(NRot);
        din.upload(img_host, nin, st);
        dh.upload(head.data(), NRot, st);
        dn.upload(npix.data(), NRot, st);
        const float NewAxis = 0.5 * (NRot - 1), OldAxis = 0.5 * (N - 1);
        for (int v = 0; v < V; ++v) {
            const float sn = (float)sin(angles[v]), cs = (float)cos(angles[v]);
            hipLaunchKernelGGL(k_rot3dpar, dim3(div_up(nout, 256)), dim3(256), 0, st, din.p,
                               dout.p + nout * v, 1, N, NRot, sn, cs, NewAxis, OldAxis, dh.p,
                               dn.p);
            HIP_CHECK(hipGetLastError());
        }
        dout.download(out_host, nout * V, st);
        HIP_CHECK(hipStreamSynchronize(st));
    }
    HIP_CHECK(hipStreamDestroy(st));
}

// vShearsRotate on host arrays (fastrotate.c:330-563), [y][z][x] volumes of
// N x Nz x N.
void dev_shears_rotate(float angle, const float *in_host, float *out_host, int Nz, int N)
{
    // odd N gives the centre row N+1 pixels (fastrotate.c:584-586), past the row end
    if (N % 2)
        fail(ETYPE_ILLEGAL_VALUE, "odd NumPixels (%d) is not supported by the shear rotation", N);
    hipStream_t st;
    HIP_CHECK(hipStreamCreate(&st));
    {
        ShearSet ss;
        ss.set({shear_geometry(angle, N)}, N, st);
        DBuf<float> din, dout;
        const long n = (long)N * Nz * N;
        din.alloc(n);
        dout.alloc(n);
        din.upload(in_host, n, st);
        hipLaunchKernelGGL(k_shear_rot, dim3(div_up(n, 256)), dim3(256), 0, st, din.p, dout.p, Nz,
                           N, ss.dev(), 0);
        HIP_CHECK(hipGetLastError());
        dout.download(out_host, n, st);
        HIP_CHECK(hipStreamSynchronize(st));
    }
    HIP_CHECK(hipStreamDestroy(st));
}

}  // namespace irlamd
