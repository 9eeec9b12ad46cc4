// esse.hip -- ESSE effective scatter source on gfx950 (krnlprj3d.c:549-954).
//
// The reference convolves the (optionally collapsed) rotated activity with
// three kernels by real FFTs on a 2x padded grid.  Here the same circular
// convolutions are evaluated as linear convolutions with the folded kernels
// on power-of-two grids (esse_host.cpp), by a batched hand-written Stockham
// FFT over all views of a subset:
//   X forward  : rows of the activity -> complex rows (zero rows skipped)
//   Z forward  : columns along the slices
//   Y conv     : forward along depth, multiply by every kernel spectrum,
//                inverse along depth, keep the cropped depth range
//   Z inverse, X inverse: back to real, cropped to the volume
// Each pass loads 8 or 16 lines into LDS (line stride padded by one complex
// element against bank conflicts), runs log2(P) radix-2 stages there and
// writes the lines back; lines of the strided axes are grouped so that every
// load and store moves 8-16 consecutive complex values.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "esse_dev.h"

namespace irlamd {

#define ESSE_CHECK(expr)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::irlamd::fail(ETYPE_DEVICE, "HIP error '%s' in %s (%s:%d)",                 \
                           hipGetErrorString(e_), #expr, __FILE__, __LINE__);            \
    } while (0)

static int ilog2(int p)
{
    int l = 0;
    while ((1 << l) < p) ++l;
    return l;
}

static long cdiv(long a, long b) { return (a + b - 1) / b; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Stockham radix-2 FFT of L lines of length P (line stride LS) held in LDS.
// Stage st: n = P >> st, m = n/2, s = 2^st; for p < m, q < s:
//   y[q + s*2p] = a + b,  y[q + s*(2p+1)] = (a - b) * exp(-+2 pi i p s / P)
// with a = x[q + s*p], b = x[q + s*(p+m)].  Output is in natural order.
template <int L>
__device__ float2 *lds_fft(float2 *a, float2 *b, int P, int logP, int LS,
                           const float2 *__restrict__ tw, bool inv)
{
    const int half = P >> 1;
    for (int st = 0; st < logP; ++st) {
        const int s = 1 << st;
        const int m = half >> st;
        for (int idx = threadIdx.x; idx < L * half; idx += blockDim.x) {
            const int line = idx >> (logP - 1);
            const int i = idx & (half - 1);
            const int p = i >> st, q = i & (s - 1);
            const float2 *x = a + line * LS;
            float2 *y = b + line * LS;
            const float2 A = x[q + s * p], B = x[q + s * (p + m)];
            float2 w = tw[p * s];
            if (inv) w.y = -w.y;
            y[q + s * 2 * p] = make_float2(A.x + B.x, A.y + B.y);
            y[q + s * (2 * p + 1)] = cmul(make_float2(A.x - B.x, A.y - B.y), w);
        }
        __syncthreads();
        float2 *t = a;
        a = b;
        b = t;
    }
    return a;
}

// X forward: real rows (r = plane*nz + z, plane = b*ny + y) -> buf rows
// [plane][z][Px]; x >= nx is zero.
template <int L>
__global__ __launch_bounds__(256) void k_fft_x_fwd(const float *__restrict__ in,
                                                   float2 *__restrict__ buf, long rows, int nz,
                                                   int nx, int Pz, int Px, int logPx,
                                                   const float2 *__restrict__ tw)
{
    extern __shared__ float2 sm[];
    const int LS = Px + 1;
    float2 *a = sm, *b = sm + L * LS;
    const long r0 = (long)blockIdx.x * L;
    for (int e = threadIdx.x; e < L * Px; e += blockDim.x) {
        const int l = e >> logPx, x = e & (Px - 1);
        const long r = r0 + l;
        a[l * LS + x] = (r < rows && x < nx) ? make_float2(in[r * nx + x], 0.0f)
                                             : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const float2 *res = lds_fft<L>(a, b, Px, logPx, LS, tw, false);
    for (int e = threadIdx.x; e < L * Px; e += blockDim.x) {
        const int l = e >> logPx, x = e & (Px - 1);
        const long r = r0 + l;
        if (r < rows) {
            const long plane = r / nz, z = r - plane * nz;
            buf[(plane * Pz + z) * Px + x] = res[l * LS + x];
        }
    }
}

// X inverse: buf rows [plane][z][Px] for z < nz -> real [plane][nz][nx].
template <int L>
__global__ __launch_bounds__(256) void k_fft_x_inv(const float2 *__restrict__ buf,
                                                   float *__restrict__ out, long rows, int nz,
                                                   int nx, int Pz, int Px, int logPx,
                                                   const float2 *__restrict__ tw, float scale)
{
    extern __shared__ float2 sm[];
    const int LS = Px + 1;
    float2 *a = sm, *b = sm + L * LS;
    const long r0 = (long)blockIdx.x * L;
    for (int e = threadIdx.x; e < L * Px; e += blockDim.x) {
        const int l = e >> logPx, x = e & (Px - 1);
        const long r = r0 + l;
        float2 v = make_float2(0.0f, 0.0f);
        if (r < rows) {
            const long plane = r / nz, z = r - plane * nz;
            v = buf[(plane * Pz + z) * Px + x];
        }
        a[l * LS + x] = v;
    }
    __syncthreads();
    const float2 *res = lds_fft<L>(a, b, Px, logPx, LS, tw, true);
    for (int e = threadIdx.x; e < L * Px; e += blockDim.x) {
        const int l = e >> logPx, x = e & (Px - 1);
        const long r = r0 + l;
        if (r < rows && x < nx) out[r * nx + x] = res[l * LS + x].x * scale;
    }
}

// Z pass on planes [plane][Pz][Px]: lines along z at L consecutive kx.
// Input z >= nzIn is zero; only z < nzOut is written back.
template <int L, bool INV>
__device__ void fft_z_body(float2 *__restrict__ data, int Pz, int logPz, int Px, int nzIn,
                           int nzOut, const float2 *__restrict__ tw)
{
    extern __shared__ float2 sm[];
    const int LS = Pz + 1;
    float2 *a = sm, *b = sm + L * LS;
    const int groups = Px / L;
    const long plane = blockIdx.x / groups;
    const int kx0 = (blockIdx.x % groups) * L;
    float2 *base = data + plane * Pz * Px + kx0;
    for (int e = threadIdx.x; e < L * Pz; e += blockDim.x) {
        const int l = e % L, z = e / L;
        a[l * LS + z] = z < nzIn ? base[(long)z * Px + l] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const float2 *res = lds_fft<L>(a, b, Pz, logPz, LS, tw, INV);
    for (int e = threadIdx.x; e < L * Pz; e += blockDim.x) {
        const int l = e % L, z = e / L;
        if (z < nzOut) base[(long)z * Px + l] = res[l * LS + z];
    }
}

template <int L>
__global__ __launch_bounds__(256) void k_fft_z_fwd(float2 *__restrict__ data, int Pz, int logPz,
                                                   int Px, int nzIn, int nzOut,
                                                   const float2 *__restrict__ tw)
{
    fft_z_body<L, false>(data, Pz, logPz, Px, nzIn, nzOut, tw);
}

template <int L>
__global__ __launch_bounds__(256) void k_fft_z_inv(float2 *__restrict__ data, int Pz, int logPz,
                                                   int Px, int nzIn, int nzOut,
                                                   const float2 *__restrict__ tw)
{
    fft_z_body<L, true>(data, Pz, logPz, Px, nzIn, nzOut, tw);
}

// Y forward only (kernel spectra): src [Py][Pz][Px] -> dst same layout.
template <int L>
__global__ __launch_bounds__(256) void k_fft_y_fwd(const float2 *__restrict__ src,
                                                   float2 *__restrict__ dst, int Py, int logPy,
                                                   int Pz, int Px,
                                                   const float2 *__restrict__ tw)
{
    extern __shared__ float2 sm[];
    const int LS = Py + 1;
    float2 *a = sm, *b = sm + L * LS;
    const int groups = Px / L;
    const int z = blockIdx.x / groups;
    const int kx0 = (blockIdx.x % groups) * L;
    const long ystride = (long)Pz * Px;
    for (int e = threadIdx.x; e < L * Py; e += blockDim.x) {
        const int l = e % L, y = e / L;
        a[l * LS + y] = src[y * ystride + (long)z * Px + kx0 + l];
    }
    __syncthreads();
    const float2 *res = lds_fft<L>(a, b, Py, logPy, LS, tw, false);
    for (int e = threadIdx.x; e < L * Py; e += blockDim.x) {
        const int l = e % L, y = e / L;
        dst[y * ystride + (long)z * Px + kx0 + l] = res[l * LS + y];
    }
}

// Y conv: per batch item b, lines along depth at (z, L consecutive kx):
// forward FFT (rows y >= ny are zero), then for every kernel t multiply by
// its spectrum, inverse FFT and keep y < ny.
template <int L>
__global__ __launch_bounds__(256) void k_fft_y_conv(const float2 *__restrict__ buf,
                                                    const float2 *__restrict__ khat,
                                                    float2 *__restrict__ outk, int T, int ny,
                                                    int Py, int logPy, int Pz, int Px,
                                                    const float2 *__restrict__ tw)
{
    extern __shared__ float2 sm[];
    const int LS = Py + 1;
    float2 *spec = sm, *a = sm + L * LS, *b = sm + 2 * L * LS;
    const int groups = Px / L;
    long bid = blockIdx.x;
    const int kx0 = (int)(bid % groups) * L;
    bid /= groups;
    const int z = (int)(bid % Pz);
    const long bb = bid / Pz;
    const long ystride = (long)Pz * Px;
    const float2 *src = buf + bb * ny * ystride + (long)z * Px + kx0;
    for (int e = threadIdx.x; e < L * Py; e += blockDim.x) {
        const int l = e % L, y = e / L;
        a[l * LS + y] = y < ny ? src[y * ystride + l] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const float2 *f = lds_fft<L>(a, b, Py, logPy, LS, tw, false);
    for (int e = threadIdx.x; e < L * LS; e += blockDim.x) spec[e] = f[e];
    __syncthreads();
    for (int t = 0; t < T; ++t) {
        const float2 *kt = khat + (long)t * Py * ystride + (long)z * Px + kx0;
        for (int e = threadIdx.x; e < L * Py; e += blockDim.x) {
            const int l = e % L, y = e / L;
            a[l * LS + y] = cmul(spec[l * LS + y], kt[y * ystride + l]);
        }
        __syncthreads();
        const float2 *g = lds_fft<L>(a, b, Py, logPy, LS, tw, true);
        float2 *dst = outk + (bb * T + t) * ny * ystride + (long)z * Px + kx0;
        for (int e = threadIdx.x; e < L * Py; e += blockDim.x) {
            const int l = e % L, y = e / L;
            if (y < ny) dst[y * ystride + l] = g[l * LS + y];
        }
        __syncthreads();
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
__global__ void k_esse_source(const float *__restrict__ act, const float *__restrict__ map,
                              const float *__restrict__ terms, int B, int T, int ny, int nz,
                              int nx, float fAtnCnvFac, float fMu, float cntr, float fDensFac,
                              float primary, const float *__restrict__ afac,
                              const int *__restrict__ bview, float *__restrict__ out)
{
    const long cols = (long)nz * nx;
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= cols * B) return;
    const int b = (int)(gid / cols);
    const long c = gid % cols;
    const long vol = cols * ny;
    const float *A = act + b * vol + c;
    const float *Mp = map + b * vol + c;
    const float *E = terms + (long)b * T * vol + c;
    const float *F = afac ? afac + (long)bview[b] * vol + c : nullptr;
    float *O = out + b * vol + c;
    float sum = 0.0f;
    for (int y = 0; y < ny; ++y) {
        const long o = (long)y * cols;
        const float m = Mp[o];
        sum = (float)((double)sum + 0.5 * (double)m);
        const float d = sum * fAtnCnvFac;
        float eff = E[o];
        if (T > 1) eff = eff - d * E[vol + o];
        if (T > 2) eff = (float)((double)eff + 0.5 * (double)d * (double)d * (double)E[2 * vol + o]);
        eff = (float)((double)eff * exp((double)(d * fMu)));
        sum = (float)((double)sum + 0.5 * (double)m);
        const float a = A[o];
        if (cntr > 0) eff = eff + a * cntr;
        eff = eff * (m * fDensFac);
        if (eff < 0) eff = 0.0f;
        const float q = a * primary + eff;
        O[o] = F ? q * F[o] : q;
    }
}

static std::vector<float2> twiddles(int P)
{
    std::vector<float2> t(P / 2 > 0 ? P / 2 : 1);
    for (int k = 0; k < P / 2; ++k) {
        const double a = -2.0 * M_PI * k / P;
        t[k] = make_float2((float)cos(a), (float)sin(a));
    }
    return t;
}

template <typename T>
static T *dalloc(size_t n)
{
    T *p = nullptr;
    if (n) ESSE_CHECK(hipMalloc((void **)&p, n * sizeof(T)));
    return p;
}

static int lines_for(int P) { return P <= 256 ? 16 : 8; }

#define ESSE_L_DISPATCH(P, KERNEL, GRID, LDS1, ...)                                         \
    do {                                                                                     \
        if (lines_for(P) == 16)                                                              \
            hipLaunchKernelGGL((KERNEL<16>), dim3(GRID), dim3(256), (LDS1) * 16, stream_,    \
                               __VA_ARGS__);                                                 \
        else                                                                                 \
            hipLaunchKernelGGL((KERNEL<8>), dim3(GRID), dim3(256), (LDS1) * 8, stream_,      \
                               __VA_ARGS__);                                                 \
        ESSE_CHECK(hipGetLastError());                                                       \
    } while (0)

EsseDev::EsseDev(const EsseModel &m, int maxBatch, hipStream_t stream, bool zfast)
    : m_(m), maxB_(maxBatch), stream_(stream), zfast_(zfast)
{
    if (!m_.enabled) return;
    const int Px = m_.Px, Py = m_.Py, Pz = m_.Pz, T = m_.nterms;
    // The FFT passes run along the layout's fastest axis first ("f"), then
    // the other in-plane axis ("m"), then depth.
    Pf_ = zfast ? Pz : Px;
    Pm_ = zfast ? Px : Pz;
    const long grid = (long)Px * Py * Pz;
    auto tf = twiddles(Pf_), ty = twiddles(Py), tm = twiddles(Pm_);
    twf_ = dalloc<float2>(tf.size());
    twy_ = dalloc<float2>(ty.size());
    twm_ = dalloc<float2>(tm.size());
    ESSE_CHECK(hipMemcpyAsync(twf_, tf.data(), tf.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
    ESSE_CHECK(hipMemcpyAsync(twy_, ty.data(), ty.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
    ESSE_CHECK(hipMemcpyAsync(twm_, tm.data(), tm.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
    khat_ = dalloc<float2>(grid * T);
    buf_ = dalloc<float2>((size_t)maxB_ * m_.nyc * Pz * Px);
    outk_ = dalloc<float2>((size_t)maxB_ * T * m_.nyc * Pz * Px);
    if (m_.c > 1) {
        coll_ = dalloc<float>((size_t)maxB_ * m_.nyc * m_.nzc * m_.nxc);
        res_ = dalloc<float>((size_t)maxB_ * T * m_.nyc * m_.nzc * m_.nxc);
    }
    // Kernel spectra [T][Py][Pm][Pf]: full 3-D forward FFT of each folded
    // kernel grid (held [Py][Pz][Px], transposed for the slice-fastest layout).
    float *dgrid = dalloc<float>(grid);
    float2 *tmp = dalloc<float2>(grid);
    const int lf = ilog2(Pf_), ly = ilog2(Py), lm = ilog2(Pm_);
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
        const long rows = (long)Py * Pm_;
        ESSE_L_DISPATCH(Pf_, k_fft_x_fwd, cdiv(rows, lines_for(Pf_)), 2 * (Pf_ + 1) * sizeof(float2),
                        dgrid, tmp, rows, Pm_, Pf_, Pm_, Pf_, lf, twf_);
        ESSE_L_DISPATCH(Pm_, k_fft_z_fwd, (long)Py * (Pf_ / lines_for(Pm_)),
                        2 * (Pm_ + 1) * sizeof(float2), tmp, Pm_, lm, Pf_, Pm_, Pm_, twm_);
        ESSE_L_DISPATCH(Py, k_fft_y_fwd, (long)Pm_ * (Pf_ / lines_for(Py)),
                        2 * (Py + 1) * sizeof(float2), tmp, khat_ + grid * t, Py, ly, Pm_, Pf_, twy_);
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

void EsseDev::convolve(const float *act, int B, float *terms)
{
    if (!m_.enabled) return;
    if (B > maxB_) fail(ETYPE_ILLEGAL_VALUE, "ESSE batch %d exceeds %d", B, maxB_);
    const int Py = m_.Py, Pf = Pf_, Pm = Pm_, T = m_.nterms;
    const int ny = m_.nyc, c = m_.c;
    const int nf = zfast_ ? m_.nzc : m_.nxc, nm = zfast_ ? m_.nxc : m_.nzc;
    const int lf = ilog2(Pf), ly = ilog2(Py), lm = ilog2(Pm);
    const float *src = act;
    if (c > 1) {
        const long n = (long)B * ny * nf * nm;
        hipLaunchKernelGGL(k_collapse, dim3(cdiv(n, 256)), dim3(256), 0, stream_, act, coll_, n,
                           m_.ny, m_.nz, m_.nx, c, zfast_);
        ESSE_CHECK(hipGetLastError());
        src = coll_;
    }
    const long rows = (long)B * ny * nm;
    ESSE_L_DISPATCH(Pf, k_fft_x_fwd, cdiv(rows, lines_for(Pf)), 2 * (Pf + 1) * sizeof(float2), src,
                    buf_, rows, nm, nf, Pm, Pf, lf, twf_);
    ESSE_L_DISPATCH(Pm, k_fft_z_fwd, (long)B * ny * (Pf / lines_for(Pm)),
                    2 * (Pm + 1) * sizeof(float2), buf_, Pm, lm, Pf, nm, Pm, twm_);
    ESSE_L_DISPATCH(Py, k_fft_y_conv, (long)B * Pm * (Pf / lines_for(Py)),
                    3 * (Py + 1) * sizeof(float2), buf_, khat_, outk_, T, ny, Py, ly, Pm, Pf, twy_);
    ESSE_L_DISPATCH(Pm, k_fft_z_inv, (long)B * T * ny * (Pf / lines_for(Pm)),
                    2 * (Pm + 1) * sizeof(float2), outk_, Pm, lm, Pf, Pm, nm, twm_);
    const float scale = (float)(1.0 / ((double)m_.Px * Py * m_.Pz));
    const long orows = (long)B * T * ny * nm;
    float *dst = c > 1 ? res_ : terms;
    ESSE_L_DISPATCH(Pf, k_fft_x_inv, cdiv(orows, lines_for(Pf)), 2 * (Pf + 1) * sizeof(float2),
                    outk_, dst, orows, nm, nf, Pm, Pf, lf, twf_, scale);
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

void EsseDev::source(const float *act, const float *map, const float *terms, int B,
                     float primary, const float *afac, const int *bview, float *out)
{
    const long cols = (long)m_.nz * m_.nx;
    const float fAtnCnvFac = 1.0 / (m_.mu0);
    const float fMu = m_.mu0 - m_.muavg;
    const float fDens = fAtnCnvFac / m_.pixsize;
    hipLaunchKernelGGL(k_esse_source, dim3(cdiv(cols * B, 256)), dim3(256), 0, stream_, act, map,
                       terms, B, m_.nterms, m_.ny, m_.nz, m_.nx, fAtnCnvFac, fMu, m_.cntrPixFac,
                       fDens, primary, afac, bview, out);
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
