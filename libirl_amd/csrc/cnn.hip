// cnn.hip -- scatter CNN on gfx950 (no counterpart in the reference; it
// stands in for the FFT convolutions of CalcEffScatterSource,
// krnlprj3d.c:826-948, which remain the fallback and the check).
//
// Network (half-resolution grid, tensors [Y][X][Z][C], channels fastest):
//   input  2 ch: 2x2x2 mean of x / mean(x), 2x2x2 mean of mu * fDens
//   L1 2->32, L2 32->32, L3 32->32 (ReLU), L4 32->2 (linear);
//   3x3x3 taps at dilations 1, 2, 4, 8 (receptive field 31 coarse voxels)
//   output 2 ch: E0 / mean(x), kappa (fMu E0 - E1) / mean(x), 2x2x2 means.
// Inference: L1 and L4 on the VALU (FP32 accumulation, BF16-rounded weights),
// L2 and L3 as implicit GEMMs on v_mfma_f32_16x16x32_bf16 (BF16 weights and
// activations, FP32 accumulation).  Training: every layer in FP32 on the
// VALU, the input gradient as the same convolution with flipped and
// transposed weights, weight gradients reduced per tap in LDS and added to
// global memory with atomics, Adam on FP32 master weights.
#include <hip/hip_runtime.h>
#include <climits>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <type_traits>

#include "cnn_dev.h"
#include "esse_dev.h"

namespace irlamd {

#define CNN_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::irlamd::fail(::irlamd::ETYPE_DEVICE, "HIP error '%s' in %s (%s:%d)",                 \
                           hipGetErrorString(e_), #expr, __FILE__, __LINE__);            \
    } while (0)

typedef __bf16 cbf16x8_t __attribute__((ext_vector_type(8)));
typedef float cfloatx4_t __attribute__((ext_vector_type(4)));

long cnn_num_params()
{
    long n = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) n += cnn_layer_params(l);
    return n;
}

static unsigned cdiv_u(long a, long b) { return (unsigned)((a + b - 1) / b); }

__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(uint16_t v)
{
    return __uint_as_float((uint32_t)v << 16);
}
__device__ __forceinline__ void from_f(float &d, float v) { d = v; }
__device__ __forceinline__ void from_f(uint16_t &d, float v)
{
    d = __builtin_bit_cast(uint16_t, (__bf16)v);
}

__device__ __forceinline__ void tap_off(int t, int dil, int &dy, int &dx, int &dz)
{
    dy = dil * (t / 9 - 1);
    dx = dil * ((t / 3) % 3 - 1);
    dz = dil * (t % 3 - 1);
}

// Block index in an XCD-contiguous order: workgroups are dispatched to the
// 8 XCDs round-robin (workgroup i on XCD i % 8), so block i takes logical
// block (i % 8) * (n / 8) + i / 8 and each XCD works through one contiguous
// range of rows, whose neighbour rows then stay in that XCD's L2.
__device__ __forceinline__ unsigned xcd_contiguous_block()
{
    const unsigned n = gridDim.x, i = blockIdx.x;
    return n % 8 == 0 ? (i % 8) * (n / 8) + i / 8 : i;
}

// ---------------------------------------------------------------------------
// Mean of a volume (FP64 accumulation): partial sums per block, then 1/mean.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sum_partial(const float *__restrict__ v, long n,
                                                     double *__restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        s += (double)v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// 1/mean and mean (1 where the mean is not positive) of n values from nb
// FP64 block sums, by a block of 256 threads with 256 doubles of LDS: every
// caller adds the sums in the same order, so all get the same floats.
__device__ __forceinline__ void inv_mean_block(const double *__restrict__ part, int nb, long n,
                                               double *red, float &inv, float &mean)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double m = red[0] / (double)n;
    mean = m > 0 ? (float)m : 1.0f;
    inv = 1.0f / mean;
}

__global__ __launch_bounds__(256) void k_inv_mean(const double *__restrict__ part, int nb, long n,
                                                  float *__restrict__ scal)
{
    __shared__ double red[256];
    float inv, mean;
    inv_mean_block(part, nb, n, red, inv, mean);
    if (threadIdx.x == 0) {
        scal[0] = inv;
        scal[1] = mean;
    }
}

// Network input from a fine volume pair (engine layout [y][x][z]).
__global__ __launch_bounds__(256) void k_cnn_input(const float *__restrict__ x,
                                                   const float *__restrict__ mu, int N, int Nz,
                                                   const float *__restrict__ scal, float densFac,
                                                   float *__restrict__ in)
{
    const int Y = N / 2, X = N / 2, Z = Nz / 2;
    const long v = blockIdx.x * 256L + threadIdx.x;
    if (v >= (long)Y * X * Z) return;
    const int z = (int)(v % Z);
    const long r = v / Z;
    const int xx = (int)(r % X), y = (int)(r / X);
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const long o = ((long)(2 * y + dy) * N + 2 * xx + dx) * Nz + 2 * z;
            s0 += x[o] + x[o + 1];
            if (mu) s1 += mu[o] + mu[o + 1];
        }
    in[2 * v] = s0 * 0.125f * scal[0];
    in[2 * v + 1] = s1 * 0.125f * densFac;
}

// Inference input: 2x2x2 means of x (not yet divided by mean(x): layer 1
// applies 1/mean(x) when it stages its input) and of mu * densFac, plus the
// block sums of x (FP64) for that mean.
__global__ __launch_bounds__(256) void k_cnn_pool(const float *__restrict__ x,
                                                  const float *__restrict__ mu, int N, int Nz,
                                                  float densFac, float *__restrict__ in,
                                                  double *__restrict__ part)
{
    __shared__ double red[256];
    const int Y = N / 2, X = N / 2, Z = Nz / 2;
    const int v = blockIdx.x * 256 + threadIdx.x;
    float s0 = 0.0f, s1 = 0.0f;
    if (v < Y * X * Z) {
        const int z = v % Z;
        const int r = v / Z;
        const int xx = r % X, y = r / X;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int o = ((2 * y + dy) * N + 2 * xx + dx) * Nz + 2 * z;
                s0 += x[o] + x[o + 1];
                if (mu) s1 += mu[o] + mu[o + 1];
            }
        in[2 * v] = s0 * 0.125f;
        in[2 * v + 1] = s1 * 0.125f * densFac;
    }
    red[threadIdx.x] = (double)s0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Training targets from the ESSE terms (E0, E1 volumes of one sample).
__global__ __launch_bounds__(256) void k_cnn_target(const float *__restrict__ E0,
                                                    const float *__restrict__ E1, int N, int Nz,
                                                    const float *__restrict__ scal, float fMu,
                                                    float kappa, float *__restrict__ tgt)
{
    const int Y = N / 2, X = N / 2, Z = Nz / 2;
    const long v = blockIdx.x * 256L + threadIdx.x;
    if (v >= (long)Y * X * Z) return;
    const int z = (int)(v % Z);
    const long r = v / Z;
    const int xx = (int)(r % X), y = (int)(r / X);
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx)
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                const long o = ((long)(2 * y + dy) * N + 2 * xx + dx) * Nz + 2 * z + dz;
                s0 += E0[o];
                s1 += fMu * E0[o] - E1[o];
            }
    tgt[2 * v] = s0 * 0.125f * scal[0];
    tgt[2 * v + 1] = kappa * s1 * 0.125f * scal[0];
}

// ---------------------------------------------------------------------------
// VALU convolution: one thread per output voxel, CO accumulators.  Weights
// [27][CI][CO] (uniform over the wave: scalar loads).  Optional bias, ReLU
// and a mask (out *= mask > 0, the ReLU derivative of the layer below).
// ---------------------------------------------------------------------------
template <int CO, typename TI, typename TO>
__global__ __launch_bounds__(256) void k_conv_valu(const TI *__restrict__ in,
                                                   const float *__restrict__ W,
                                                   const float *__restrict__ bias,
                                                   const float *__restrict__ mask,
                                                   TO *__restrict__ out, int Y, int X, int Z,
                                                   int CI, int dil, int relu)
{
    const long V = (long)Y * X * Z;
    const long v = blockIdx.x * 256L + threadIdx.x;
    if (v >= V) return;
    const int z = (int)(v % Z);
    const long r = v / Z;
    const int x = (int)(r % X), y = (int)(r / X);
    float acc[CO];
#pragma unroll
    for (int co = 0; co < CO; ++co) acc[co] = bias ? bias[co] : 0.0f;
    for (int t = 0; t < 27; ++t) {
        int dy, dx, dz;
        tap_off(t, dil, dy, dx, dz);
        const int yy = y + dy, xx = x + dx, zz = z + dz;
        const bool ok = yy >= 0 && yy < Y && xx >= 0 && xx < X && zz >= 0 && zz < Z;
        const TI *p = in + (((long)(ok ? yy : y) * X + (ok ? xx : x)) * Z + (ok ? zz : z)) * CI;
        const float *w = W + (long)t * CI * CO;
        for (int ci = 0; ci < CI; ++ci) {
            const float a = ok ? to_f(p[ci]) : 0.0f;
#pragma unroll
            for (int co = 0; co < CO; ++co) acc[co] = fmaf(w[ci * CO + co], a, acc[co]);
        }
    }
    if constexpr (std::is_same<TO, uint16_t>::value && CO % 8 == 0) {
        // bf16 outputs, 8 channels per 128-bit store
#pragma unroll
        for (int c8 = 0; c8 < CO; c8 += 8) {
            uint32_t q[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float v0 = acc[c8 + 2 * k], v1 = acc[c8 + 2 * k + 1];
                if (relu) {
                    v0 = fmaxf(v0, 0.0f);
                    v1 = fmaxf(v1, 0.0f);
                }
                uint16_t h0, h1;
                from_f(h0, v0);
                from_f(h1, v1);
                q[k] = (uint32_t)h0 | ((uint32_t)h1 << 16);
            }
            *(uint4 *)(out + v * CO + c8) = make_uint4(q[0], q[1], q[2], q[3]);
        }
    } else {
#pragma unroll
        for (int co = 0; co < CO; ++co) {
            float val = acc[co];
            if (relu) val = fmaxf(val, 0.0f);
            if (mask && !(mask[v * CO + co] > 0.0f)) val = 0.0f;
            from_f(out[v * CO + co], val);
        }
    }
}

// ---------------------------------------------------------------------------
// Layer 1 (2 -> 32 channels, dilation 1) of the inference path on MFMA: the
// 27 taps x 2 channels form K = 54 (padded to 64, two K = 32 steps) of
// out^T[co][voxel] = sum_k W1[co][k] in[k][voxel], k = 2 tap + channel.  A
// wave owns 64 slices of one (y, x) row, stages the 9 neighbour rows' slices
// [z0 - 1, z0 + 65) of both channels in LDS once (channel 0 times scal[0]),
// and builds each lane's B fragment (voxel n, k = 32 s + 8 (l/16) .. +7,
// i.e. 4 taps x 2 channels) from there, rounded to bf16.  A fragments
// (bf16 [2 s][2 co tiles][64 lanes][8], k >= 54 zero) come from the caches.
// Bias + ReLU, bf16 [V][32] out.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_conv_l1_mfma(const float2 *__restrict__ in,
                                                      const uint4 *__restrict__ Wf,
                                                      const float *__restrict__ bias,
                                                      float *__restrict__ scal,
                                                      uint16_t *__restrict__ out, int Y, int X,
                                                      int Z, const double *__restrict__ part,
                                                      int nb, long nvals)
{
    constexpr int SL = 64 + 2;
    __shared__ float2 seg[4][9 * SL];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const int zt = (Z + 63) / 64;
    const long task = (long)xcd_contiguous_block() * 4 + w;
    const bool valid = task < (long)Y * X * zt;
    const long row = valid ? task / zt : 0;
    const int z0 = valid ? (int)(task % zt) * 64 : 0;
    const int y = (int)(row / X), x = (int)(row % X);
    // the input's scale 1/mean(x): from scal, or (part) from the block sums
    // of x, every block adding them as k_inv_mean does and block 0 storing
    // both factors in scal for the later kernels
    float s0;
    if (part) {
        __shared__ double red[256];
        float mean;
        inv_mean_block(part, nb, nvals, red, s0, mean);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal[0] = s0;
            scal[1] = mean;
        }
    } else {
        s0 = scal[0];
    }
    for (int e = lane; e < 9 * SL; e += 64) {
        const int q = e / SL, k = e - q * SL;
        const int yy = y + q / 3 - 1, xx = x + q % 3 - 1, zz = z0 - 1 + k;
        const bool ok = valid && yy >= 0 && yy < Y && xx >= 0 && xx < X && zz >= 0 && zz < Z;
        float2 v = ok ? in[((long)yy * X + xx) * Z + zz] : make_float2(0.0f, 0.0f);
        v.x = v.x * s0;
        seg[w][e] = v;
    }
    __syncthreads();
    cfloatx4_t acc[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[c][j] = cfloatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        cbf16x8_t A[2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
            A[c] = __builtin_bit_cast(cbf16x8_t, Wf[(st * 2 + c) * 64 + lane]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cbf16x8_t B;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int t = 16 * st + 4 * g + h;   // tap (27 and up: padding)
                float2 v = make_float2(0.0f, 0.0f);
                if (t < 27) v = seg[w][(t / 3) * SL + 16 * j + n + t % 3];
                B[2 * h] = (__bf16)v.x;
                B[2 * h + 1] = (__bf16)v.y;
            }
#pragma unroll
            for (int c = 0; c < 2; ++c)
                acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[c], B, acc[c][j], 0, 0, 0);
        }
    }
    if (!valid) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int z = z0 + 16 * j + n;
        if (z >= Z) continue;
        const long v = row * Z + z;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int co = 16 * c + 4 * g;
            uint16_t h[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) from_f(h[i], fmaxf(acc[c][j][i] + bias[co + i], 0.0f));
            *(uint2 *)(out + v * 32 + co) = make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16),
                                                       (uint32_t)h[2] | ((uint32_t)h[3] << 16));
        }
    }
}

// ---------------------------------------------------------------------------
// Implicit-GEMM convolution of 32 input channels on MFMA (inference).
// out^T[co][voxel] = sum_tap W_tap[co][ci] in^T[ci][voxel + off]: per tap one
// K = 32 step; A = weights (bf16 [27][16 NCT co][32 ci], lane l holds
// co = l%16 (+16c), ci = 8(l/16)..+7), B = 16 voxels' channels (lane l:
// voxel l%16, ci = 8(l/16)..+7, 128 bits); C lane l, element i = (co
// 4(l/16)+i, voxel l%16).  A wave owns 64 consecutive slices of one (y, x)
// row.  For each of the 9 in-plane tap offsets the wave stages the
// neighbour row's slices [z0 - dil, z0 + 64 + dil) in LDS once (one
// coalesced pass), and the three slice taps of that offset read their B
// fragments from there: 9 instead of 27 global passes over the input.  A
// fragments (identical for every wave) come from the caches.
// NCT = 2: 32 outputs, bias + ReLU, bf16 [V][32]; NCT = 1 with COUT = 2:
// the linear output layer, fp32 [V][2] (rows 2..15 of A are zero).  ROWZ
// (64, 32 or 16) is the slices per wave: shorter rows give more waves for
// the same volume, longer rows less halo per slice.
// ---------------------------------------------------------------------------
constexpr int CNN_ROWZ = 64;   // the layer-1 kernel's rows
constexpr int CNN_MAXDIL = 1 << (CNN_LAYERS - 1);

// Orders a wave's own LDS writes and reads (a wave's LDS instructions are
// performed in issue order; this keeps the compiler from reordering them).
__device__ __forceinline__ void cnn_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NCT, int COUT, int ROWZ>
__global__ __launch_bounds__(256, 4) void k_conv_mfma_rows(const uint16_t *__restrict__ in,
                                                        const uint4 *__restrict__ Wg,
                                                        const float *__restrict__ bias,
                                                        void *__restrict__ outp, int Y, int X,
                                                        int Z, int dil, int relu)
{
    constexpr int WQ = 16 * NCT * 32 * 2 / 16;   // uint4 per tap
    extern __shared__ uint4 seg[];               // [4 waves][ROWZ + 2 dil][4]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    constexpr int NJ = ROWZ / 16;
    const int zt = (Z + ROWZ - 1) / ROWZ;
    const long task = (long)xcd_contiguous_block() * 4 + w;
    const long rows = (long)Y * X;
    const bool valid = task < rows * zt;
    const long row = valid ? task / zt : 0;
    const int z0 = valid ? (int)(task % zt) * ROWZ : 0;
    const int y = (int)(row / X), x = (int)(row % X);
    const int SL = ROWZ + 2 * dil;
    uint4 *my = seg + (long)w * SL * 4;
    cfloatx4_t acc[NCT][NJ];
#pragma unroll
    for (int c = 0; c < NCT; ++c)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[c][j] = cfloatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
    // the neighbour row segment of offset q, SL*4 <= PRE*64 uint4 per wave
    // (dil <= CNN_MAXDIL): loaded into registers one offset ahead, so its
    // latency overlaps the previous offset's products
    constexpr int PRE = ((ROWZ + 2 * CNN_MAXDIL) * 4 + 63) / 64;
    uint4 pre[PRE];
    auto fetch = [&](int q) {
        const int yy = y + dil * (q / 3 - 1), xx = x + dil * (q % 3 - 1);
        const bool rok = valid && yy >= 0 && yy < Y && xx >= 0 && xx < X;
        const uint16_t *src = in + ((long)(rok ? yy : 0) * X + (rok ? xx : 0)) * Z * 32;
#pragma unroll
        for (int k = 0; k < PRE; ++k) {
            const int e = lane + 64 * k;
            const int vz = z0 - dil + (e >> 2);
            pre[k] = (e < SL * 4 && rok && vz >= 0 && vz < Z)
                         ? *(const uint4 *)(src + (long)vz * 32 + 8 * (e & 3))
                         : make_uint4(0, 0, 0, 0);
        }
    };
    // LDS slot of voxel r's channel quarter c (8 channels, 16 bytes): the
    // quarters rotate by r/4, so the 16 lanes of a B-fragment read (16
    // consecutive voxels, one quarter) cover all 64 banks once
    auto slot = [](int r, int c) { return r * 4 + ((c + (r >> 2)) & 3); };
    fetch(0);
    for (int q = 0; q < 9; ++q) {
        // each wave stages and reads only its own segment: the wave's LDS
        // accesses are performed in order, so only the compiler must keep
        // them in program order (no block barrier)
        cnn_wave_sync();   // the previous offset's readers are done
#pragma unroll
        for (int k = 0; k < PRE; ++k) {
            const int e = lane + 64 * k;
            if (e < SL * 4) my[slot(e >> 2, e & 3)] = pre[k];
        }
        cnn_wave_sync();
        // the offset's A fragments before the next segment's loads: vector
        // loads retire in order, so waiting for A does not wait for the
        // prefetch
        cbf16x8_t A[3][NCT];
#pragma unroll
        for (int dzi = 0; dzi < 3; ++dzi)
#pragma unroll
            for (int c = 0; c < NCT; ++c)
                A[dzi][c] = __builtin_bit_cast(cbf16x8_t,
                                               Wg[(3 * q + dzi) * WQ + (16 * c + n) * 4 + g]);
        if (q + 1 < 9) fetch(q + 1);
#pragma unroll
        for (int dzi = 0; dzi < 3; ++dzi) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const cbf16x8_t B =
                    __builtin_bit_cast(cbf16x8_t, my[slot(16 * j + n + dzi * dil, g)]);
#pragma unroll
                for (int c = 0; c < NCT; ++c)
                    acc[c][j] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[dzi][c], B, acc[c][j], 0, 0, 0);
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int z = z0 + 16 * j + n;
        if (z >= Z) continue;
        const long v = row * Z + z;
        if constexpr (COUT == 2) {
            if (g == 0)
                *(float2 *)((float *)outp + v * 2) =
                    make_float2(acc[0][j][0] + bias[0], acc[0][j][1] + bias[1]);
        } else {
            uint16_t *out = (uint16_t *)outp;
#pragma unroll
            for (int c = 0; c < NCT; ++c) {
                const int co = 16 * c + 4 * g;
                uint16_t h[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float val = acc[c][j][i] + bias[co + i];
                    if (relu) val = fmaxf(val, 0.0f);
                    from_f(h[i], val);
                }
                *(uint2 *)(out + v * 32 + co) =
                    make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16),
                               (uint32_t)h[2] | ((uint32_t)h[3] << 16));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradients of one tap per blockIdx.y over a chunk of voxels:
// dW[t][ci][co] += sum_v in[v + off_t][ci] dout[v][co].  Threads own IB x OB
// blocks of (ci, co); 256/G voxel subgroups are reduced in LDS.
// ---------------------------------------------------------------------------
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_conv_dw(const float *__restrict__ in,
                                                 const float *__restrict__ dout,
                                                 float *__restrict__ dW, int Y, int X, int Z,
                                                 int dil, int chunk)
{
    constexpr int IB = CI < 4 ? CI : 4, OB = CO < 4 ? CO : 4;
    constexpr int G = (CI / IB) * (CO / OB);
    constexpr int NSG = 256 / G;
    __shared__ float red[NSG * CI * CO];
    const int t = blockIdx.y;
    int dy, dx, dz;
    tap_off(t, dil, dy, dx, dz);
    const int sg = threadIdx.x / G, e = threadIdx.x % G;
    const int ci0 = (e / (CO / OB)) * IB, co0 = (e % (CO / OB)) * OB;
    float acc[IB][OB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int o = 0; o < OB; ++o) acc[i][o] = 0.0f;
    const long V = (long)Y * X * Z;
    const long v0 = (long)blockIdx.x * chunk;
    const long v1 = v0 + chunk < V ? v0 + chunk : V;
    for (long v = v0 + sg; v < v1; v += NSG) {
        const int z = (int)(v % Z);
        const long r = v / Z;
        const int x = (int)(r % X), y = (int)(r / X);
        const int yy = y + dy, xx = x + dx, zz = z + dz;
        if (yy < 0 || yy >= Y || xx < 0 || xx >= X || zz < 0 || zz >= Z) continue;
        const float *pi = in + (((long)yy * X + xx) * Z + zz) * CI + ci0;
        const float *po = dout + v * CO + co0;
        float a[IB], gg[OB];
#pragma unroll
        for (int i = 0; i < IB; ++i) a[i] = pi[i];
#pragma unroll
        for (int o = 0; o < OB; ++o) gg[o] = po[o];
#pragma unroll
        for (int i = 0; i < IB; ++i)
#pragma unroll
            for (int o = 0; o < OB; ++o) acc[i][o] = fmaf(a[i], gg[o], acc[i][o]);
    }
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int o = 0; o < OB; ++o) red[sg * CI * CO + (ci0 + i) * CO + co0 + o] = acc[i][o];
    __syncthreads();
    for (int k = threadIdx.x; k < CI * CO; k += 256) {
        float s = 0.0f;
        for (int q = 0; q < NSG; ++q) s += red[q * CI * CO + k];
        atomicAdd(dW + (long)t * CI * CO + k, s);
    }
}

// Bias gradients: db[co] += sum_v dout[v][co].
template <int CO>
__global__ __launch_bounds__(256) void k_bias_grad(const float *__restrict__ dout, long V,
                                                   float *__restrict__ db, int chunk)
{
    constexpr int NVL = 256 / CO;
    __shared__ float red[256];
    const int co = threadIdx.x % CO, vl = threadIdx.x / CO;
    const long v0 = (long)blockIdx.x * chunk;
    const long v1 = v0 + chunk < V ? v0 + chunk : V;
    float s = 0.0f;
    for (long v = v0 + vl; v < v1; v += NVL) s += dout[v * CO + co];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < CO) {
        float t = 0.0f;
        for (int q = 0; q < NVL; ++q) t += red[q * CO + threadIdx.x];
        atomicAdd(db + threadIdx.x, t);
    }
}

// Squared-error loss and its gradient gscale*(out - tgt) (gscale = 2/count).
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ out,
                                             const float *__restrict__ tgt, long n, float gscale,
                                             float *__restrict__ grad, double *__restrict__ loss)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const float d = out[i] - tgt[i];
        grad[i] = gscale * d;
        s += (double)d * (double)d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Adam (Kingma & Ba 2015) on FP32 master weights; clears the gradient.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2)
{
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    g[i] = 0.0f;
}

// Input-gradient weights: Wt[26 - t][co][ci] = W[t][ci][co] (layout [27][CI'][CO']
// of the convolution with CI' = CO, CO' = CI).
__global__ void k_flip(const float *__restrict__ W, float *__restrict__ Wt, int CI, int CO)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 27 * CI * CO) return;
    const int t = i / (CI * CO), r = i % (CI * CO), ci = r / CO, co = r % CO;
    Wt[((long)(26 - t) * CO + co) * CI + ci] = W[i];
}

// BF16 copy of master weights [27][CI][CO] in MFMA order [27][COP][CI], rows
// co >= CO of the padded COP zero.
__global__ void k_to_bf16_t(const float *__restrict__ W, uint16_t *__restrict__ Wb, int CI,
                            int CO, int COP)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 27 * CI * COP) return;
    const int t = i / (CI * COP), r = i % (CI * COP), co = r / CI, ci = r % CI;
    const float v = co < CO ? W[((long)t * CI + ci) * CO + co] : 0.0f;
    from_f(Wb[i], v);
}

// Layer 1's A fragments from its master weights [27][2][32]: entry
// ((s * 2 + c) * 64 + l) * 8 + i is W[tap][ch][co] for co = 16 c + l % 16,
// k = 32 s + 8 (l / 16) + i = 2 tap + ch (zero for tap >= 27), in bf16.
__global__ void k_l1_fragments(const float *__restrict__ W, uint16_t *__restrict__ Wf)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= 2 * 2 * 64 * 8) return;
    const int i = e % 8, l = (e / 8) % 64, c = (e / 512) % 2, st = e / 1024;
    const int co = 16 * c + l % 16, k = 32 * st + 8 * (l / 16) + i;
    const int t = k / 2, ch = k % 2;
    from_f(Wf[e], t < 27 ? W[(t * 2 + ch) * 32 + co] : 0.0f);
}

// ---------------------------------------------------------------------------
// Source volumes (engine layout [y][x][z]) from the network output: trilinear
// upsampling by 2 (fine sample j at coarse coordinate j/2 - 1/4, clamped).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void up_coord(int j, int n, int &i0, int &i1, float &w)
{
    const float u = 0.5f * (float)j - 0.25f;
    if (u < 0.0f) {
        i0 = i1 = 0;
        w = 0.0f;
        return;
    }
    i0 = (int)u;
    w = u - (float)i0;
    i1 = i0 + 1 < n ? i0 + 1 : n - 1;
    if (i0 > n - 1) i0 = n - 1;
}

// Source volumes out: g0[i], g1[i] in FP32, or (gp) both as bf16 in one
// 32-bit word, g0 in the low half.  The bf16 values are rounded
// stochastically with a fixed per-voxel dither (the 16 bits below the bf16
// mantissa plus a hash of the voxel index, truncated): the rounding errors
// have zero mean and do not repeat across voxels of equal value (an early,
// nearly uniform estimate), so the depth sums of the projector average them
// out instead of adding them up.  The hash is the lowbias32 integer mixer.
__device__ __forceinline__ uint32_t dither_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t bf16_dithered(float v, uint32_t r16)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, v);
    // finite values only (|v| below the largest bf16): the carry into the
    // exponent is the correct rounding up to the next binade
    return (u + r16) >> 16;
}
__device__ __forceinline__ void store_sources(float *g0, float *g1, uint32_t *gp, long i, float a,
                                              float b)
{
    if (gp) {
        const uint32_t h = dither_hash((uint32_t)i);
        gp[i] = bf16_dithered(a, h & 0xffffu) | (bf16_dithered(b, h >> 16) << 16);
    } else {
        g0[i] = a;
        g1[i] = b;
    }
}

__global__ __launch_bounds__(256) void k_cnn_source(const float *__restrict__ o,
                                                    const float *__restrict__ x,
                                                    const float *__restrict__ mu, int N, int Nz,
                                                    const float *__restrict__ scal, float kappa,
                                                    float fDens, float cntr, float primary,
                                                    float *__restrict__ g0,
                                                    float *__restrict__ g1,
                                                    uint32_t *__restrict__ gp)
{
    const int n = N * N * Nz;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int z = i % Nz;
    const int r = i / Nz;
    const int xx = r % N, y = r / N;
    const int Y = N / 2, X = N / 2, Z = Nz / 2;
    int y0, y1, x0, x1, z0, z1;
    float wy, wx, wz;
    up_coord(y, Y, y0, y1, wy);
    up_coord(xx, X, x0, x1, wx);
    up_coord(z, Z, z0, z1, wz);
    float c[2] = {0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const float wgt = (a ? wy : 1.0f - wy) * (b ? wx : 1.0f - wx) * (d ? wz : 1.0f - wz);
                const int q = (((a ? y1 : y0) * X + (b ? x1 : x0)) * Z + (d ? z1 : z0)) * 2;
                c[0] = fmaf(wgt, o[q], c[0]);
                c[1] = fmaf(wgt, o[q + 1], c[1]);
            }
    const float sx = scal[1];
    const float dens = mu ? mu[i] * fDens : 0.0f;
    const float xi = x[i];
    store_sources(g0, g1, gp, i, primary * xi + (sx * c[0] + cntr * xi) * dens,
                  (sx / kappa) * c[1] * dens);
}

__global__ __launch_bounds__(256) void k_exact_source(const float *__restrict__ E0,
                                                      const float *__restrict__ E1,
                                                      const float *__restrict__ x,
                                                      const float *__restrict__ mu, long n,
                                                      float fMu, float fDens, float cntr,
                                                      float primary, float *__restrict__ g0,
                                                      float *__restrict__ g1,
                                                      uint32_t *__restrict__ gp)
{
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const float dens = mu ? mu[i] * fDens : 0.0f;
    const float xi = x[i], e0 = E0[i];
    store_sources(g0, g1, gp, i, primary * xi + (e0 + cntr * xi) * dens,
                  (fMu * e0 - E1[i]) * dens);
}

void cnn_exact_sources(const float *E0, const float *E1, const float *x, const float *mu, int N,
                       int Nz, const CnnPhys &ph, float primary, float *g0, float *g1,
                       hipStream_t s, uint32_t *packed)
{
    const long n = (long)N * N * Nz;
    hipLaunchKernelGGL(k_exact_source, dim3(cdiv_u(n, 256)), dim3(256), 0, s, E0, E1, x, mu, n,
                       ph.fMu, ph.fDens, ph.cntr, primary, g0, g1, packed);
    CNN_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Training phantoms: ellipsoids rotated about the slice axis.
// ---------------------------------------------------------------------------
struct Ellipsoid {
    float cy, cx, cz, iay, iax, iaz, c, s, val;
};

__global__ __launch_bounds__(256) void k_phantom(const Ellipsoid *__restrict__ e, int ne, int N,
                                                 int Nz, float *__restrict__ out)
{
    const long n = (long)N * N * Nz;
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const int z = (int)(i % Nz);
    const long r = i / Nz;
    const int x = (int)(r % N), y = (int)(r / N);
    float v = 0.0f;
    for (int k = 0; k < ne; ++k) {
        const float dy = (float)y + 0.5f - e[k].cy, dx = (float)x + 0.5f - e[k].cx;
        const float dz = (float)z + 0.5f - e[k].cz;
        const float u = (dy * e[k].c + dx * e[k].s) * e[k].iay;
        const float w = (-dy * e[k].s + dx * e[k].c) * e[k].iax;
        const float q = dz * e[k].iaz;
        if (u * u + w * w + q * q <= 1.0f) v += e[k].val;
    }
    out[i] = fmaxf(v, 0.0f);
}

// Gaussian blur along one axis (0 = y, 1 = x, 2 = z) of [y][x][z], radius 3 sigma.
__global__ __launch_bounds__(256) void k_blur(const float *__restrict__ in, float *__restrict__ out,
                                              int N, int Nz, int axis, float sigma)
{
    const long n = (long)N * N * Nz;
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const int z = (int)(i % Nz);
    const long r = i / Nz;
    const int x = (int)(r % N), y = (int)(r / N);
    const int pos = axis == 0 ? y : axis == 1 ? x : z;
    const int len = axis == 2 ? Nz : N;
    const long stride = axis == 0 ? (long)N * Nz : axis == 1 ? Nz : 1;
    const int R = (int)ceilf(3.0f * sigma);
    float s = 0.0f, ws = 0.0f;
    for (int k = -R; k <= R; ++k) {
        const int p = pos + k;
        const float wgt = expf(-0.5f * (float)(k * k) / (sigma * sigma));
        ws += wgt;
        if (p >= 0 && p < len) s += wgt * in[i + (long)k * stride];
    }
    out[i] = s / ws;
}

// Multiplicative noise x *= 1 + eps * sqrt(3) (2u - 1), u uniform from a hash.
__global__ __launch_bounds__(256) void k_noise(float *__restrict__ x, long n, float eps,
                                               uint32_t seed)
{
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    uint32_t h = (uint32_t)i * 0x9E3779B1u ^ seed;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    const float u = (float)(h >> 8) * (1.0f / 16777216.0f);
    x[i] = fmaxf(x[i] * (1.0f + eps * 1.7320508f * (2.0f * u - 1.0f)), 0.0f);
}

// ---------------------------------------------------------------------------
// ScatterCnn
// ---------------------------------------------------------------------------
template <typename T>
static T *cnn_alloc(size_t n)
{
    T *p = nullptr;
    if (n) CNN_CHECK(hipMalloc((void **)&p, n * sizeof(T)));
    return p;
}

template <typename T>
static void cnn_free(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

static long wcount(int l) { return 27L * cnn_cin(l) * cnn_cout(l); }

void ScatterCnn::alloc_state()
{
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const long nw = wcount(l), nb = cnn_cout(l);
        W_[l] = cnn_alloc<float>(nw);
        b_[l] = cnn_alloc<float>(nb);
        gW_[l] = cnn_alloc<float>(nw);
        gb_[l] = cnn_alloc<float>(nb);
        mW_[l] = cnn_alloc<float>(nw);
        vW_[l] = cnn_alloc<float>(nw);
        mb_[l] = cnn_alloc<float>(nb);
        vb_[l] = cnn_alloc<float>(nb);
        Wbf_[l] = cnn_alloc<uint16_t>(27L * cnn_cin(l) * (cnn_cout(l) < 16 ? 16 : cnn_cout(l)));
        CNN_CHECK(hipMemset(gW_[l], 0, nw * 4));
        CNN_CHECK(hipMemset(gb_[l], 0, nb * 4));
        CNN_CHECK(hipMemset(mW_[l], 0, nw * 4));
        CNN_CHECK(hipMemset(vW_[l], 0, nw * 4));
        CNN_CHECK(hipMemset(mb_[l], 0, nb * 4));
        CNN_CHECK(hipMemset(vb_[l], 0, nb * 4));
    }
    Wt_ = cnn_alloc<float>(27L * 32 * 32);
    Wl1_ = cnn_alloc<uint16_t>(2 * 2 * 64 * 8);
    dsum_ = cnn_alloc<double>(1);
    one_ = cnn_alloc<float>(2);
    const float ones[2] = {1.0f, 1.0f};
    CNN_CHECK(hipMemcpy(one_, ones, sizeof ones, hipMemcpyHostToDevice));
}

ScatterCnn::ScatterCnn(float kappa, unsigned seed) : kappa_(kappa)
{
    alloc_state();
    std::mt19937 rng(seed);
    std::vector<float> packed(cnn_num_params());
    long o = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * cnn_cin(l))));
        for (long k = 0; k < wcount(l); ++k) packed[o++] = nd(rng);
        for (int k = 0; k < cnn_cout(l); ++k) packed[o++] = 0.0f;
    }
    setParams(packed);
}

ScatterCnn::ScatterCnn(float kappa, const std::vector<float> &packed) : kappa_(kappa)
{
    alloc_state();
    setParams(packed);
    prepareInference();
}

void ScatterCnn::prepareInference()
{
    upload_bf16_weights(nullptr);
    CNN_CHECK(hipDeviceSynchronize());
}

CnnWork::CnnWork(int N, int Nz)
{
    if (N < 2 || Nz < 2 || N % 2 || Nz % 2)
        fail(ETYPE_ILLEGAL_VALUE, "scatter CNN needs even volume sizes (%d, %d)", N, Nz);
    V = (long)(N / 2) * (N / 2) * (Nz / 2);
    in = cnn_alloc<float>(V * 2);
    out = cnn_alloc<float>(V * 2);
    for (auto &a : actb) a = cnn_alloc<uint16_t>(V * 32);
    nb = cdiv_u(V, 256);
    psum = cnn_alloc<double>(nb);
    scal = cnn_alloc<float>(2);
}

CnnWork::~CnnWork()
{
    cnn_free(in);
    cnn_free(out);
    for (auto &a : actb) cnn_free(a);
    cnn_free(psum);
    cnn_free(scal);
}

ScatterCnn::~ScatterCnn()
{
    for (int l = 0; l < CNN_LAYERS; ++l) {
        cnn_free(W_[l]);
        cnn_free(b_[l]);
        cnn_free(gW_[l]);
        cnn_free(gb_[l]);
        cnn_free(mW_[l]);
        cnn_free(vW_[l]);
        cnn_free(mb_[l]);
        cnn_free(vb_[l]);
        cnn_free(Wbf_[l]);
    }
    for (auto &a : act_) cnn_free(a);
    for (auto &a : grad_) cnn_free(a);
    for (auto &a : actb_) cnn_free(a);
    cnn_free(Wt_);
    cnn_free(Wl1_);
    cnn_free(dsum_);
    cnn_free(one_);
}

// packed (oracle) order per layer: W [27][Cout][Cin], b [Cout]
void ScatterCnn::setParams(const std::vector<float> &packed)
{
    if ((long)packed.size() != cnn_num_params())
        fail(ETYPE_ILLEGAL_VALUE, "scatter CNN: %ld parameters given, %ld expected",
             (long)packed.size(), cnn_num_params());
    long o = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const int CI = cnn_cin(l), CO = cnn_cout(l);
        std::vector<float> w(wcount(l));
        for (int t = 0; t < 27; ++t)
            for (int co = 0; co < CO; ++co)
                for (int ci = 0; ci < CI; ++ci)
                    w[((long)t * CI + ci) * CO + co] = packed[o + ((long)t * CO + co) * CI + ci];
        o += wcount(l);
        CNN_CHECK(hipMemcpy(W_[l], w.data(), w.size() * 4, hipMemcpyHostToDevice));
        CNN_CHECK(hipMemcpy(b_[l], packed.data() + o, CO * 4, hipMemcpyHostToDevice));
        o += CO;
    }
    bfDirty_ = true;
}

void ScatterCnn::getParams(std::vector<float> &packed) const
{
    packed.assign(cnn_num_params(), 0.0f);
    long o = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const int CI = cnn_cin(l), CO = cnn_cout(l);
        std::vector<float> w(wcount(l));
        CNN_CHECK(hipMemcpy(w.data(), W_[l], w.size() * 4, hipMemcpyDeviceToHost));
        for (int t = 0; t < 27; ++t)
            for (int co = 0; co < CO; ++co)
                for (int ci = 0; ci < CI; ++ci)
                    packed[o + ((long)t * CO + co) * CI + ci] = w[((long)t * CI + ci) * CO + co];
        o += wcount(l);
        CNN_CHECK(hipMemcpy(packed.data() + o, b_[l], CO * 4, hipMemcpyDeviceToHost));
        o += CO;
    }
}

void ScatterCnn::getGrads(std::vector<float> &packed) const
{
    packed.assign(cnn_num_params(), 0.0f);
    long o = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const int CI = cnn_cin(l), CO = cnn_cout(l);
        std::vector<float> w(wcount(l));
        CNN_CHECK(hipMemcpy(w.data(), gW_[l], w.size() * 4, hipMemcpyDeviceToHost));
        for (int t = 0; t < 27; ++t)
            for (int co = 0; co < CO; ++co)
                for (int ci = 0; ci < CI; ++ci)
                    packed[o + ((long)t * CO + co) * CI + ci] = w[((long)t * CI + ci) * CO + co];
        o += wcount(l);
        CNN_CHECK(hipMemcpy(packed.data() + o, gb_[l], CO * 4, hipMemcpyDeviceToHost));
        o += CO;
    }
}

void ScatterCnn::ensure_work(long V)
{
    if (V <= workV_) return;
    for (auto &a : act_) cnn_free(a);
    for (auto &a : grad_) cnn_free(a);
    for (auto &a : actb_) cnn_free(a);
    for (int l = 0; l <= CNN_LAYERS; ++l)
        act_[l] = cnn_alloc<float>(V * (l == 0 ? 2 : l == CNN_LAYERS ? 2 : 32));
    for (auto &a : grad_) a = cnn_alloc<float>(V * 32);
    for (auto &a : actb_) a = cnn_alloc<uint16_t>(V * 32);
    workV_ = V;
}

// Inference weights: BF16 MFMA operands of the 32-input layers (the output
// layer padded to 16 rows) and the BF16-rounded FP32 weights of layer 1.
void ScatterCnn::upload_bf16_weights(hipStream_t s)
{
    if (!bfDirty_) return;
    for (int l = 1; l < CNN_LAYERS; ++l) {
        const int COP = cnn_cout(l) < 16 ? 16 : cnn_cout(l);
        const long n = 27L * cnn_cin(l) * COP;
        hipLaunchKernelGGL(k_to_bf16_t, dim3(cdiv_u(n, 256)), dim3(256), 0, s, W_[l], Wbf_[l],
                           cnn_cin(l), cnn_cout(l), COP);
    }
    static_assert(cnn_cin(0) == 2 && cnn_cout(0) == 32, "layer 1 fragments");
    hipLaunchKernelGGL(k_l1_fragments, dim3(8), dim3(256), 0, s, W_[0], Wl1_);
    CNN_CHECK(hipGetLastError());
    bfDirty_ = false;
}

// Network forward: FP32 throughout (training arithmetic), or the inference
// arithmetic (bf16 weights and activations on MFMA, FP32 accumulation).
void ScatterCnn::forward(const float *in, int Y, int X, int Z, float *out, bool bf16,
                         hipStream_t s)
{
    const long V = (long)Y * X * Z;
    ensure_work(V);
    const dim3 grid(cdiv_u(V, 256)), blk(256);
    if (!bf16) {
        const float *a = in;
        for (int l = 0; l < CNN_LAYERS; ++l) {
            float *o = l == CNN_LAYERS - 1 ? out : act_[l + 1];
            const int relu = l < CNN_LAYERS - 1;
            if (cnn_cout(l) == 32)
                hipLaunchKernelGGL((k_conv_valu<32, float, float>), grid, blk, 0, s, a, W_[l],
                                   b_[l], nullptr, o, Y, X, Z, cnn_cin(l), cnn_dil(l), relu);
            else
                hipLaunchKernelGGL((k_conv_valu<2, float, float>), grid, blk, 0, s, a, W_[l],
                                   b_[l], nullptr, o, Y, X, Z, cnn_cin(l), cnn_dil(l), relu);
            a = o;
        }
        CNN_CHECK(hipGetLastError());
        return;
    }
    upload_bf16_weights(s);
    forward_bf16(in, Y, X, Z, out, actb_[0], actb_[1], nullptr, rowz_, s);
}

// Inference arithmetic, activations in the given buffers ([V][32] bf16 each);
// channel 0 of the input is multiplied by scal[0] (1 for normalised input).
void ScatterCnn::forward_bf16(const float *in, int Y, int X, int Z, float *out, uint16_t *a0,
                              uint16_t *a1, float *scal, int rowz, hipStream_t s,
                              const double *part, int nb, long n)
{
    const dim3 blk(256);
    const long tasks = (long)Y * X * ((Z + CNN_ROWZ - 1) / CNN_ROWZ);
    hipLaunchKernelGGL(k_conv_l1_mfma, dim3(cdiv_u(tasks, 4)), blk, 0, s, (const float2 *)in,
                       (const uint4 *)Wl1_, b_[0], scal ? scal : one_, a0, Y, X, Z, part, nb, n);
    static_assert(cnn_dil(CNN_LAYERS - 1) == CNN_MAXDIL, "largest dilation");
    switch (rowz) {
    case 64: mfma_layers<64>(Y, X, Z, out, a0, a1, s); break;
    case 32: mfma_layers<32>(Y, X, Z, out, a0, a1, s); break;
    case 16: mfma_layers<16>(Y, X, Z, out, a0, a1, s); break;
    default: fail(ETYPE_ILLEGAL_VALUE, "scatter CNN row length %d (64, 32 or 16)", rowz);
    }
    CNN_CHECK(hipGetLastError());
}

// Layers 2-4 with rows of ROWZ slices per wave (a0 holds layer 1's output).
template <int ROWZ>
void ScatterCnn::mfma_layers(int Y, int X, int Z, float *out, uint16_t *a0, uint16_t *a1,
                             hipStream_t s)
{
    const dim3 blk(256);
    const long tasks = (long)Y * X * ((Z + ROWZ - 1) / ROWZ);
    const dim3 rgrid(cdiv_u(tasks, 4));
    auto lds = [](int dil) { return (size_t)4 * (ROWZ + 2 * dil) * 4 * sizeof(uint4); };
    hipLaunchKernelGGL((k_conv_mfma_rows<2, 32, ROWZ>), rgrid, blk, lds(cnn_dil(1)), s, a0,
                       (const uint4 *)Wbf_[1], b_[1], (void *)a1, Y, X, Z, cnn_dil(1), 1);
    hipLaunchKernelGGL((k_conv_mfma_rows<2, 32, ROWZ>), rgrid, blk, lds(cnn_dil(2)), s, a1,
                       (const uint4 *)Wbf_[2], b_[2], (void *)a0, Y, X, Z, cnn_dil(2), 1);
    hipLaunchKernelGGL((k_conv_mfma_rows<1, 2, ROWZ>), rgrid, blk, lds(cnn_dil(3)), s, a0,
                       (const uint4 *)Wbf_[3], b_[3], (void *)out, Y, X, Z, cnn_dil(3), 0);
}

void ScatterCnn::setRowLength(int rowz)
{
    if (rowz != 64 && rowz != 32 && rowz != 16)
        fail(ETYPE_ILLEGAL_VALUE, "scatter CNN row length %d (64, 32 or 16)", rowz);
    rowz_ = rowz;
}

void ScatterCnn::zeroGrads(hipStream_t s)
{
    for (int l = 0; l < CNN_LAYERS; ++l) {
        CNN_CHECK(hipMemsetAsync(gW_[l], 0, wcount(l) * 4, s));
        CNN_CHECK(hipMemsetAsync(gb_[l], 0, cnn_cout(l) * 4, s));
    }
}

template <int CI, int CO>
static void launch_dw(const float *in, const float *dout, float *dW, int Y, int X, int Z,
                      int dil, hipStream_t s)
{
    const long V = (long)Y * X * Z;
    const int chunk = 4096;
    hipLaunchKernelGGL((k_conv_dw<CI, CO>), dim3(cdiv_u(V, chunk), 27), dim3(256), 0, s, in, dout,
                       dW, Y, X, Z, dil, chunk);
}

double ScatterCnn::lossGrad(const float *in, const float *tgt, int Y, int X, int Z, float gscale,
                            hipStream_t s)
{
    const long V = (long)Y * X * Z;
    ensure_work(V);
    const dim3 grid(cdiv_u(V, 256)), blk(256);
    // forward, keeping the activations
    CNN_CHECK(hipMemcpyAsync(act_[0], in, V * 2 * 4, hipMemcpyDeviceToDevice, s));
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const int relu = l < CNN_LAYERS - 1;
        if (cnn_cout(l) == 32)
            hipLaunchKernelGGL((k_conv_valu<32, float, float>), grid, blk, 0, s, act_[l], W_[l],
                               b_[l], nullptr, act_[l + 1], Y, X, Z, cnn_cin(l), cnn_dil(l), relu);
        else
            hipLaunchKernelGGL((k_conv_valu<2, float, float>), grid, blk, 0, s, act_[l], W_[l],
                               b_[l], nullptr, act_[l + 1], Y, X, Z, cnn_cin(l), cnn_dil(l), relu);
    }
    CNN_CHECK(hipMemsetAsync(dsum_, 0, sizeof(double), s));
    const long nout = V * 2;
    hipLaunchKernelGGL(k_mse, dim3(std::min<long>(1024, cdiv_u(nout, 256))), blk, 0, s,
                       act_[CNN_LAYERS], tgt, nout, gscale, grad_[0], dsum_);
    // backward: grad_[cur] holds d(loss)/d(output of layer l) (ReLU mask applied)
    int cur = 0;
    for (int l = CNN_LAYERS - 1; l >= 0; --l) {
        const int CI = cnn_cin(l), CO = cnn_cout(l), dil = cnn_dil(l);
        const float *g = grad_[cur];
        if (CI == 32 && CO == 32) launch_dw<32, 32>(act_[l], g, gW_[l], Y, X, Z, dil, s);
        else if (CI == 32) launch_dw<32, 2>(act_[l], g, gW_[l], Y, X, Z, dil, s);
        else launch_dw<2, 32>(act_[l], g, gW_[l], Y, X, Z, dil, s);
        if (CO == 32)
            hipLaunchKernelGGL((k_bias_grad<32>), dim3(cdiv_u(V, 4096)), blk, 0, s, g, V, gb_[l], 4096);
        else
            hipLaunchKernelGGL((k_bias_grad<2>), dim3(cdiv_u(V, 4096)), blk, 0, s, g, V, gb_[l], 4096);
        if (l == 0) break;
        // input gradient = convolution with flipped, transposed weights,
        // masked by the ReLU of layer l-1 (its output act_[l] > 0)
        hipLaunchKernelGGL(k_flip, dim3(cdiv_u(27L * CI * CO, 256)), blk, 0, s, W_[l], Wt_, CI, CO);
        float *gn = grad_[1 - cur];   // CI (= 32 for l > 0) channels
        hipLaunchKernelGGL((k_conv_valu<32, float, float>), grid, blk, 0, s, g, Wt_, nullptr,
                           act_[l], gn, Y, X, Z, CO, dil, 0);
        cur = 1 - cur;
    }
    CNN_CHECK(hipGetLastError());
    double loss = 0;
    CNN_CHECK(hipMemcpyAsync(&loss, dsum_, sizeof(double), hipMemcpyDeviceToHost, s));
    CNN_CHECK(hipStreamSynchronize(s));
    return loss / (double)nout;
}

void ScatterCnn::adamStep(float lr, hipStream_t s)
{
    ++step_;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - std::pow(b1, (float)step_), bc2 = 1.0f - std::pow(b2, (float)step_);
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const long nw = wcount(l), nb = cnn_cout(l);
        hipLaunchKernelGGL(k_adam, dim3(cdiv_u(nw, 256)), dim3(256), 0, s, W_[l], gW_[l], mW_[l],
                           vW_[l], nw, lr, b1, b2, eps, bc1, bc2);
        hipLaunchKernelGGL(k_adam, dim3(cdiv_u(nb, 256)), dim3(256), 0, s, b_[l], gb_[l], mb_[l],
                           vb_[l], nb, lr, b1, b2, eps, bc1, bc2);
    }
    CNN_CHECK(hipGetLastError());
    bfDirty_ = true;
}

static void mean_scale(const float *x, long n, double *&psum, long &nb, float *scal, hipStream_t s)
{
    const long want = std::min<long>(1024, cdiv_u(n, 256));
    if (want > nb) {
        cnn_free(psum);
        psum = cnn_alloc<double>(want);
        nb = want;
    }   // (a CnnWork is made with enough partial sums: no allocation here)
    hipLaunchKernelGGL(k_sum_partial, dim3((unsigned)want), dim3(256), 0, s, x, n, psum);
    hipLaunchKernelGGL(k_inv_mean, dim3(1), dim3(256), 0, s, psum, (int)want, n, scal);
}

void ScatterCnn::sources(const float *x, const float *mu, int N, int Nz, const CnnPhys &ph,
                         float primary, float *g0, float *g1, CnnWork &w, hipStream_t s,
                         uint32_t *packed)
{
    const int Y = N / 2, X = N / 2, Z = Nz / 2;
    const long V = (long)Y * X * Z, n = (long)N * N * Nz;
    if (N % 2 || Nz % 2 || V != w.V)
        fail(ETYPE_ILLEGAL_VALUE, "scatter CNN work buffers made for another size");
    if (n > INT_MAX)   // the source kernels index in 32 bits
        fail(ETYPE_ILLEGAL_VALUE, "scatter CNN volume too large (%ld voxels)", n);
    if (bfDirty_) fail(ETYPE_ILLEGAL_VALUE, "scatter CNN inference weights not prepared");
    const unsigned pb = cdiv_u(V, 256);
    if (pb > w.nb) fail(ETYPE_ILLEGAL_VALUE, "scatter CNN work buffers too small");
    hipLaunchKernelGGL(k_cnn_pool, dim3(pb), dim3(256), 0, s, x, mu, N, Nz, ph.fDens, w.in, w.psum);
    // layer 1 takes mean(x) from the pooling's block sums and stores it
    forward_bf16(w.in, Y, X, Z, w.out, w.actb[0], w.actb[1], w.scal, w.rowz, s, w.psum,
                 (int)pb, n);
    hipLaunchKernelGGL(k_cnn_source, dim3(cdiv_u(n, 256)), dim3(256), 0, s, w.out, x, mu, N, Nz,
                       w.scal, kappa_, ph.fDens, ph.cntr, primary, g0, g1, packed);
    CNN_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Training on ESSE-generated pairs
// ---------------------------------------------------------------------------
namespace {

struct PhantomGen {
    std::mt19937 rng;
    int N, Nz;
    float pix;
    PhantomGen(unsigned seed, int n, int nz, float p) : rng(seed), N(n), Nz(nz), pix(p) {}
    float U(float a, float b) { return std::uniform_real_distribution<float>(a, b)(rng); }
    Ellipsoid ell(float cy, float cx, float cz, float ay, float ax, float az, float ang, float v)
    {
        return Ellipsoid{cy, cx, cz, 1.0f / ay, 1.0f / ax, 1.0f / az, std::cos(ang), std::sin(ang), v};
    }
    // activity ellipsoids and attenuation ellipsoids (mu per pixel)
    void make(std::vector<Ellipsoid> &act, std::vector<Ellipsoid> &mu)
    {
        act.clear();
        mu.clear();
        const float h = 0.5f * N, hz = 0.5f * Nz;
        const float by = U(0.28f, 0.44f) * N, bx = U(0.28f, 0.44f) * N, bz = U(0.3f, 0.5f) * Nz;
        const float cy = h + U(-0.05f, 0.05f) * N, cx = h + U(-0.05f, 0.05f) * N;
        const float cz = hz + U(-0.05f, 0.05f) * Nz, ang = U(0.0f, 3.14159265f);
        act.push_back(ell(cy, cx, cz, by, bx, bz, ang, 1.0f));
        mu.push_back(ell(cy, cx, cz, by, bx, bz, ang, 0.15f * pix));
        const int nin = 3 + (int)(rng() % 8);
        for (int k = 0; k < nin; ++k) {
            const float r = U(0.03f, 0.16f) * N;
            const float a = U(0.0f, 0.6f), b = U(0.0f, 6.28318f);
            act.push_back(ell(cy + a * by * std::sin(b), cx + a * bx * std::cos(b),
                              cz + U(-0.6f, 0.6f) * bz, r * U(0.6f, 1.4f), r * U(0.6f, 1.4f),
                              r * U(0.6f, 1.4f), U(0.0f, 3.14159f), U(-1.0f, 7.0f)));
        }
        // lung-like low-density region
        const float lr = U(0.08f, 0.16f) * N;
        mu.push_back(ell(cy + U(-0.3f, 0.3f) * by, cx + U(-0.3f, 0.3f) * bx, cz, lr, lr * U(0.6f, 1.2f),
                         U(0.2f, 0.4f) * Nz, U(0.0f, 3.14f), -0.10f * pix));
    }
};

}  // namespace

ScatterCnn *ScatterCnn::train(const EsseModel &m, const CnnTrainParms &p, CnnTrainStats &st)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (!m.enabled) fail(ETYPE_ILLEGAL_VALUE, "scatter CNN: the scatter kernel has scatFac 0");
    if (m.nterms < 2) fail(ETYPE_ILLEGAL_VALUE, "scatter CNN: the ESSE model needs 2 terms");
    if (m.c != 1) fail(ETYPE_ILLEGAL_VALUE, "scatter CNN: collapse factor 1 only");
    if (p.N % 2 || p.Nz % 2) fail(ETYPE_ILLEGAL_VALUE, "scatter CNN: even sizes only");
    if (p.pairs <= p.val || p.val < 1 || p.batch < 1 || p.epochs < 0)
        fail(ETYPE_ILLEGAL_VALUE, "scatter CNN: bad training sizes");
    const int N = p.N, Nz = p.Nz;
    const long n = (long)N * N * Nz;
    const int Y = N / 2, X = N / 2, Z = Nz / 2;
    const long V = (long)Y * X * Z;
    hipStream_t s;
    CNN_CHECK(hipStreamCreate(&s));
    const int GB = 8;   // phantoms per ESSE batch
    EsseDev esse(m, GB, s, true);
    const float fMu = m.mu0 - m.muavg;
    const float fDens = (1.0f / m.mu0) / m.pixsize;

    float *xs = cnn_alloc<float>(n * GB), *mus = cnn_alloc<float>(n * GB), *tmp = cnn_alloc<float>(n);
    float *terms = cnn_alloc<float>(n * GB * m.nterms);
    float *scal = cnn_alloc<float>(2);
    double *psum = nullptr;
    long nb = 0;
    float *dsIn = cnn_alloc<float>((size_t)p.pairs * V * 2);
    float *dsTgt = cnn_alloc<float>((size_t)p.pairs * V * 2);
    Ellipsoid *dEll = cnn_alloc<Ellipsoid>(64);
    PhantomGen gen(p.seed, N, Nz, p.pixsize);
    std::vector<Ellipsoid> ea, em;
    float kappa = 0.0f;
    const dim3 fgrid(cdiv_u(n, 256)), cgrid(cdiv_u(V, 256)), blk(256);
    for (int p0 = 0; p0 < p.pairs; p0 += GB) {
        const int B = std::min(GB, p.pairs - p0);
        for (int b = 0; b < B; ++b) {
            gen.make(ea, em);
            float *xb = xs + b * n, *mb = mus + b * n;
            CNN_CHECK(hipMemcpyAsync(dEll, ea.data(), ea.size() * sizeof(Ellipsoid),
                                     hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_phantom, fgrid, blk, 0, s, dEll, (int)ea.size(), N, Nz, xb);
            // estimates of early OSEM iterations are smooth, of late ones noisy
            const float sig = gen.U(0.0f, 1.0f) < 0.7f ? gen.U(0.3f, 3.0f) : 0.0f;
            if (sig > 0) {
                for (int ax = 0; ax < 3; ++ax) {
                    hipLaunchKernelGGL(k_blur, fgrid, blk, 0, s, xb, tmp, N, Nz, ax, sig);
                    CNN_CHECK(hipMemcpyAsync(xb, tmp, n * 4, hipMemcpyDeviceToDevice, s));
                }
            }
            const float eps = gen.U(0.0f, 1.0f) < 0.5f ? gen.U(0.0f, 0.5f) : 0.0f;
            if (eps > 0) hipLaunchKernelGGL(k_noise, fgrid, blk, 0, s, xb, n, eps, (uint32_t)gen.rng());
            // the ellipsoid list must stay valid until the kernel has read it
            CNN_CHECK(hipStreamSynchronize(s));
            CNN_CHECK(hipMemcpyAsync(dEll, em.data(), em.size() * sizeof(Ellipsoid),
                                     hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_phantom, fgrid, blk, 0, s, dEll, (int)em.size(), N, Nz, mb);
            CNN_CHECK(hipStreamSynchronize(s));
        }
        esse.convolve(xs, B, terms);
        for (int b = 0; b < B; ++b) {
            const long q = p0 + b;
            mean_scale(xs + b * n, n, psum, nb, scal, s);
            hipLaunchKernelGGL(k_cnn_input, cgrid, blk, 0, s, xs + b * n, mus + b * n, N, Nz, scal,
                               fDens, dsIn + q * V * 2);
            if (kappa == 0.0f) {
                // output scale: both target channels of the first pair at unit rms
                hipLaunchKernelGGL(k_cnn_target, cgrid, blk, 0, s, terms + (long)b * m.nterms * n,
                                   terms + ((long)b * m.nterms + 1) * n, N, Nz, scal, fMu, 1.0f,
                                   dsTgt + q * V * 2);
                std::vector<float> h(V * 2);
                CNN_CHECK(hipMemcpyAsync(h.data(), dsTgt + q * V * 2, V * 2 * 4,
                                         hipMemcpyDeviceToHost, s));
                CNN_CHECK(hipStreamSynchronize(s));
                double s0 = 0, s1 = 0;
                for (long i = 0; i < V; ++i) {
                    s0 += (double)h[2 * i] * h[2 * i];
                    s1 += (double)h[2 * i + 1] * h[2 * i + 1];
                }
                kappa = s1 > 0 ? (float)std::sqrt(s0 / s1) : 1.0f;
            }
            hipLaunchKernelGGL(k_cnn_target, cgrid, blk, 0, s, terms + (long)b * m.nterms * n,
                               terms + ((long)b * m.nterms + 1) * n, N, Nz, scal, fMu, kappa,
                               dsTgt + q * V * 2);
        }
        CNN_CHECK(hipGetLastError());
    }
    CNN_CHECK(hipStreamSynchronize(s));
    cnn_free(xs);
    cnn_free(mus);
    cnn_free(tmp);
    cnn_free(terms);
    cnn_free(dEll);

    ScatterCnn *net = new ScatterCnn(kappa, p.seed);   // prepared below, after training
    const int ntrain = p.pairs - p.val;
    std::vector<int> order(ntrain);
    for (int i = 0; i < ntrain; ++i) order[i] = i;
    std::mt19937 rng(p.seed ^ 0x5bd1e995u);
    const float gscale = 2.0f / (float)(V * 2) / (float)p.batch;
    double epochLoss = 0;
    const long steps = (long)p.epochs * ((ntrain + p.batch - 1) / p.batch);
    long step = 0;
    for (int e = 0; e < p.epochs; ++e) {
        std::shuffle(order.begin(), order.end(), rng);
        epochLoss = 0;
        for (int i0 = 0; i0 < ntrain; i0 += p.batch) {
            const int B = std::min(p.batch, ntrain - i0);
            for (int b = 0; b < B; ++b) {
                const long q = order[i0 + b];
                epochLoss += net->lossGrad(dsIn + q * V * 2, dsTgt + q * V * 2, Y, X, Z, gscale, s);
            }
            // cosine decay of the learning rate
            const float lr = 0.5f * p.lr * (1.0f + std::cos(3.14159265f * (float)step / (float)steps));
            net->adamStep(lr, s);
            ++step;
        }
        epochLoss /= ntrain;
        if (std::getenv("IRLAMD_CNN_VERBOSE"))
            fprintf(stderr, "scatter CNN: epoch %d loss %.4g\n", e + 1, epochLoss);
    }
    // held-out relative L2 of the two outputs (inference path)
    double e0 = 0, t0s = 0, e1 = 0, t1s = 0;
    std::vector<float> ho(V * 2), ht(V * 2);
    float *pred = cnn_alloc<float>(V * 2);
    for (int q = ntrain; q < p.pairs; ++q) {
        net->forward(dsIn + (long)q * V * 2, Y, X, Z, pred, true, s);
        CNN_CHECK(hipMemcpyAsync(ho.data(), pred, V * 8, hipMemcpyDeviceToHost, s));
        CNN_CHECK(hipMemcpyAsync(ht.data(), dsTgt + (long)q * V * 2, V * 8, hipMemcpyDeviceToHost, s));
        CNN_CHECK(hipStreamSynchronize(s));
        for (long i = 0; i < V; ++i) {
            const double d0 = ho[2 * i] - ht[2 * i], d1 = ho[2 * i + 1] - ht[2 * i + 1];
            e0 += d0 * d0;
            e1 += d1 * d1;
            t0s += (double)ht[2 * i] * ht[2 * i];
            t1s += (double)ht[2 * i + 1] * ht[2 * i + 1];
        }
    }
    cnn_free(pred);
    net->prepareInference();
    cnn_free(dsIn);
    cnn_free(dsTgt);
    cnn_free(scal);
    cnn_free(psum);
    CNN_CHECK(hipStreamDestroy(s));
    st.loss = (float)epochLoss;
    st.valNrmse0 = (float)std::sqrt(e0 / std::max(t0s, 1e-300));
    st.valNrmse1 = (float)std::sqrt(e1 / std::max(t1s, 1e-300));
    st.kappa = kappa;
    st.seconds = std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
    return net;
}

}  // namespace irlamd

// ---------------------------------------------------------------------------
// C API (include/irl_amd.h)
// ---------------------------------------------------------------------------
#include "../../include/irl_amd.h"
#include "cnn.h"
#include "irlprj.h"

namespace {

int cnn_report(const irlamd::IrlError &e)
{
    irlamd::set_error_string(e.msg);
    irlamd::set_last_error(e.code);
    fprintf(stderr, "scatter CNN: %s\n", e.msg.c_str());
    return e.code;
}

struct CnnFileHeader {
    char magic[8];       // "IRLCNN1"
    int32_t nparams;
    float kappa;
    int32_t trainN, trainNz;
    float trainPix;
    int32_t reserved;
};

// Device buffer of a host array for the test entry points.
struct HostDev {
    float *p = nullptr;
    explicit HostDev(long n) { CNN_CHECK(hipMalloc((void **)&p, n * sizeof(float))); }
    ~HostDev() { (void)hipFree(p); }
};

}  // namespace

extern "C" {

irlamd_cnn *irlamd_cnn_create(float fKappa, unsigned uSeed)
{
    irlamd::set_last_error(0);
    try {
        auto *h = new irlamd_cnn;
        h->net.reset(new irlamd::ScatterCnn(fKappa, uSeed));
        // usable at once by reconstructions (BF16 inference form of the weights)
        h->net->prepareInference();
        return h;
    } catch (const irlamd::IrlError &e) {
        cnn_report(e);
        return nullptr;
    }
}

irlamd_cnn *irlamd_cnn_train(const char *krnl, int N, int Nz, float pix, int pairs, int val,
                             int epochs, int batch, float lr, unsigned seed, float *stats)
{
    irlamd::set_last_error(0);
    try {
        if (!krnl) irlamd::fail(irlamd::ETYPE_ILLEGAL_VALUE, "scatter CNN: no kernel file");
        const irlamd::KrnlData k = irlamd::read_krnl_file(krnl);
        const irlamd::EsseModel m = irlamd::make_esse_model(k, "2 2 2 3", pix, N, Nz, 1);
        irlamd::CnnTrainParms p;
        p.N = N;
        p.Nz = Nz;
        p.pixsize = pix;
        p.pairs = pairs;
        p.val = val;
        p.epochs = epochs;
        p.batch = batch;
        p.lr = lr;
        p.seed = seed;
        irlamd::CnnTrainStats st;
        auto *h = new irlamd_cnn;
        h->net.reset(irlamd::ScatterCnn::train(m, p, st));
        if (stats) {
            stats[0] = st.loss;
            stats[1] = st.valNrmse0;
            stats[2] = st.valNrmse1;
            stats[3] = st.seconds;
            stats[4] = st.kappa;
        }
        return h;
    } catch (const irlamd::IrlError &e) {
        cnn_report(e);
        return nullptr;
    }
}

int irlamd_cnn_save(irlamd_cnn *h, const char *fname)
{
    irlamd::set_last_error(0);
    try {
        if (!h || !fname) irlamd::fail(irlamd::ETYPE_ILLEGAL_VALUE, "scatter CNN: bad argument");
        std::vector<float> w;
        h->net->getParams(w);
        CnnFileHeader hd{};
        std::memcpy(hd.magic, "IRLCNN1", 8);
        hd.nparams = (int32_t)w.size();
        hd.kappa = h->net->kappa();
        FILE *f = fopen(fname, "wb");
        if (!f) irlamd::fail(irlamd::ETYPE_IO, "scatter CNN: cannot write %s", fname);
        const bool ok = fwrite(&hd, sizeof hd, 1, f) == 1 &&
                        fwrite(w.data(), sizeof(float), w.size(), f) == w.size();
        fclose(f);
        if (!ok) irlamd::fail(irlamd::ETYPE_IO, "scatter CNN: short write to %s", fname);
        return 0;
    } catch (const irlamd::IrlError &e) {
        return cnn_report(e);
    }
}

irlamd_cnn *irlamd_cnn_load(const char *fname)
{
    irlamd::set_last_error(0);
    try {
        FILE *f = fname ? fopen(fname, "rb") : nullptr;
        if (!f) irlamd::fail(irlamd::ETYPE_IO, "scatter CNN: cannot open %s", fname ? fname : "");
        CnnFileHeader hd{};
        std::vector<float> w;
        bool ok = fread(&hd, sizeof hd, 1, f) == 1 && std::memcmp(hd.magic, "IRLCNN1", 8) == 0 &&
                  hd.nparams == irlamd::cnn_num_params();
        if (ok) {
            w.resize(hd.nparams);
            ok = fread(w.data(), sizeof(float), w.size(), f) == w.size();
        }
        fclose(f);
        if (!ok) irlamd::fail(irlamd::ETYPE_IO, "scatter CNN: %s is not a network file", fname);
        auto *h = new irlamd_cnn;
        h->net.reset(new irlamd::ScatterCnn(hd.kappa, w));
        return h;
    } catch (const irlamd::IrlError &e) {
        cnn_report(e);
        return nullptr;
    }
}

void irlamd_cnn_free(irlamd_cnn *h) { delete h; }

int irlamd_cnn_num_params(void) { return (int)irlamd::cnn_num_params(); }

int irlamd_cnn_get_params(irlamd_cnn *h, float *out, float *kappa)
{
    irlamd::set_last_error(0);
    try {
        std::vector<float> w;
        h->net->getParams(w);
        std::memcpy(out, w.data(), w.size() * sizeof(float));
        if (kappa) *kappa = h->net->kappa();
        return 0;
    } catch (const irlamd::IrlError &e) {
        return cnn_report(e);
    }
}

int irlamd_cnn_set_params(irlamd_cnn *h, const float *in)
{
    irlamd::set_last_error(0);
    try {
        h->net->setParams(std::vector<float>(in, in + irlamd::cnn_num_params()));
        h->net->prepareInference();
        return 0;
    } catch (const irlamd::IrlError &e) {
        return cnn_report(e);
    }
}

int irlamd_cnn_set_row_length(irlamd_cnn *h, int rowz)
{
    irlamd::set_last_error(0);
    try {
        h->net->setRowLength(rowz);
        return 0;
    } catch (const irlamd::IrlError &e) {
        return cnn_report(e);
    }
}

int irlamd_cnn_forward(irlamd_cnn *h, const float *in, int Y, int X, int Z, int bf16, float *out)
{
    irlamd::set_last_error(0);
    try {
        if (Y < 1 || X < 1 || Z < 1) irlamd::fail(irlamd::ETYPE_ILLEGAL_VALUE, "bad CNN grid");
        const long n = (long)Y * X * Z * 2;
        HostDev di(n), dout(n);
        CNN_CHECK(hipMemcpy(di.p, in, n * 4, hipMemcpyHostToDevice));
        h->net->forward(di.p, Y, X, Z, dout.p, bf16 != 0, nullptr);
        CNN_CHECK(hipMemcpy(out, dout.p, n * 4, hipMemcpyDeviceToHost));
        return 0;
    } catch (const irlamd::IrlError &e) {
        return cnn_report(e);
    }
}

int irlamd_cnn_loss_grad(irlamd_cnn *h, const float *in, const float *tgt, int Y, int X, int Z,
                         double *loss, float *grads)
{
    irlamd::set_last_error(0);
    try {
        if (Y < 1 || X < 1 || Z < 1) irlamd::fail(irlamd::ETYPE_ILLEGAL_VALUE, "bad CNN grid");
        const long n = (long)Y * X * Z * 2;
        HostDev di(n), dt(n);
        CNN_CHECK(hipMemcpy(di.p, in, n * 4, hipMemcpyHostToDevice));
        CNN_CHECK(hipMemcpy(dt.p, tgt, n * 4, hipMemcpyHostToDevice));
        h->net->zeroGrads(nullptr);
        const double l = h->net->lossGrad(di.p, dt.p, Y, X, Z, 2.0f / (float)n, nullptr);
        if (loss) *loss = l;
        std::vector<float> g;
        h->net->getGrads(g);
        std::memcpy(grads, g.data(), g.size() * sizeof(float));
        h->net->zeroGrads(nullptr);
        CNN_CHECK(hipDeviceSynchronize());
        return 0;
    } catch (const irlamd::IrlError &e) {
        return cnn_report(e);
    }
}

int irlamd_cnn_sources(irlamd_cnn *h, const float *x, const float *mu, int N, int Nz,
                       float fDens, float fCntr, float primary, float *g0, float *g1)
{
    irlamd::set_last_error(0);
    try {
        if (N < 2 || Nz < 2) irlamd::fail(irlamd::ETYPE_ILLEGAL_VALUE, "bad volume size");
        const long n = (long)N * N * Nz;
        HostDev dx(n), dm(n), d0(n), d1(n);
        CNN_CHECK(hipMemcpy(dx.p, x, n * 4, hipMemcpyHostToDevice));
        CNN_CHECK(hipMemcpy(dm.p, mu, n * 4, hipMemcpyHostToDevice));
        irlamd::CnnPhys ph;
        ph.fDens = fDens;
        ph.cntr = fCntr;
        h->net->prepareInference();
        irlamd::CnnWork w(N, Nz);
        h->net->sources(dx.p, dm.p, N, Nz, ph, primary, d0.p, d1.p, w, nullptr);
        CNN_CHECK(hipMemcpy(g0, d0.p, n * 4, hipMemcpyDeviceToHost));
        CNN_CHECK(hipMemcpy(g1, d1.p, n * 4, hipMemcpyDeviceToHost));
        return 0;
    } catch (const irlamd::IrlError &e) {
        return cnn_report(e);
    }
}

}  // extern "C"
