// fft_dev.h -- line FFTs held in registers for gfx950 kernels (HIP
// translation units only): complex helpers, direct 2- to 16-point, 10- and
// 12-point DFTs, and mixed-radix Stockham line transforms of the sizes in
// FftPlan, exchanging through LDS between passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace irlamd {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// cos and sin of 2*pi*k/16, k = 0..7.
__device__ constexpr float kCos16[8] = {1.0f, 0.923879532511286756f, 0.707106781186547524f,
                                        0.382683432365089772f, 0.0f, -0.382683432365089772f,
                                        -0.707106781186547524f, -0.923879532511286756f};
__device__ constexpr float kSin16[8] = {0.0f, 0.382683432365089772f, 0.707106781186547524f,
                                        0.923879532511286756f, 1.0f, 0.923879532511286756f,
                                        0.707106781186547524f, 0.382683432365089772f};

// z * exp(-+2 pi i K/16) for a compile-time K in [0, 8).
template <int K, bool INV>
__device__ __forceinline__ float2 rot16(float2 z)
{
    if (K == 0) return z;
    if (K == 4) return INV ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
    const float c = kCos16[K], sn = INV ? kSin16[K] : -kSin16[K];
    return make_float2(fmaf(z.x, c, -z.y * sn), fmaf(z.x, sn, z.y * c));
}

// One radix-2 Stockham stage of an r-point DFT held in registers:
// y[q + 2sp] = a + b, y[q + s(2p+1)] = (a - b) W_r^{ps}, a = x[q + sp],
// b = x[q + s(p + m)], m = r/(2s).
template <int r, int s, bool INV>
__device__ __forceinline__ void dft_stage(float2 *x)
{
    constexpr int m = r / (2 * s);
    float2 y[r];
#pragma unroll
    for (int p = 0; p < m; ++p)
#pragma unroll
        for (int q = 0; q < s; ++q) {
            const float2 a = x[q + s * p], b = x[q + s * (p + m)];
            y[q + 2 * s * p] = cadd(a, b);
            y[q + s * (2 * p + 1)] = csub(a, b);
        }
#pragma unroll
    for (int p = 0; p < m; ++p)
#pragma unroll
        for (int q = 0; q < s; ++q) {
            // twiddle index p*s in units of 1/r turns = p*s*(16/r) sixteenths
            constexpr int dummy = 0;
            (void)dummy;
            const int k = p * s * (16 / r);
            float2 &z = y[q + s * (2 * p + 1)];
            switch (k) {
            case 0: break;
            case 1: z = rot16<1, INV>(z); break;
            case 2: z = rot16<2, INV>(z); break;
            case 3: z = rot16<3, INV>(z); break;
            case 4: z = rot16<4, INV>(z); break;
            case 5: z = rot16<5, INV>(z); break;
            case 6: z = rot16<6, INV>(z); break;
            case 7: z = rot16<7, INV>(z); break;
            }
        }
#pragma unroll
    for (int i = 0; i < r; ++i) x[i] = y[i];
}

// In-place r-point DFT (r = 2, 4, 8, 16), natural order in and out:
// x_v <- sum_u x_u exp(-+2 pi i uv/r).
template <int r, bool INV>
__device__ __forceinline__ void dft(float2 *x)
{
    if constexpr (r >= 2) dft_stage<r, 1, INV>(x);
    if constexpr (r >= 4) dft_stage<r, 2, INV>(x);
    if constexpr (r >= 8) dft_stage<r, 4, INV>(x);
    if constexpr (r >= 16) dft_stage<r, 8, INV>(x);
}

// cos and sin of 2*pi*n/60 (the radix-10 and radix-12 twiddles are n = 6k, 5k).
__device__ constexpr float kCos60[60] = {1.000000000f, 0.994521895f, 0.978147601f, 0.951056516f, 0.913545458f, 0.866025404f, 0.809016994f, 0.743144825f, 0.669130606f, 0.587785252f, 0.500000000f, 0.406736643f, 0.309016994f, 0.207911691f, 0.104528463f, 0.000000000f, -0.104528463f, -0.207911691f, -0.309016994f, -0.406736643f, -0.500000000f, -0.587785252f, -0.669130606f, -0.743144825f, -0.809016994f, -0.866025404f, -0.913545458f, -0.951056516f, -0.978147601f, -0.994521895f, -1.000000000f, -0.994521895f, -0.978147601f, -0.951056516f, -0.913545458f, -0.866025404f, -0.809016994f, -0.743144825f, -0.669130606f, -0.587785252f, -0.500000000f, -0.406736643f, -0.309016994f, -0.207911691f, -0.104528463f, -0.000000000f, 0.104528463f, 0.207911691f, 0.309016994f, 0.406736643f, 0.500000000f, 0.587785252f, 0.669130606f, 0.743144825f, 0.809016994f, 0.866025404f, 0.913545458f, 0.951056516f, 0.978147601f, 0.994521895f};
__device__ constexpr float kSin60[60] = {0.000000000f, 0.104528463f, 0.207911691f, 0.309016994f, 0.406736643f, 0.500000000f, 0.587785252f, 0.669130606f, 0.743144825f, 0.809016994f, 0.866025404f, 0.913545458f, 0.951056516f, 0.978147601f, 0.994521895f, 1.000000000f, 0.994521895f, 0.978147601f, 0.951056516f, 0.913545458f, 0.866025404f, 0.809016994f, 0.743144825f, 0.669130606f, 0.587785252f, 0.500000000f, 0.406736643f, 0.309016994f, 0.207911691f, 0.104528463f, 0.000000000f, -0.104528463f, -0.207911691f, -0.309016994f, -0.406736643f, -0.500000000f, -0.587785252f, -0.669130606f, -0.743144825f, -0.809016994f, -0.866025404f, -0.913545458f, -0.951056516f, -0.978147601f, -0.994521895f, -1.000000000f, -0.994521895f, -0.978147601f, -0.951056516f, -0.913545458f, -0.866025404f, -0.809016994f, -0.743144825f, -0.669130606f, -0.587785252f, -0.500000000f, -0.406736643f, -0.309016994f, -0.207911691f, -0.104528463f};

// z * exp(-+2 pi i n/60) for a compile-time n.
template <int N60, bool INV>
__device__ __forceinline__ float2 rot60(float2 z)
{
    constexpr int n = N60 % 60;
    if (n == 0) return z;
    if (n == 30) return make_float2(-z.x, -z.y);
    if (n == 15) return INV ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
    if (n == 45) return INV ? make_float2(z.y, -z.x) : make_float2(-z.y, z.x);
    const float c = kCos60[n], sn = INV ? kSin60[n] : -kSin60[n];
    return make_float2(fmaf(z.x, c, -z.y * sn), fmaf(z.x, sn, z.y * c));
}

// Direct r-point DFT (r dividing 60), natural order: y_v = sum_u x_u W_r^{uv}.
template <int r, bool INV, int V, int U>
__device__ __forceinline__ void dft_direct_term(const float2 *x, float2 &acc)
{
    if constexpr (U < r) {
        const float2 t = rot60<(U * V % r) * (60 / r), INV>(x[U]);
        acc = cadd(acc, t);
        dft_direct_term<r, INV, V, U + 1>(x, acc);
    }
}

template <int r, bool INV, int V>
__device__ __forceinline__ void dft_direct_out(const float2 *x, float2 *y)
{
    if constexpr (V < r) {
        float2 acc = x[0];
        dft_direct_term<r, INV, V, 1>(x, acc);
        y[V] = acc;
        dft_direct_out<r, INV, V + 1>(x, y);
    }
}

// x - i*y and x + i*y
__device__ __forceinline__ float2 c_mi(float2 a, float2 b) { return make_float2(a.x + b.y, a.y - b.x); }
__device__ __forceinline__ float2 c_pi(float2 a, float2 b) { return make_float2(a.x - b.y, a.y + b.x); }

// 3- and 5-point DFTs in place, natural order (y_k = sum_n x_n W^{nk},
// W = exp(-+2 pi i / r)), from the symmetric sums and differences of the
// input pairs (n, r - n).
template <bool INV>
__device__ __forceinline__ void dft3(float2 &x0, float2 &x1, float2 &x2)
{
    constexpr float h = 0.866025404f;   // sin(2 pi/3)
    const float2 t = cadd(x1, x2), d = csub(x1, x2);
    const float2 m = make_float2(x0.x - 0.5f * t.x, x0.y - 0.5f * t.y);
    const float2 b = make_float2(h * d.x, h * d.y);
    x0 = cadd(x0, t);
    x1 = INV ? c_pi(m, b) : c_mi(m, b);
    x2 = INV ? c_mi(m, b) : c_pi(m, b);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2 &x0, float2 &x1, float2 &x2, float2 &x3, float2 &x4)
{
    constexpr float c1 = 0.309016994f, c2 = -0.809016994f;   // cos(2 pi/5), cos(4 pi/5)
    constexpr float s1 = 0.951056516f, s2 = 0.587785252f;    // sin(2 pi/5), sin(4 pi/5)
    const float2 t1 = cadd(x1, x4), t2 = cadd(x2, x3), t3 = csub(x1, x4), t4 = csub(x2, x3);
    const float2 a1 = make_float2(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
    const float2 a2 = make_float2(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
    const float2 b1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const float2 b2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    x0 = cadd(x0, cadd(t1, t2));
    x1 = INV ? c_pi(a1, b1) : c_mi(a1, b1);
    x4 = INV ? c_mi(a1, b1) : c_pi(a1, b1);
    x2 = INV ? c_pi(a2, b2) : c_mi(a2, b2);
    x3 = INV ? c_mi(a2, b2) : c_pi(a2, b2);
}

// 10 = 2 x 5: DFT5 of the even and odd samples, then one radix-2 stage with
// the twiddles W10^k (= W60^{6k}).
template <bool INV>
__device__ __forceinline__ void dft10(float2 *x)
{
    float2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6], e4 = x[8];
    float2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7], o4 = x[9];
    dft5<INV>(e0, e1, e2, e3, e4);
    dft5<INV>(o0, o1, o2, o3, o4);
    o1 = rot60<6, INV>(o1);
    o2 = rot60<12, INV>(o2);
    o3 = rot60<18, INV>(o3);
    o4 = rot60<24, INV>(o4);
    x[0] = cadd(e0, o0);
    x[5] = csub(e0, o0);
    x[1] = cadd(e1, o1);
    x[6] = csub(e1, o1);
    x[2] = cadd(e2, o2);
    x[7] = csub(e2, o2);
    x[3] = cadd(e3, o3);
    x[8] = csub(e3, o3);
    x[4] = cadd(e4, o4);
    x[9] = csub(e4, o4);
}

// 12 = 3 x 4: X[k1 + 3 k2] = sum_j W4^{j k2} W12^{j k1} DFT3(x[j], x[j+4], x[j+8])[k1].
template <bool INV>
__device__ __forceinline__ void dft12(float2 *x)
{
    float2 d[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        d[j][0] = x[j];
        d[j][1] = x[j + 4];
        d[j][2] = x[j + 8];
        dft3<INV>(d[j][0], d[j][1], d[j][2]);
    }
    d[1][1] = rot60<5, INV>(d[1][1]);
    d[1][2] = rot60<10, INV>(d[1][2]);
    d[2][1] = rot60<10, INV>(d[2][1]);
    d[2][2] = rot60<20, INV>(d[2][2]);
    d[3][1] = rot60<15, INV>(d[3][1]);
    d[3][2] = rot60<30, INV>(d[3][2]);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        float2 t[4] = {d[0][k1], d[1][k1], d[2][k1], d[3][k1]};
        dft<4, INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) x[k1 + 3 * k2] = t[k2];
    }
}

template <int r, bool INV>
__device__ __forceinline__ void dft_any(float2 *x)
{
    if constexpr ((r & (r - 1)) == 0) {
        dft<r, INV>(x);
    } else if constexpr (r == 10) {
        dft10<INV>(x);
    } else if constexpr (r == 12) {
        dft12<INV>(x);
    } else {
        float2 y[r];
        dft_direct_out<r, INV, 0>(x, y);
#pragma unroll
        for (int i = 0; i < r; ++i) x[i] = y[i];
    }
}

// Line FFT plans: P = T threads x R values per thread, Stockham passes of
// the listed radices, each radix r dividing R and P/r divisible by T, so
// every thread owns R/r whole units in every pass.
template <int P>
struct FftPlan;
template <> struct FftPlan<16> { static constexpr int T = 1, R = 16, r0 = 16, r1 = 0, r2 = 0, r3 = 0, r4 = 0; };
template <> struct FftPlan<32> { static constexpr int T = 2, R = 16, r0 = 16, r1 = 2, r2 = 0, r3 = 0, r4 = 0; };
template <> struct FftPlan<64> { static constexpr int T = 4, R = 16, r0 = 16, r1 = 4, r2 = 0, r3 = 0, r4 = 0; };
template <> struct FftPlan<128> { static constexpr int T = 8, R = 16, r0 = 16, r1 = 8, r2 = 0, r3 = 0, r4 = 0; };
template <> struct FftPlan<160> { static constexpr int T = 8, R = 20, r0 = 4, r1 = 4, r2 = 10, r3 = 0, r4 = 0; };
template <> struct FftPlan<192> { static constexpr int T = 16, R = 12, r0 = 4, r1 = 4, r2 = 12, r3 = 0, r4 = 0; };
template <> struct FftPlan<256> { static constexpr int T = 16, R = 16, r0 = 16, r1 = 16, r2 = 0, r3 = 0, r4 = 0; };
template <> struct FftPlan<512> { static constexpr int T = 32, R = 16, r0 = 16, r1 = 16, r2 = 2, r3 = 0, r4 = 0; };

// Line FFT of length P: thread tt holds x[tt + T*w] in a[w] (w < R) on
// entry and X[tt + T*w] on exit.  Pass with radix r and stride S: unit
// (p, q), p < m = P/(S r), q < S, reads x[q + S(p + m u)], applies the
// r-point DFT and the twiddle W_P^{v p S}, and writes y[q + S(r p + v)].
// Between passes the line goes through LDS with a skew of one element per
// 16 against bank conflicts.
template <int P>
struct FftLine {
    static constexpr int R = FftPlan<P>::R;
    static constexpr int T = FftPlan<P>::T;
    static constexpr int LS = P + P / 16 + 1;  // LDS line stride (float2)
    static_assert(T * R == P, "plan");
};

__device__ __forceinline__ int fft_skew(int i) { return (int)((unsigned)i + ((unsigned)i >> 4)); }

template <int P, int IDX>
__host__ __device__ constexpr int fft_radix()
{
    return IDX == 0   ? FftPlan<P>::r0
           : IDX == 1 ? FftPlan<P>::r1
           : IDX == 2 ? FftPlan<P>::r2
           : IDX == 3 ? FftPlan<P>::r3
           : IDX == 4 ? FftPlan<P>::r4
                      : 0;
}

template <int P, int IDX, int S, bool INV>
struct FftPass {
    static constexpr int r = fft_radix<P, IDX>();
    static constexpr int m = P / (S * r);
    static constexpr int R = FftPlan<P>::R;
    static constexpr int U = R / r;           // units per thread
    static constexpr int T = FftPlan<P>::T;
    static_assert(U * r == R && (P / r) % T == 0, "plan");
    // unsigned index arithmetic: S is a power of two in every plan, so the
    // divisions are shifts
    __device__ static int in_pos(int tt, int k, int u)
    {
        const unsigned j = (unsigned)tt + T * k, p = j / S, q = j % S;
        return (int)(q + S * (p + m * u));
    }
    __device__ static int out_pos(int tt, int k, int v)
    {
        const unsigned j = (unsigned)tt + T * k, p = j / S, q = j % S;
        return (int)(q + S * (r * p + v));
    }
    __device__ static void compute(float2 *a, int tt, const float2 *__restrict__ tw)
    {
#pragma unroll
        for (int k = 0; k < U; ++k) {
            dft_any<r, INV>(a + k * r);
            if (m > 1) {
                const unsigned p = ((unsigned)tt + T * k) / S;
#pragma unroll
                for (int v = 1; v < r; ++v) {
                    // v p S < r m S = P: no reduction needed
                    float2 w = tw[v * p * S];
                    if (INV) w.y = -w.y;
                    a[k * r + v] = cmul(a[k * r + v], w);
                }
            }
        }
    }
};

// Synchronisation between the passes' LDS writes and reads.  Block-wide by
// default.  With WAVE, a line's threads are consecutive lanes of one wave
// (T divides 64, lines start at multiples of T): a wave's LDS instructions
// are performed in issue order, so only the compiler must keep the accesses
// in program order.
template <bool WAVE>
__device__ __forceinline__ void fft_sync()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

template <int P, int IDX, int S, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft_passes(float2 *a, float2 *ldsl, int tt,
                                           const float2 *__restrict__ tw)
{
    using PS = FftPass<P, IDX, S, INV>;
    constexpr int R = PS::R;
    if constexpr (IDX == 0) {
        if constexpr (PS::U > 1) {
            // unit order: a_u[k r + u] = x[tt + T(k + U u)] = a[k + U u]
            float2 b[R];
#pragma unroll
            for (int k = 0; k < PS::U; ++k)
#pragma unroll
                for (int u = 0; u < PS::r; ++u) b[k * PS::r + u] = a[k + PS::U * u];
#pragma unroll
            for (int w = 0; w < R; ++w) a[w] = b[w];
        }
    } else {
#pragma unroll
        for (int k = 0; k < PS::U; ++k)
#pragma unroll
            for (int u = 0; u < PS::r; ++u) a[k * PS::r + u] = ldsl[fft_skew(PS::in_pos(tt, k, u))];
    }
    PS::compute(a, tt, tw);
    if constexpr (PS::m == 1) {
        // final pass: a[k r + v] holds X[tt + T k + S v]; reorder to a[w] = X[tt + T w]
        float2 b[R];
#pragma unroll
        for (int k = 0; k < PS::U; ++k)
#pragma unroll
            for (int v = 0; v < PS::r; ++v) b[k + PS::U * v] = a[k * PS::r + v];
#pragma unroll
        for (int w = 0; w < R; ++w) a[w] = b[w];
    } else {
        fft_sync<WAVE>();   // earlier readers of the LDS line are done
#pragma unroll
        for (int k = 0; k < PS::U; ++k)
#pragma unroll
            for (int v = 0; v < PS::r; ++v) ldsl[fft_skew(PS::out_pos(tt, k, v))] = a[k * PS::r + v];
        fft_sync<WAVE>();
        fft_passes<P, IDX + 1, S * PS::r, INV, WAVE>(a, ldsl, tt, tw);
    }
}

template <int P, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft_line(float2 *a, float2 *ldsl, int tt,
                                         const float2 *__restrict__ tw)
{
    static_assert(!WAVE || 64 % FftPlan<P>::T == 0, "a line within one wave");
    fft_passes<P, 0, 1, INV, WAVE>(a, ldsl, tt, tw);
}

// W_P^k = exp(-2 pi i k / P), k < P, rounded from double: the twiddles the
// line transforms read.
inline std::vector<float2> twiddles(int P)
{
    std::vector<float2> t(P);
    for (int k = 0; k < P; ++k) {
        const double a = -2.0 * M_PI * k / P;
        t[k] = make_float2((float)cos(a), (float)sin(a));
    }
    return t;
}

}  // namespace irlamd
