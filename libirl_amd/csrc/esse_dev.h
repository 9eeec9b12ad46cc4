// esse_dev.h -- device side of ESSE (HIP translation units only).
#pragma once

#include <hip/hip_runtime.h>

#include "esse.h"

namespace irlamd {

// Batched ESSE: for B rotated activity volumes computes the FFT
// convolutions with the kernel grids (K, K(mu-muavg), K(mu-muavg)^2),
// collapsed and expanded as the model asks.  Volumes are [B][ny][nz][nx]
// (the reference's image order) or, with zfast, [B][ny][nx][nz] (slices
// fastest, the reconstruction engine's order); terms are [B][T] volumes of
// the same layout.
class EsseDev {
public:
    EsseDev(const EsseModel &m, int maxBatch, hipStream_t stream, bool zfast);
    ~EsseDev();
    EsseDev(const EsseDev &) = delete;
    EsseDev &operator=(const EsseDev &) = delete;

    // act: [B] volumes on the device; terms: [B][T] volumes.
    void convolve(const float *act, int B, float *terms);

    // Effective scatter source plus primaries (krnlprj3d.c:826-948 and
    // prjbck.c:164-176): q = act*primary + eff(act, map), optionally times
    // afac (the attenuation factors of the view, [V] volumes).
    // Writes q * afac into out ([B] volumes); all in the same layout.
    // mapByView: map holds the rotated maps of every view, indexed by bview.
    void source(const float *act, const float *map, const float *terms, int B, float primary,
                const float *afac, const int *bview, float *out, bool mapByView = false);

    // convolve + source in one pass over the FFT output, where the model and
    // the buffers allow it: T = 2 terms, no collapse, slice-fastest layout,
    // maps and factors of every view (indexed by bview) with afac = exp(-S),
    // S the running half-pixel sum of the map along depth (what the engine's
    // factors are for its rotated mu).  The two terms' inverse transforms
    // along the slices run as one complex transform and the source is formed
    // from them in registers, with the depth taken from the factors
    // (-ln afac); the terms are not stored.  The factors must have been
    // built with an attenuation factor of 1 (the engine's precomputed ones
    // are); a factor of 0 (a depth past the float range) gives a zero
    // source.  Returns false, having done nothing, when the conditions do
    // not hold.
    bool convolveSource(const float *act, const float *map, int B, float primary,
                        const float *afac, const int *bview, float *out);
    bool canConvolveSource(const float *map, const float *afac, const int *bview) const
    {
        return fused_ && m_.enabled && m_.nterms == 2 && m_.c == 1 && zfast_ && afac && bview &&
               map;
    }
    // off: the line passes everywhere and no fused source
    void setFused(bool on) { fused_ = on; }
    // the stream later calls launch on (the work buffers are shared: calls
    // on different streams must not overlap)
    void setStream(hipStream_t s) { stream_ = s; }

    const EsseModel &model() const { return m_; }
    int terms() const { return m_.nterms; }

private:
    void forward2(const float *in, long planes, int nm, int nf);
    void fft_f_fwd(const float *in, float2 *out, long rows, int nm, int nf);
    void fft_f_inv(const float2 *in, float *out, long rows, int nm, int nf, float scale);
    void fft_strided(int P, bool inv, const float2 *src, float2 *dst, long nbases, long bstride,
                     long nstride, int W, int nIn, int nOut, const float2 *tw);

    EsseModel m_;
    int maxB_;
    hipStream_t stream_;
    bool zfast_;
    int Pf_ = 0, Pm_ = 0;      // FFT lengths along the fastest and the other in-plane axis
    int Wc_ = 0;               // stored half-spectrum width along the fastest axis
    float2 *khat_ = nullptr;   // [T][Py][Pm][Wc]
    float2 *buf_ = nullptr;    // [B][nyc][Pm][Wc]
    float2 *outk_ = nullptr;   // [B][T][nyc][Pm][Wc]
    float *coll_ = nullptr;    // collapsed activity [B][nyc][nzc][nxc]
    float *res_ = nullptr;     // collapsed terms [B][T][nyc][nzc][nxc] (c > 1)
    float2 *twf_ = nullptr, *twy_ = nullptr, *twm_ = nullptr;
    bool fused_ = true;
};

}  // namespace irlamd
