// esse_dev.h -- device side of ESSE (HIP translation units only).
#pragma once

#include <hip/hip_runtime.h>

#include "esse.h"

namespace irlamd {

// Batched ESSE: for B rotated activity volumes [B][ny][nz][nx] computes the
// FFT convolutions with the kernel grids (K, K(mu-muavg), K(mu-muavg)^2),
// collapsed and expanded as the model asks, into terms [B][T][ny][nz][nx].
class EsseDev {
public:
    EsseDev(const EsseModel &m, int maxBatch, hipStream_t stream);
    ~EsseDev();
    EsseDev(const EsseDev &) = delete;
    EsseDev &operator=(const EsseDev &) = delete;

    // act: [B][ny][nz][nx] device; terms: [B][T][ny][nz][nx] device.
    void convolve(const float *act, int B, float *terms);

    // Effective scatter source plus primaries (krnlprj3d.c:826-948 and
    // prjbck.c:164-176): q = act*primary + eff(act, map), optionally times
    // afac (the attenuation factors of the view, [V][ny][nz][nx]).
    // Writes q * afac into out [B][ny][nz][nx].
    void source(const float *act, const float *map, const float *terms, int B, float primary,
                const float *afac, const int *bview, float *out);

    const EsseModel &model() const { return m_; }
    int terms() const { return m_.nterms; }

private:
    EsseModel m_;
    int maxB_;
    hipStream_t stream_;
    float2 *khat_ = nullptr;   // [T][Py][Pz][Px]
    float2 *buf_ = nullptr;    // [B][nyc][Pz][Px]
    float2 *outk_ = nullptr;   // [B][T][nyc][Pz][Px]
    float *coll_ = nullptr;    // collapsed activity [B][nyc][nzc][nxc]
    float *res_ = nullptr;     // collapsed terms [B][T][nyc][nzc][nxc] (c > 1)
    float2 *twx_ = nullptr, *twy_ = nullptr, *twz_ = nullptr;
};

}  // namespace irlamd
