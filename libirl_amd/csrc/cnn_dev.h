// cnn_dev.h -- scatter CNN (device side, HIP translation units only).
//
// A 4-layer dilated 3-D convolution network on the half-resolution grid that
// predicts the two FFT convolution terms of the ESSE effective source,
// E0 = K * act and E1 = K(mu - muavg) * act (krnlprj3d.c:826-948), in the
// unrotated frame once per sub-iteration.  The per-view source is then
//   att_v = (rot g0 + d_v rot g1) afac_v,
//   g0 = primary x + (E0 + cntr x) dens,  g1 = (fMu E0 - E1) dens,
// with dens = mu * fDens and d_v = -ln(afac_v) / mu0 the view's attenuation
// depth: the reference's (E0 - d E1) exp(d fMu) to first order in d fMu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "esse.h"

namespace irlamd {

constexpr int CNN_LAYERS = 4;
constexpr int cnn_cin(int l) { return l == 0 ? 2 : 32; }
constexpr int cnn_cout(int l) { return l == CNN_LAYERS - 1 ? 2 : 32; }
constexpr int cnn_dil(int l) { return 1 << l; }
// parameters of layer l: W [27][Cout][Cin] then b [Cout] (packed order)
constexpr long cnn_layer_params(int l) { return 27L * cnn_cout(l) * cnn_cin(l) + cnn_cout(l); }
long cnn_num_params();

// Constants of the ESSE model the network serves (from the reconstruction's
// scatter model): fMu = mu0 - muavg, fDens = 1/(mu0 pixsize), cntr = fCntrPixFac.
struct CnnPhys {
    float mu0 = 0.15f, fMu = 0.0f, fDens = 1.0f, cntr = 0.0f;
};

struct CnnTrainParms {
    int N = 128, Nz = 128;       // training volume size (even)
    float pixsize = 0.33f;
    int pairs = 2000, val = 100; // generated pairs, of which the last 'val' are held out
    int epochs = 10, batch = 4;
    float lr = 2e-3f;
    unsigned seed = 1;
};

struct CnnTrainStats {
    float loss = 0;              // mean training loss of the last epoch
    float valNrmse0 = 0, valNrmse1 = 0;  // held-out rel. L2 of the two outputs
    float seconds = 0;           // generation + training wall time
    float kappa = 0;
};

// Inference work buffers of one user of a network (each reconstruction has
// its own, so reconstructions on different streams share only the weights).
struct CnnWork {
    long V = 0;                          // coarse voxels
    float *in = nullptr, *out = nullptr; // [V][2]
    uint16_t *actb[2] = {};              // bf16 activations [V][32]
    double *psum = nullptr;              // block sums of x (one per 256 coarse voxels)
    long nb = 0;
    float *scal = nullptr;               // [0] 1/mean(x), [1] mean(x)
    int rowz = 64;                       // slices per wave of the MFMA layers (64, 32, 16)
    CnnWork(int N, int Nz);
    ~CnnWork();
    CnnWork(const CnnWork &) = delete;
    CnnWork &operator=(const CnnWork &) = delete;
};

class ScatterCnn {
public:
    // Untrained network (He-normal weights from 'seed'), output scale kappa.
    ScatterCnn(float kappa, unsigned seed);
    // From packed parameters (oracle order).
    ScatterCnn(float kappa, const std::vector<float> &packed);
    ~ScatterCnn();
    ScatterCnn(const ScatterCnn &) = delete;
    ScatterCnn &operator=(const ScatterCnn &) = delete;

    float kappa() const { return kappa_; }
    void getParams(std::vector<float> &packed) const;
    void setParams(const std::vector<float> &packed);

    // Network on an input [Y][X][Z][2] (device), output [Y][X][Z][2] (device).
    // bf16: inference path (BF16 MFMA for the 32->32 layers); else FP32.
    void forward(const float *in, int Y, int X, int Z, float *out, bool bf16, hipStream_t s);

    // Loss mean((out - tgt)^2) of one sample and the accumulated gradients
    // (added to the gradient buffers; zeroGrads() clears them).
    double lossGrad(const float *in, const float *tgt, int Y, int X, int Z, float gscale,
                    hipStream_t s);
    void zeroGrads(hipStream_t s);
    void getGrads(std::vector<float> &packed) const;   // oracle order
    void adamStep(float lr, hipStream_t s);

    // Source volumes of a reconstruction (engine layout [y][x][z], N x N x Nz):
    // g0, g1 from the current estimate x and attenuation map mu (per pixel),
    // in the caller's work buffers (made for the same N, Nz).  Reads only the
    // weights, so users on different streams may run it concurrently.  With
    // 'packed' the volumes are written there instead, rounded to bf16 (g0 in
    // the low, g1 in the high half of each 32-bit word).
    void sources(const float *x, const float *mu, int N, int Nz, const CnnPhys &ph,
                 float primary, float *g0, float *g1, CnnWork &w, hipStream_t s,
                 uint32_t *packed = nullptr);

    // BF16 inference weights from the master weights (after loading or
    // training; synchronous).
    void prepareInference();

    // Slices per wave of the inference path's MFMA layers (64, 32 or 16;
    // the result does not depend on it) for forward(); sources() takes the
    // work buffers' own choice (CnnWork::rowz).
    void setRowLength(int rowz);
    int rowLength() const { return rowz_; }

    // Train from ESSE-generated pairs (the model's kernel at the given size).
    static ScatterCnn *train(const EsseModel &m, const CnnTrainParms &p, CnnTrainStats &st);

private:
    void alloc_state();
    void ensure_work(long V);
    void upload_bf16_weights(hipStream_t s);
    // scal: the input's {1/mean, mean}; with 'part' (nb FP64 block sums of
    // the n values) layer 1 computes them and stores them in scal
    void forward_bf16(const float *in, int Y, int X, int Z, float *out, uint16_t *a0,
                      uint16_t *a1, float *scal, int rowz, hipStream_t s,
                      const double *part = nullptr, int nb = 0, long n = 0);
    template <int ROWZ>
    void mfma_layers(int Y, int X, int Z, float *out, uint16_t *a0, uint16_t *a1, hipStream_t s);

    int rowz_ = 64;

    float kappa_ = 1.0f;
    int step_ = 0;
    // master parameters, layout [27][Cin][Cout] (Cout fastest) per layer, and bias
    float *W_[CNN_LAYERS] = {}, *b_[CNN_LAYERS] = {};
    float *gW_[CNN_LAYERS] = {}, *gb_[CNN_LAYERS] = {};
    float *mW_[CNN_LAYERS] = {}, *vW_[CNN_LAYERS] = {}, *mb_[CNN_LAYERS] = {}, *vb_[CNN_LAYERS] = {};
    float *Wt_ = nullptr;            // flipped/transposed weights for the input gradient
    uint16_t *Wbf_[CNN_LAYERS] = {}; // bf16 [27][max(Cout,16)][Cin] (MFMA layers 2-4)
    uint16_t *Wl1_ = nullptr;        // layer 1 A fragments (bf16 [2][2][64][8])
    bool bfDirty_ = true;
    // work buffers for one sample
    long workV_ = 0;
    float *act_[CNN_LAYERS + 1] = {};   // fp32 activations (training / FP32 forward)
    float *grad_[2] = {};
    uint16_t *actb_[2] = {};             // bf16 activations (inference)
    double *dsum_ = nullptr;
    float *one_ = nullptr;               // {1, 1}: the scale of normalised input
};

// g0, g1 from the scatter model's own terms E0, E1 of the unrotated
// estimate (engine layout): the learned source with exact terms ('packed'
// as in ScatterCnn::sources).
void cnn_exact_sources(const float *E0, const float *E1, const float *x, const float *mu, int N,
                       int Nz, const CnnPhys &ph, float primary, float *g0, float *g1,
                       hipStream_t s, uint32_t *packed = nullptr);

}  // namespace irlamd
