// engine.h -- device engine of libirl_amd (host-side interface, no HIP types).
//
// The engine owns every device buffer of one reconstruction and launches the
// hand-written gfx950 kernels of engine.hip.  Host code (api.cpp) builds the
// per-view tables in the reference's float arithmetic and hands them over;
// the engine never computes geometry decisions itself.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace irlamd {

// Error raised by the engine and the host layers; mapped to the C API's
// nonzero return code plus pchIrlErrorString() at the boundary.
struct IrlError {
    int code;
    std::string msg;
};

enum ErrType {
    ETYPE_ILLEGAL_VALUE = 2,
    ETYPE_IO = 3,
    ETYPE_MALLOC = 4,
    ETYPE_SYNTAX = 5,
    ETYPE_UNKNOWN = 6,
    ETYPE_DEVICE = 7,
};

[[noreturn]] void fail(int code, const char *fmt, ...);

// DRF table padded to one common (KY x KX) support, centred.
struct PaddedDrf {
    int nDist = 0;
    int KX = 1, KY = 1;           // padded full sizes (odd)
    std::vector<float> psf;       // nDist * KY * KX, x fastest
    std::vector<int> cy;          // true half-size in y per distance
};

// Per-view constants, all computed on the host exactly as the reference.
struct ViewTab {
    float fsin, fcos;             // forward rotation (0.5*pi + Angle)
    float bsin, bcos;             // back rotation (-Angle - 0.5*pi)
    float viewtime, sensitivity;
    float fang = 0, bang = 0;     // the two rotation angles as passed (float)
};

// Geometry of the three-shear rotation of an N x N slice by one angle
// (fastrotate.c:330-563), computed on the host in the reference's float
// arithmetic.  The angle is folded into (-pi/4, pi/4] with the 90/180
// degree index flips b90/b180; the intermediate grid is mDim = N + 2*iAdd.
// Row j of the x shifts uses shift fl[j] with weights f1[j], f2[j] (the
// first and the last pass use the same values); column i of the y shift
// uses jT[i], g1[i], g2[i] and writes rows [jS[i], jE[i]).
struct ShearGeom {
    int iAdd = 0, mDim = 0, b90 = 0, b180 = 0;
    std::vector<int> fl;
    std::vector<float> f1, f2;
    std::vector<int> jT, jS, jE;
    std::vector<float> g1, g2;
};
ShearGeom shear_geometry(float fAngle, int N);
// Row extents [start, end) of the shear rotation's circle (fastrotate.c:573-589).
void shear_extents(int N, std::vector<int> &start, std::vector<int> &end);

struct Geometry {
    int N = 0;      // NumPixels
    int Nz = 0;     // NumSlices
    int NRot = 0;   // NumRotPixs (== N through IrlOsem)
    int V = 0;      // NumViews
};

struct EsseModel;
class ScatterCnn;

// Scatter-model handling of one sub-iteration (osemrecon.c:652-676, 734-788).
enum SrfMode {
    SRF_NONE = 0,        // no scatter model in the projector
    SRF_MODEL = 1,       // ESSE in the projector
    SRF_MODEL_STORE = 2, // ESSE, and store the scatter projection estimate
    SRF_ADD_STORED = 3,  // no ESSE; add the stored estimate
};

// One device-resident OSEM reconstruction (IrlOsem's state between setup and
// teardown).
class Recon {
public:
    struct Config {
        Geometry g;
        int S = 1, Vs = 1;
        int prjModel = 0, bckModel = 0;
        std::vector<int> subsets;          // S*Vs view indices
        std::vector<ViewTab> views;        // V
        PaddedDrf drf;                     // empty when DRF not modelled
        std::vector<int> rowKernel;        // V*NRot: DRF index per row, -1 = none
        std::vector<int> rowPad;           // V*NRot: axial pad rows per row
        std::vector<int> rowKernelBck;     // backprojector's, when it has its own table
        std::vector<int> rowPadBck;        //   (empty: the projector's)
        int axialAvg = 0;
        std::vector<uint8_t> colMeasOne;   // V*NRot (norm mask)
        std::vector<uint8_t> colKeep;      // V*NRot (ratio mask)
        bool hasScatterEstimate = false;
        bool hasAtn = false;
        bool shears = false;               // three-shear rotation (bFastRotate) instead of rot3dpar
        int precision = 0;                 // DRF convolution: 0 FP32, 1 BF16 (hi+lo split where
                                           // the estimated per-plane error exceeds 1e-3),
                                           // 2 BF16 hi+lo always, 3 BF16 plain always
        std::shared_ptr<const EsseModel> esse;  // scatter model, or null
        // Learned scatter: the network's terms replace the per-view FFT
        // convolutions of the scatter model while the check agrees with the
        // model (every cnnCheckEvery-th modelled subset also runs the model on
        // two of its views and compares the projected sources; rel. L2 >
        // cnnMaxNrmse switches back to the model).
        std::shared_ptr<ScatterCnn> cnn;
        int cnnCheckEvery = 0;             // 0: check the first modelled subset only
        float cnnMaxNrmse = 0.01f;
        int cnnExact = 0;                  // 1: the terms from one FFT convolution instead
        bool autotune = false;             // autotuneTiles() at precompute
        // FP32 DRF convolution: 0 spatial kernel, 1 2-D FFT kernel (where its
        // sizes fit), 2 the faster of the two per direction, timed at
        // precompute
        int drfFft = 0;
        // scatter model: 2-D plane transforms and the source formed in the
        // last FFT pass where the sizes allow (else the line passes and the
        // separate source kernel with the reference's FP64 exponential)
        bool esseFused = true;
    };

    struct CnnStatus {
        bool configured = false, active = false;
        int checks = 0, used = 0;          // checked subsets, subsets run on the network
        float lastNrmse = -1, worstNrmse = 0;   // projected source (primary + scatter)
        float lastScatterNrmse = -1, worstScatterNrmse = 0;   // its scatter part alone
    };
    CnnStatus cnnStatus() const;

    explicit Recon(const Config &cfg);
    ~Recon();

    // Host -> device uploads (volumes in [z][y][x] order, projections
    // [view][slice][bin]).
    void uploadEstimate(const float *x);
    void uploadAtn(const float *mu);
    void uploadProjections(const float *prj);
    void uploadScatterEstimate(const float *scat);

    // Attenuation factors of every view (reused by all sub-iterations) and,
    // when 'norms' is set, the FP64 sensitivity images of every subset.
    void precompute(bool norms = true);

    // Projection of the current estimate for every view (IrlGenprj's
    // per-view output before the 1/NumViews scaling), [view][slice][bin].
    void forwardProject(float primary, float *out_host, bool srf = false);

    // One OSEM sub-iteration over subset s (osemrecon.c:687-863 body).  With
    // setGraphs(true) a sub-iteration is replayed from a captured graph from
    // the second occurrence of its (subset, mode) on.
    void subIteration(int s, bool keepUpdate, SrfMode srf = SRF_NONE);
    void setGraphs(bool on);
    // With graphs on: captures and instantiates subset s's sub-iteration for
    // the mode now, so that its first occurrence already replays (subsets
    // with a learned-scatter check still run eagerly).  Executes nothing.
    void prepareGraph(int s, bool keepUpdate, SrfMode srf);

    // Pieces of subIteration for drivers of several reconstructions: the
    // learned-scatter bookkeeping around a sub-iteration (beginSubset
    // returns true when the subset must run eagerly: its check reads a
    // result on the host), the launches themselves, and the launches
    // captured into a graph (a hipGraph_t, not executed; owned by the caller).
    bool beginSubset(SrfMode srf);
    void endSubset(SrfMode srf);
    void runSubIteration(int s, SrfMode srf) { runViews(s, 0, cfg_.Vs, srf, false, true, false); }
    void *captureSubIteration(int s, bool keepUpdate, SrfMode srf);
    // The sub-iteration's launches on the engine's stream while that stream
    // takes part in a capture begun elsewhere (joined by an event wait).
    void recordSubIteration(int s, SrfMode srf);
    void *streamHandle() const;       // the engine's hipStream_t
    bool learnedScatterActive() const;

    // Tile shapes of the DRF kernels chosen by timing the candidates on
    // subset 0 (slice groups per block of the FP32 kernel, slice tiles per
    // wave of the BF16 kernel), for the projector and the backprojector.
    void autotuneTiles();
    int cnnRowLength() const;   // the scatter network's row length in use (0: none)
    int prjRowsPerChunk() const;   // the spatial projector's rows per partial plane
    void tiles(int *prj, int *bck) const;
    void setTiles(int prj, int bck);   // 0: size heuristic; drops captured graphs
    // FP32 DRF algorithm in use per direction: 0 spatial, 1 FFT (length P)
    void drfAlgorithms(int *prj, int *bck, int *P) const;
    // BF16 DRF arithmetic per table entry (the delta of unconvolved rows
    // last): 0 plain, 1 weights hi+lo, 2 weights and data hi+lo, with the
    // measured relative errors of plain BF16 and of the weight split (zeros
    // when not measured).  Empty on the FP32 path.
    void bf16Planes(std::vector<int> &modes, std::vector<float> &err) const;

    // One view at a time (for per-angle hooks): zeroUpdate at the start of
    // the subset, then member m of subset s is projected and backprojected
    // into U; the update of x follows with finishSubset.
    void zeroUpdate();
    void angleStep(int s, int m, SrfMode srf = SRF_NONE);
    void finishSubset(int s);

    // Stored scatter projection estimate (pfPrjScatEst), [view][slice][bin].
    void downloadScatterProjections(float *h) const;

    // Device -> host downloads.
    void downloadEstimate(float *x) const;
    void downloadUpdate(float *u) const;
    void downloadNorm(int s, float *ninv) const;
    void uploadEstimateFromHost(const float *x) { uploadEstimate(x); }
    void uploadUpdate(const float *u);
    void uploadNorm(int s, const float *ninv);

    // Per-kernel timing of the last sub-iterations (HIP events on the
    // engine's stream).  Enabled by setTiming(true).
    void setTiming(bool on);
    struct StageTime {
        std::string name;
        double ms_total = 0;
        int launches = 0;
    };
    std::vector<StageTime> stageTimes() const;
    void resetTimes();
    void synchronize() const;

    const Config &config() const { return cfg_; }

private:
    void runViews(int s, int first, int B, SrfMode srf, bool accumulate, bool update,
                  bool keepUpdate);
    struct Impl;
    Impl *impl_;
    Config cfg_;
};

// Several reconstructions stepped together: one sub-iteration of each per
// step.  The ones that can be replayed run from captured graphs, each on its
// own stream (concurrently), or with setParentGraph(true) as child graphs
// of one parent graph (one launch per step); the others (first occurrence
// of their subset and mode, learned-scatter checks) run eagerly on their
// own streams after them.
class ReconBatch {
public:
    explicit ReconBatch(const std::vector<Recon *> &recons);
    ~ReconBatch();
    ReconBatch(const ReconBatch &) = delete;
    ReconBatch &operator=(const ReconBatch &) = delete;
    void step(const std::vector<int> &subset, const std::vector<SrfMode> &srf);
    // 0: each member's graph on its own stream; 1: the members' graphs as
    // child graphs of one parent graph; 2: one graph captured with the
    // members' launches as parallel branches.  One launch per step in 1, 2.
    void setParentGraph(int mode);
    void synchronize();
    int graphLaunches() const;       // one-launch-per-step launches so far
    int replays() const;             // sub-iterations replayed from graphs so far
    int eagerSubIterations() const;  // sub-iterations run eagerly so far

private:
    struct Impl;
    Impl *impl_;
};

// One view of the projector-level entry points.  Tables come from the host
// layer, computed with the reference's float arithmetic.
struct PrjCall {
    int Nz, N, NRot, model;
    float fsin, fcos, bsin, bcos;  // rotation constants of this view
    float primary, atnFac;
    int avg;
    int atnRotated, actRotated;
    const int *rowKernel;          // NRot entries: DRF index or -1
    const int *rowPad;             // NRot entries
    const void *drfKey;            // identity of the DRF table (cache key)
    const PaddedDrf *drf;          // null when DRF is not modelled
    const void *esse;              // pvInitESSEScatterModel handle (MODEL_SRF), may be null
    float primaryFac;              // fPrimaryFac as passed (scatter source scaling)
    int fast;                      // bFastRotate: three-shear rotation
    float fang, bang;              // forward and back rotation angles (float)
};

void prjws_init(int Nz, int N, int NRot, int model);
void prjws_free();
void prjws_project(const PrjCall &c, const float *atn, const float *recon, float *calcprj);
void prjws_backproject(const PrjCall &c, const float *atn, const float *ratio, float *image);
// The same with no device allocation (a file-scope device array holds the
// staging); psf is KY x KX, x fastest.
void dev_convolve_plane_noalloc(const float *in, int inX, int inY, const float *psf, int KX, int KY,
                                int pad, int avg, float *out);
void dev_convolve_plane(const float *in, int inX, int inY, const PaddedDrf &d, int pad, int avg,
                        float *out);

// Stateless single-call helpers behind the prjbck-level C entry points.
void dev_rot3dpar(float angle, const float *in_host, float *out_host, int pbFlag,
                  int Nz, int N, int NRot);
void dev_shears_rotate(float angle, const float *in_host, float *out_host, int Nz, int N);
void dev_rot2d_views(const float *img_host, int N, int NRot, const float *angles, int V,
                     float *out_host);

}  // namespace irlamd
