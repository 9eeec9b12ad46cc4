// esse.h -- effective scatter source estimation (ESSE) on the GPU.
//
// Host side: the kernel grids of krnlprj3d.c:403-547 built from a KRNL file
// and folded onto a power-of-two FFT grid large enough that the circular
// convolution equals the reference's (itself circular on a 2x padded grid).
// Device side: batched FFT convolution of the rotated activity with the
// kernel, the kernel times (mu - muavg) and its square, and the collapse /
// expand steps (krnlprj3d.c:549-954).
#pragma once

#include <string>
#include <vector>

#include "engine.h"

namespace irlamd {

// Contents of a KRNL file (krnlio.h:10-30).
struct KrnlData {
    std::string title, isotope;
    int nXYoffs = 0, nDists = 0;
    float mu0 = 0, pixSize = 0, hiE = 0, loE = 0, priFac = 0, scatFac = 0, cntrPixFac = 0;
    std::vector<float> krnl, krnlMu;   // nDists x nXYoffs x nXYoffs, x fastest
};

KrnlData read_krnl_file(const char *fname);

// ESSE model prepared for one volume size (pvInitESSEScatterModel with the
// parameter string "xypad zpad nterms 3").
struct EsseModel {
    bool enabled = false;              // false when scatFac == 0 (krnlprj3d.c:428)
    int nx = 0, ny = 0, nz = 0;        // bins, depth rows, slices (full resolution)
    int c = 1;                         // collapse factor
    int nxc = 0, nyc = 0, nzc = 0;     // collapsed sizes
    int Px = 0, Py = 0, Pz = 0;        // FFT grid (line lengths the device FFT supports)
    int nterms = 2;
    float mu0 = 0, muavg = 0, pixsize = 0, cntrPixFac = 0;
    std::vector<float> grid[3];        // folded kernels on the FFT grid, [Py][Pz][Px]
};

EsseModel make_esse_model(const KrnlData &k, const char *parString, float pixsize, int nPix,
                          int nSlices, int collapse);

// Opaque handle of pvInitESSEScatterModel (device work in esse.hip).
void *esse_handle_create(const EsseModel &m);
void esse_handle_free(void *h);
bool esse_handle_enabled(const void *h);
const EsseModel *esse_handle_model(const void *h);
void esse_handle_effsrc(void *h, const float *act, const float *map, int nx, int ny, int nz,
                        float *eff);

}  // namespace irlamd
