// orbit.cpp -- ComputeOrbit (computeorbit.c:21-150): camera-to-centre
// distances of an orbit from a 2-D image of the object.  The image is
// rotated for every view on the GPU (rot3dpar, one batch); the host then
// finds the first row, counted from the detector side, that holds any pixel
// above the threshold.
#include <cmath>
#include <vector>

#include "irlprj.h"

using namespace irlamd;

extern "C" int ComputeOrbit(float *pfPixels, IrlParms_t *psParms, float fMargin,
                            float *pfMargins, float fThresh, int OrbitType, PrjView_t *psViews)
{
    const int V = psParms->NumViews, N = psParms->NumPixels;
    const float pix = psParms->BinWidth;
    if (OrbitType == 0) {   // circular orbit: the margin is the radius
        for (int v = 0; v < V; ++v) psViews[v].CFCR = fMargin;
        return 0;
    }
    if (OrbitType > 2 || OrbitType < 0) return 2;
    // threshold in place (computeorbit.c:82-95)
    float thr = 0.0f;
    if (fThresh <= 0.0) {
        thr = -fThresh;
    } else {
        float mx = pfPixels[0];
        for (long i = 1; i < (long)N * N; ++i) mx = pfPixels[i] > mx ? pfPixels[i] : mx;
        thr = mx * fThresh;
    }
    for (long i = 0; i < (long)N * N; ++i)
        if (pfPixels[i] <= thr) pfPixels[i] = 0.0f;
    // rotated image grown to cover an off-centre object (computeorbit.c:97-99)
    int NRot = (int)ceil(sqrt(2.0) * (double)N);
    if ((NRot % 2) != (N % 2)) NRot++;
    std::vector<float> angles(V), rot;
    for (int v = 0; v < V; ++v) angles[v] = (float)(0.5 * M_PI + psViews[v].Angle);
    try {
        rot.resize((size_t)V * NRot * NRot);
        dev_rot2d_views(pfPixels, N, NRot, angles.data(), V, rot.data());
    } catch (const IrlError &e) {
        set_error_string(e.msg);
        return 1;
    } catch (const std::bad_alloc &) {
        return 1;
    }
    for (int v = 0; v < V; ++v) {
        const float *img = rot.data() + (size_t)v * NRot * NRot;
        // first row from the detector side with a pixel above the threshold;
        // the loop index ends one past it, or at NRot/2 (computeorbit.c:118-121)
        float sum = 0.0f;
        int row = 0;
        for (; row < NRot / 2 && sum <= 0.0f; ++row) {
            sum = 0.0f;
            for (int x = 0; x < NRot; ++x) sum += img[(size_t)row * NRot + x];
        }
        float r = ((float)NRot / 2.0 - (float)row) * pix;
        r += pfMargins ? pfMargins[v] : fMargin;
        psViews[v].CFCR = r;
    }
    if (OrbitType == 2) {   // L-mode: views v and v+V/2 share the larger radius
        for (int v = 0; v < V / 2; ++v) {
            if (psViews[v].CFCR > psViews[v + V / 2].CFCR)
                psViews[v + V / 2].CFCR = psViews[v].CFCR;
            else
                psViews[v].CFCR = psViews[v + V / 2].CFCR;
        }
    }
    return 0;
}
