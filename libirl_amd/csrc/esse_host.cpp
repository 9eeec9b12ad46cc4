// esse_host.cpp -- KRNL file reader and the host-side ESSE kernel grids.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "esse.h"
#include "irlprj.h"

namespace irlamd {

static void swap4(void *p, size_t n)
{
    unsigned char *b = (unsigned char *)p;
    for (size_t i = 0; i < n; ++i, b += 4) {
        std::swap(b[0], b[3]);
        std::swap(b[1], b[2]);
    }
}

// krnlio.c:196-331.  Format: "KRNL", int version 1, char title[80], char
// isotope[20], float mu0, hiE, loE, pixsize, priFac, scatFac, cntrPixFac,
// int nXYoffs, nDists, then float krnl[nDists][nXYoffs][nXYoffs] and
// krnlMu[...].  As in the reference, a byte-swapped file has its nXYoffs left
// unswapped (krnlio.c:293 swaps nDists there), which the size checks reject.
KrnlData read_krnl_file(const char *fname)
{
    KrnlData k;
    FILE *fp = fopen(fname, "rb");
    if (!fp) fail(ETYPE_IO, "Error opening Krnl Table file: %s", fname);
    struct Closer {
        FILE *f;
        ~Closer() { fclose(f); }
    } closer{fp};
    auto rd = [&](void *dst, size_t sz, size_t n, const char *what) {
        if (fread(dst, sz, n, fp) != n || ferror(fp))
            fail(ETYPE_IO, "IO Error for file %s : reading %s", fname, what);
    };
    char sig[5] = {0};
    rd(sig, 1, 4, "signature");
    if (strncmp(sig, "KRNL", 4))
        fail(ETYPE_ILLEGAL_VALUE, "Signature in Krnl File %s is %s, should be KRNL", fname, sig);
    int version;
    rd(&version, sizeof(int), 1, "version");
    bool swapped = false;
    if (version != 1) {
        swap4(&version, 1);
        if (version != 1) {
            swap4(&version, 1);
            fail(ETYPE_ILLEGAL_VALUE, "Version in Krnl File %s is  %x hex, should be 1", fname,
                 version);
        }
        swapped = true;
    }
    char t[81] = {0}, iso[21] = {0};
    rd(t, 1, 80, "title");
    rd(iso, 1, 20, "isotope name");
    k.title = std::string(t, strnlen(t, 80));
    k.isotope = std::string(iso, strnlen(iso, 20));
    float f[7];
    rd(f, sizeof(float), 7, "header floats");
    if (swapped) swap4(f, 7);
    k.mu0 = f[0];
    k.hiE = f[1];
    k.loE = f[2];
    k.pixSize = f[3];
    k.priFac = f[4];
    k.scatFac = f[5];
    k.cntrPixFac = f[6];
    rd(&k.nXYoffs, sizeof(int), 1, "NumXYoffs");
    rd(&k.nDists, sizeof(int), 1, "NumDistances");
    if (swapped) swap4(&k.nDists, 1);
    if (k.nDists <= 0 || k.nXYoffs < 0)
        fail(ETYPE_ILLEGAL_VALUE, "sizes of kernel must be > 0 not numdists=%d, numXYoffs=%d",
             k.nDists, k.nXYoffs);
    const long long n = (long long)k.nDists * k.nXYoffs * k.nXYoffs;
    if (n > (1LL << 28))
        fail(ETYPE_ILLEGAL_VALUE, "kernel in %s is too large (%d x %d x %d)", fname, k.nDists,
             k.nXYoffs, k.nXYoffs);
    k.krnl.resize(n);
    k.krnlMu.resize(n);
    rd(k.krnl.data(), sizeof(float), n, "Krnl");
    rd(k.krnlMu.data(), sizeof(float), n, "KrnlMu");
    if (swapped) {
        swap4(k.krnl.data(), n);
        swap4(k.krnlMu.data(), n);
    }
    return k;
}

// krnlprj3d.c:43-127: mirror the one-quadrant kernel into the wrap-around
// grid [outz=depth][outy=slices][outx=offs], depth flipped and centred.
static std::vector<float> reorder_krnl(const float *in, int noffs, int ndepths, int outx,
                                       int outy, int outz)
{
    std::vector<float> o((size_t)outx * outy * outz, 0.0f);
    const int xmin = std::min(noffs, outx / 2), ymin = std::min(noffs, outy / 2);
    for (int izin = 0; izin < ndepths; ++izin) {
        int izout = outz / 2 - izin + ndepths / 2 - 1;
        if (izout < 0 || izout >= outz) continue;
        izout += izout < outz / 2 ? outz / 2 : -outz / 2;
        for (int iyin = 0; iyin < ymin; iyin++) {
            for (int s = 0; s < 2; ++s) {
                if (s == 1 && iyin == 0) continue;
                const int iyout = s == 0 ? iyin : outy - iyin;
                for (int ixin = 0; ixin < xmin; ixin++) {
                    const float v = in[ixin + (size_t)noffs * (iyin + (size_t)noffs * izin)];
                    o[ixin + (size_t)outx * (iyout + (size_t)outy * izout)] = v;
                    if (ixin > 0) o[outx - ixin + (size_t)outx * (iyout + (size_t)outy * izout)] = v;
                }
            }
        }
    }
    return o;
}

// Smallest line-FFT length of the device FFT (esse.hip) that is >= v, or 0.
static int fft_length(int v)
{
    for (int p : {16, 32, 64, 128, 160, 192, 256, 512})
        if (p >= v) return p;
    return 0;
}

// krnlprj3d.c:956-1009 and 403-547.
EsseModel make_esse_model(const KrnlData &k, const char *parString, float pixsize, int nPix,
                          int nSlices, int collapse)
{
    EsseModel m;
    float xypad = 0, zpad = 0;
    int nterms = 0, geom = 0;
    if (sscanf(parString, "%f %f %d %d", &xypad, &zpad, &nterms, &geom) < 0)
        fail(ETYPE_ILLEGAL_VALUE, "illegal srf par string");
    if (geom != 2 && geom != 3) fail(ETYPE_ILLEGAL_VALUE, "geometry must be 2 or 3");
    if (nterms < 1 || nterms > 3) fail(ETYPE_ILLEGAL_VALUE, "numexpterms must be 1, 2, or 3");
    if (fabs(k.pixSize - pixsize * collapse) > 0.02)
        warn("Pixel size after collapse=%.3g, Pixel size in kernel=%.3g", collapse * pixsize,
             k.pixSize);
    if (k.scatFac == 0.0f) return m;  // no scatter data: effective source is zero
    if (geom != 3) fail(ETYPE_ILLEGAL_VALUE, "2D not supported yet, sorry");
    if (nPix % collapse || nSlices % collapse)
        fail(ETYPE_ILLEGAL_VALUE,
             "factor does not evenly divide image dimensions \n     iFac=%d, nx=%d, ny=%d, nz=%d",
             collapse, nPix, nSlices, nPix);
    m.enabled = true;
    m.nx = nPix;
    m.ny = nPix;
    m.nz = nSlices;
    m.c = collapse;
    m.nxc = nPix / collapse;
    m.nyc = nPix / collapse;
    m.nzc = nSlices / collapse;
    m.nterms = nterms;
    m.mu0 = k.mu0;
    m.pixsize = pixsize;
    m.cntrPixFac = k.cntrPixFac;
    const int Kx = (int)ceil(xypad * (float)m.nxc);
    const int Ky = (int)ceil(xypad * (float)m.nyc);
    const int Kz = (int)ceil(zpad * (float)m.nzc);
    std::vector<float> G[3];
    G[0] = reorder_krnl(k.krnl.data(), k.nXYoffs, k.nDists, Kx, Kz, Ky);
    std::vector<float> Mu = reorder_krnl(k.krnlMu.data(), k.nXYoffs, k.nDists, Kx, Kz, Ky);
    const size_t n = G[0].size();
    m.muavg = k.mu0;
    if (nterms > 1) {
        float sK = 0.0f, sKM = 0.0f;
        for (size_t i = 0; i < n; ++i) {
            sK += G[0][i];
            sKM += G[0][i] * Mu[i];
        }
        if (sK == 0.0f)
            warn("Sum of Krnl is 0!");
        else
            m.muavg = sKM / sK;
        G[1].resize(n);
        for (size_t i = 0; i < n; ++i) {
            Mu[i] -= m.muavg;
            G[1][i] = Mu[i] * G[0][i];
        }
        if (nterms > 2) {
            G[2].resize(n);
            for (size_t i = 0; i < n; ++i) G[2][i] = Mu[i] * G[1][i];
        }
    }
    // Fold each grid onto the offsets a cropped output can see, d in
    // (-(n'-1), n'-1) per axis, F[d] = G[d mod K]; find the support box.
    const int n3[3] = {m.nyc, m.nzc, m.nxc}, K3[3] = {Ky, Kz, Kx};
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int t = 0; t < nterms; ++t)
        for (int gy = 0; gy < Ky; ++gy)
            for (int gz = 0; gz < Kz; ++gz)
                for (int gx = 0; gx < Kx; ++gx) {
                    if (G[t][((size_t)gy * Kz + gz) * Kx + gx] == 0.0f) continue;
                    const int g3[3] = {gy, gz, gx};
                    for (int a = 0; a < 3; ++a)
                        for (int d : {g3[a], g3[a] - K3[a], g3[a] + K3[a]})
                            if (d > -n3[a] && d < n3[a]) {
                                lo[a] = std::min(lo[a], d);
                                hi[a] = std::max(hi[a], d);
                            }
                }
    for (int a = 0; a < 3; ++a)
        if (lo[a] > hi[a]) lo[a] = hi[a] = 0;
    // Circular convolution of length P of a line zero-padded from n equals
    // the linear one on the cropped outputs i < n when no input index i - d,
    // d in [lo, hi], wraps into [0, n): P >= n + hi and P >= n - lo (then
    // also P > hi - lo, so the kernel's offsets land on distinct cells).
    int P3[3];
    for (int a = 0; a < 3; ++a) {
        const int need = n3[a] + std::max(std::max(hi[a], -lo[a]), 0);
        P3[a] = fft_length(need);
        if (P3[a] == 0)
            fail(ETYPE_ILLEGAL_VALUE, "ESSE FFT length %d along axis %d exceeds 512", need, a);
    }
    m.Py = P3[0];
    m.Pz = P3[1];
    m.Px = P3[2];
    for (int t = 0; t < nterms; ++t) {
        m.grid[t].assign((size_t)m.Px * m.Py * m.Pz, 0.0f);
        for (int dy = -(m.nyc - 1); dy <= m.nyc - 1; ++dy) {
            const int gy = ((dy % Ky) + Ky) % Ky, py = ((dy % m.Py) + m.Py) % m.Py;
            for (int dz = -(m.nzc - 1); dz <= m.nzc - 1; ++dz) {
                const int gz = ((dz % Kz) + Kz) % Kz, pz = ((dz % m.Pz) + m.Pz) % m.Pz;
                const float *src = &G[t][((size_t)gy * Kz + gz) * Kx];
                float *dst = &m.grid[t][((size_t)py * m.Pz + pz) * m.Px];
                for (int dx = -(m.nxc - 1); dx <= m.nxc - 1; ++dx) {
                    const float v = src[((dx % Kx) + Kx) % Kx];
                    if (v != 0.0f) dst[((dx % m.Px) + m.Px) % m.Px] += v;
                }
            }
        }
    }
    return m;
}

}  // namespace irlamd
