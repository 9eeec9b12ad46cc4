// host_util.cpp -- host-side plumbing of libirl_amd: status messages and
// warnings, DRF table I/O, and the reference-arithmetic geometry tables that
// the kernels consume.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "irlprj.h"

namespace irlamd {

static int g_msgLevel = 4;
static FILE *g_msgFile = nullptr;
static FILE *g_logFile = nullptr;
static std::string g_errString;
static int g_lastErr = 0;

void set_msg_level(int level) { g_msgLevel = level; }
void set_msg_file(FILE *fp) { g_msgFile = fp; }
void set_log_file(FILE *fp) { g_logFile = fp; }
void set_error_string(const std::string &s) { g_errString = s; }
void set_last_error(int code) { g_lastErr = code; }
int last_error() { return g_lastErr; }
const std::string &error_string() { return g_errString; }

void msg(int level, const char *fmt, ...)
{
    if (!g_msgFile || level > g_msgLevel) return;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(g_msgFile, fmt, ap);
    va_end(ap);
    fflush(g_msgFile);
}

void warn(const char *fmt, ...)
{
    if (!g_logFile) return;
    va_list ap;
    va_start(ap, fmt);
    fprintf(g_logFile, "WARNING: ");
    vfprintf(g_logFile, fmt, ap);
    fprintf(g_logFile, "\n");
    va_end(ap);
    fflush(g_logFile);
}

// prjbck.c:186-204.  Float arithmetic as in the reference: the start distance
// is formed with a double half-pixel term and stored to float, the plane
// index is floor of a float quotient plus 0.5 (double).
int drf_row_index(const DrfTab_t *t, float fCFCR, float fPixelWidth, int iNumRotPixs, int iRow,
                  bool *skip)
{
    float fStartDist = fCFCR - fPixelWidth * iNumRotPixs / 2 + 0.5 * fPixelWidth;
    float fDist = fStartDist + iRow * fPixelWidth;
    int iDist = (int)floor((fDist - t->fStartDistance) / t->fDrfSpacing + 0.5);
    if (iDist < 0) iDist = 0;
    if (iDist > t->iNumDistances - 1) iDist = t->iNumDistances - 1;
    *skip = !(fDist > 0.0);
    return iDist;
}

// Pads every PSF of the table to the table's largest support, centred, so
// one kernel instantiation serves all planes.  Zero taps add exact zeros.
PaddedDrf pad_drf_table(const DrfTab_t *t)
{
    PaddedDrf p;
    p.nDist = t->iNumDistances;
    int cxm = 0, cym = 0;
    for (int d = 0; d < t->iNumDistances; ++d) {
        const Drf_t &D = t->psDrfs[d];
        if (D.bFFTdone)
            fail(ETYPE_ILLEGAL_VALUE, "DRF %d is already Fourier transformed", d);
        if (D.MaxOffsX < 1 || D.MaxOffsY < 1)
            fail(ETYPE_ILLEGAL_VALUE, "DRF %d has illegal offsets %d x %d", d, D.MaxOffsX,
                 D.MaxOffsY);
        cxm = std::max(cxm, D.MaxOffsX - 1);
        cym = std::max(cym, D.MaxOffsY - 1);
    }
    p.KX = 2 * cxm + 1;
    p.KY = 2 * cym + 1;
    p.psf.assign((size_t)p.nDist * p.KX * p.KY, 0.0f);
    p.cy.resize(p.nDist);
    for (int d = 0; d < p.nDist; ++d) {
        const Drf_t &D = t->psDrfs[d];
        const int cx = D.MaxOffsX - 1, cy = D.MaxOffsY - 1;
        const int w = 2 * cx + 1, h = 2 * cy + 1;
        p.cy[d] = cy;
        float *dst = p.psf.data() + (size_t)d * p.KX * p.KY;
        for (int j = 0; j < h; ++j)
            for (int i = 0; i < w; ++i)
                dst[(j + cym - cy) * p.KX + (i + cxm - cx)] = D.Psf[j * w + i];
    }
    return p;
}

// Rotation constants of one view: forward angle 0.5*pi+Angle (prjbck.c:135)
// and backward angle -Angle-0.5*pi (prjbck.c:413), each passed to rot3dpar as
// float and turned into float sin/cos computed in double (rot3dpar.c:51-52).
void view_trig(float angle, ViewTab &vt)
{
    const float fa = 0.5 * M_PI + angle;
    const float ba = -angle - 0.5 * M_PI;
    vt.fsin = (float)sin(fa);
    vt.fcos = (float)cos(fa);
    vt.bsin = (float)sin(ba);
    vt.bcos = (float)cos(ba);
    vt.fang = fa;
    vt.bang = ba;
}

void shear_extents(int N, std::vector<int> &start, std::vector<int> &end)
{
    start.resize(N);
    end.resize(N);
    const float c = (N - 1) * 0.5;
    for (int j = 0; j < N; ++j) {
        const float d = (float)j - c;
        const int len = 2 * floor(sqrt(c * c - d * d)) + 2;
        start[j] = (N - len) / 2;
        end[j] = start[j] + len;
    }
}

// fastrotate.c:358-388 (folding, grid size) and the per-row / per-column
// shift decisions of its three passes (:396-404, :489-500, :520-530).
ShearGeom shear_geometry(float fAngle, int N)
{
    ShearGeom g;
    while (fAngle > 0.75 * M_PI) {
        fAngle -= M_PI;
        g.b180 ^= 1;
    }
    while (fAngle <= -M_PI / 4.0) {
        fAngle += M_PI;
        g.b180 ^= 1;
    }
    if (fAngle > M_PI / 4.0) {
        g.b90 = 1;
        fAngle -= M_PI / 2.0;
    }
    const float fCntr = ((float)N - 1.0) * 0.5;
    int iAdd = 0;
    for (int j = N / 4; j < N / 2; ++j) {
        const float yDist = (float)j - fCntr;
        float fDelta = fabs(yDist * tan((double)(fAngle * 0.5)));
        fDelta -= fCntr - sqrt(fCntr * fCntr - yDist * yDist);
        const int t = ceil(fDelta);
        if (t > iAdd) iAdd = t;
    }
    g.iAdd = iAdd;
    g.mDim = N + 2 * iAdd;
    const float tCntr = ((float)g.mDim - 1.0) * 0.5;
    g.fl.resize(N);
    g.f1.resize(N);
    g.f2.resize(N);
    for (int j = 0; j < N; ++j) {
        const float yDist = (float)j - fCntr;
        const float fDelta = -1.0 * yDist * tan((double)(fAngle * 0.5));
        const int fl = floorf(fDelta);
        const float f2 = fDelta - (float)fl;
        const float f1 = 1.0 - f2;
        g.fl[j] = fl;
        g.f1[j] = f1;
        g.f2[j] = f2;
    }
    g.jT.resize(g.mDim);
    g.jS.resize(g.mDim);
    g.jE.resize(g.mDim);
    g.g1.resize(g.mDim);
    g.g2.resize(g.mDim);
    for (int i = 0; i < g.mDim; ++i) {
        const float xDist = (float)i - tCntr;
        const float fDelta = xDist * sin((double)fAngle);
        const int jT = floorf(fDelta);
        const float f2 = fDelta - (float)jT;
        const float f1 = 1.0 - f2;
        g.jT[i] = jT;
        g.g1[i] = f1;
        g.g2[i] = f2;
        g.jS[i] = std::max(0, -jT);
        g.jE[i] = std::min(g.mDim, g.mDim - jT - 1);
    }
    return g;
}

// ---------------------------------------------------------------------------
// DRFT file (drftabio.c:21-46 format): "DRFT", int version 1 (byte order
// detected from it), char title[80], char col[20], float pixsize, spacing,
// start, int nDist, int bIncKrnl; per distance int xoffs, yoffs and
// (2xoffs-1)(2yoffs-1) floats, x fastest.
// ---------------------------------------------------------------------------
static void swap4(void *p, size_t n)
{
    unsigned char *b = (unsigned char *)p;
    for (size_t i = 0; i < n; ++i, b += 4) {
        std::swap(b[0], b[3]);
        std::swap(b[1], b[2]);
    }
}

DrfTab_t read_drf_tab(const char *fname, std::string *title, std::string *col)
{
    DrfTab_t t;
    memset(&t, 0, sizeof t);
    t.bDrfFromFile = 1;
    FILE *fp = fopen(fname, "rb");
    if (!fp) fail(ETYPE_IO, "Error opening DRF Table file: %s", fname);
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
    if (strncmp(sig, "DRFT", 4))
        fail(ETYPE_ILLEGAL_VALUE, "Signature in Drf File %s is %s, should be DRFT", fname, sig);
    int version;
    rd(&version, sizeof(int), 1, "version");
    bool swapped = false;
    if (version != 1) {
        swap4(&version, 1);
        if (version != 1) {
            swap4(&version, 1);
            fail(ETYPE_ILLEGAL_VALUE, "Version in Drf File %s is  %x hex, should be 1", fname,
                 version);
        }
        swapped = true;
    }
    char tbuf[81] = {0}, cbuf[21] = {0};
    rd(tbuf, 1, 80, "title");
    rd(cbuf, 1, 20, "colname");
    if (title) *title = std::string(tbuf, strnlen(tbuf, 80));
    if (col) *col = std::string(cbuf, strnlen(cbuf, 20));
    rd(&t.fPixSize, sizeof(float), 1, "PixSize");
    rd(&t.fDrfSpacing, sizeof(float), 1, "DrfSpacing");
    rd(&t.fStartDistance, sizeof(float), 1, "StartDistance");
    rd(&t.iNumDistances, sizeof(int), 1, "NumDistances");
    rd(&t.bIncKrnl, sizeof(int), 1, "bIncKrnl");
    if (swapped) {
        swap4(&t.fPixSize, 1);
        swap4(&t.fDrfSpacing, 1);
        swap4(&t.fStartDistance, 1);
        swap4(&t.iNumDistances, 1);
        swap4(&t.bIncKrnl, 1);
    }
    if (t.iNumDistances <= 0 || t.iNumDistances > 1000000)
        fail(ETYPE_ILLEGAL_VALUE, "DRF table %s has %d distances", fname, t.iNumDistances);
    t.psDrfs = (Drf_t *)calloc(t.iNumDistances, sizeof(Drf_t));
    if (!t.psDrfs) fail(ETYPE_MALLOC, "ReadDrfTab:Drfs");
    try {
        for (int i = 0; i < t.iNumDistances; ++i) {
            int xo, yo;
            rd(&xo, sizeof(int), 1, "drf offsets");
            rd(&yo, sizeof(int), 1, "drf offsets");
            if (swapped) {
                swap4(&xo, 1);
                swap4(&yo, 1);
            }
            if (xo < 1 || yo < 1 || xo > 4096 || yo > 4096)
                fail(ETYPE_ILLEGAL_VALUE, "DRF %d in %s has illegal offsets %d x %d", i, fname,
                     xo, yo);
            Drf_t &D = t.psDrfs[i];
            D.MaxOffsX = xo;
            D.MaxOffsY = yo;
            D.bFFTdone = 0;
            const size_t n = (size_t)(2 * xo - 1) * (2 * yo - 1);
            D.Psf = (float *)malloc(n * sizeof(float));
            if (!D.Psf) fail(ETYPE_MALLOC, "ReadDrfTab:Drfs[i].Psf");
            char what[64];
            snprintf(what, sizeof what, "drf %d", i);
            rd(D.Psf, sizeof(float), n, what);
            if (swapped) swap4(D.Psf, n);
        }
    } catch (...) {
        free_drf_tab_contents(&t);
        throw;
    }
    return t;
}

void free_drf_tab_contents(DrfTab_t *t)
{
    if (!t || !t->psDrfs) return;
    for (int i = 0; i < t->iNumDistances; ++i) free(t->psDrfs[i].Psf);
    free(t->psDrfs);
    t->psDrfs = nullptr;
}

}  // namespace irlamd

extern "C" {

char *pchIrlErrorString(void) { return (char *)irlamd::error_string().c_str(); }

DrfTab_t sReadDrfTab(char *pchInFname, char **ppchInTitle, char **ppchInColName)
{
    std::string title, col;
    DrfTab_t t;
    memset(&t, 0, sizeof t);
    try {
        t = irlamd::read_drf_tab(pchInFname, &title, &col);
        irlamd::set_error_string("");
    } catch (const irlamd::IrlError &e) {
        irlamd::set_error_string(e.msg);
        fprintf(stderr, "sReadDrfTab: %s\n", e.msg.c_str());
        if (ppchInTitle) *ppchInTitle = nullptr;
        if (ppchInColName) *ppchInColName = nullptr;
        return t;
    }
    if (ppchInTitle) *ppchInTitle = strdup(title.c_str());
    if (ppchInColName) *ppchInColName = strdup(col.c_str());
    return t;
}

void vWriteDrfTab(DrfTab_t *t, char *pchOutFname, char *pchTitle, char *pchColName)
{
    FILE *fp = fopen(pchOutFname, "wb");
    if (!fp) {
        // drftabio.c:73-77
        irlamd::set_error_string(std::string("Error opening DRF Table file: ") + pchOutFname);
        return;
    }
    char tbuf[80] = {0}, cbuf[20] = {0};
    if (pchTitle) strncpy(tbuf, pchTitle, 79);
    if (pchColName) strncpy(cbuf, pchColName, 19);
    const int version = 1;
    fwrite("DRFT", 1, 4, fp);
    fwrite(&version, sizeof(int), 1, fp);
    fwrite(tbuf, 1, 80, fp);
    fwrite(cbuf, 1, 20, fp);
    fwrite(&t->fPixSize, sizeof(float), 1, fp);
    fwrite(&t->fDrfSpacing, sizeof(float), 1, fp);
    fwrite(&t->fStartDistance, sizeof(float), 1, fp);
    fwrite(&t->iNumDistances, sizeof(int), 1, fp);
    fwrite(&t->bIncKrnl, sizeof(int), 1, fp);
    for (int i = 0; i < t->iNumDistances; ++i) {
        const Drf_t &D = t->psDrfs[i];
        fwrite(&D.MaxOffsX, sizeof(int), 1, fp);
        fwrite(&D.MaxOffsY, sizeof(int), 1, fp);
        fwrite(D.Psf, sizeof(float), (size_t)(2 * D.MaxOffsX - 1) * (2 * D.MaxOffsY - 1), fp);
    }
    fclose(fp);
}

}  // extern "C"
