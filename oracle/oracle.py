"""ctypes bindings of the CPU oracle (oracle/irl_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product (libirl_amd) never imports it.

Two builds: 'f32' (reference semantics, float values) and 'f64' (FP64
oracle).  Volumes are the reference's internal [y][z][x] order unless noted.
"""
import ctypes
import os

import numpy as np

from . import build as _build

_libs = {}


class OrcView(ctypes.Structure):
    _fields_ = [("angle", ctypes.c_float), ("left", ctypes.c_float),
                ("right", ctypes.c_float), ("cfcr", ctypes.c_float),
                ("viewtime", ctypes.c_float), ("sensitivity", ctypes.c_float)]


def lib(real="f64"):
    if real in _libs:
        return _libs[real]
    path = _build.lib_path(real)
    if not os.path.exists(path):
        _build.build(force=False)
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.orc_drftab_new.restype = vp
    L.orc_drftab_new.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, vp, vp]
    L.orc_drftab_free.argtypes = [vp]
    L.orc_rot3dpar.argtypes = [ctypes.c_float, vp, vp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int]
    L.orc_conv2d_ypad.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    L.orc_init_prj.argtypes = [ctypes.c_int] * 4
    L.orc_drf_index.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int]
    L.orc_project.argtypes = ([ctypes.c_int] * 4 + [ctypes.c_float] * 3 +
                              [vp, ctypes.c_int, vp, vp, vp, vp] + [ctypes.c_float] * 2 +
                              [ctypes.c_int] * 5)
    L.orc_backproject.argtypes = ([ctypes.c_int] * 4 + [ctypes.c_float] * 3 +
                                  [vp, ctypes.c_int, vp, vp, vp, vp] + [ctypes.c_float] * 2 +
                                  [ctypes.c_int] * 4)
    L.orc_shears_rotate.argtypes = [ctypes.c_float, vp, vp, ctypes.c_int, ctypes.c_int]
    L.orc_shears_extents.argtypes = [ctypes.c_int, vp, vp]
    L.orc_subsets.argtypes = [ctypes.c_int, ctypes.c_int, vp]
    L.orc_osem.argtypes = ([ctypes.c_int] * 8 + [ctypes.c_float] + [vp, vp, ctypes.c_int, vp] +
                           [ctypes.c_int] * 3 + [vp] * 4 + [vp, vp, ctypes.c_int, vp])
    L.orc_esse_init.restype = vp
    L.orc_esse_init.argtypes = ([vp, vp, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 5 +
                                [ctypes.c_int] * 4 + [ctypes.c_float] * 2)
    L.orc_esse_free.argtypes = [vp]
    L.orc_esse_effsrc.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    L.orc_reorder.argtypes = [ctypes.c_int] * 3 + [vp, vp]
    L.orc_compute_orbit.argtypes = ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_float, vp, ctypes.c_float, ctypes.c_int, vp, vp])
    L.orc_compute_orbit.restype = ctypes.c_int
    L.orc_real_size.restype = ctypes.c_int
    L.orc_measured_to_one.argtypes = [vp, vp] + [ctypes.c_int] * 3 + [ctypes.c_float] * 3
    L.orc_nonmeasured_to_zero.argtypes = [vp, vp] + [ctypes.c_int] * 3 + [ctypes.c_float] * 2
    L.orc_initial_estimate.argtypes = ([ctypes.c_int] * 4 + [ctypes.c_float, ctypes.c_int,
                                        ctypes.c_float] + [vp] * 6)
    L.orc_initial_estimate.restype = ctypes.c_int
    L.orc_esse_effsrc_cols.argtypes = [vp, vp, vp] + [ctypes.c_int] * 4 + [vp, vp]
    L.orc_esse_effsrc_cols.restype = ctypes.c_int
    L.orc_esse_info.argtypes = [vp, vp, vp]
    L.orc_esse_grid.argtypes = [vp, ctypes.c_int, vp]
    _libs[real] = L
    return L


def _dt(real):
    return np.float32 if real == "f32" else np.float64


def _arr(a, real):
    return np.ascontiguousarray(a, dtype=_dt(real))


def _p(a):
    return None if a is None else a.ctypes.data


class DrfTable:
    """DRF table in the reference layout: offs[(MaxOffsX, MaxOffsY)], psfs
    of shape (2*MaxOffsY-1, 2*MaxOffsX-1)."""

    def __init__(self, start, spacing, offs, psfs, fft=False):
        self.start, self.spacing, self.fft = float(start), float(spacing), bool(fft)
        self.offs = np.ascontiguousarray(np.array(offs, dtype=np.int32).reshape(-1))
        self.psf = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float32).reshape(-1)
                                                        for p in psfs]))
        self.n = len(psfs)
        self._h = {}

    def handle(self, real):
        if real not in self._h:
            self._h[real] = lib(real).orc_drftab_new(self.n, self.start, self.spacing,
                                                     self.offs.ctypes.data, self.psf.ctypes.data)
        return self._h[real]

    def __del__(self):
        for real, h in self._h.items():
            try:
                lib(real).orc_drftab_free(h)
            except Exception:
                pass


def rot3dpar(angle, vol, pb_flag, nz, n, nrot, real="f64"):
    vol = _arr(vol, real)
    out = np.zeros((nrot, nz, nrot), _dt(real))
    lib(real).orc_rot3dpar(angle, vol.ctypes.data, out.ctypes.data, pb_flag, nz, n, nrot)
    return out


def shears_rotate(angle, vol, nz, n, real="f64"):
    """vShearsRotate of a [y][z][x] volume (n x nz x n)."""
    vol = _arr(vol, real)
    out = np.zeros((n, nz, n), _dt(real))
    lib(real).orc_shears_rotate(angle, vol.ctypes.data, out.ctypes.data, nz, n)
    return out


def shears_extents(n):
    s = np.zeros(n, np.int32)
    e = np.zeros(n, np.int32)
    lib("f64").orc_shears_extents(n, s.ctypes.data, e.ctypes.data)
    return s, e


def conv2d_ypad(plane, psf, pad_y=0, avg_y=0, real="f64"):
    plane = _arr(plane, real)
    psf = np.ascontiguousarray(psf, np.float32)
    ny, nx = plane.shape
    out = np.zeros_like(plane)
    lib(real).orc_conv2d_ypad(plane.ctypes.data, nx, ny, psf.ctypes.data,
                              (psf.shape[1] - 1) // 2, (psf.shape[0] - 1) // 2, pad_y, avg_y,
                              out.ctypes.data)
    return out


def init_prj(nz, n, nrot, model=7, real="f64"):
    lib(real).orc_init_prj(nz, n, nrot, model)


def project(nz, n, nrot, model, pw, angle, cfcr, drf, atn, recon, primary=1.0, atnfac=1.0,
            pad=0, avg=0, atn_rotated=0, act_rotated=0, real="f64", esse=None, fast=0):
    L = lib(real)
    atn = None if atn is None else _arr(atn, real)
    recon = _arr(recon, real)
    out = np.zeros((nz, nrot), _dt(real))
    L.orc_project(nz, n, nrot, model, pw, angle, cfcr, drf.handle(real) if drf else None,
                  int(drf.fft) if drf else 0, esse, _p(atn), recon.ctypes.data, out.ctypes.data,
                  primary, atnfac, pad, avg, atn_rotated, act_rotated, int(fast))
    return out


def backproject(nz, n, nrot, model, pw, angle, cfcr, drf, atn, ratio, image, primary=1.0,
                atnfac=1.0, pad=0, avg=0, atn_rotated=0, real="f64", esse=None, fast=0):
    """Adds the backprojection of 'ratio' into 'image' (in place, with the EPS
    clamp); 'image' must be a contiguous array of the oracle's dtype."""
    L = lib(real)
    atn = None if atn is None else _arr(atn, real)
    ratio = _arr(ratio, real)
    assert image.dtype == _dt(real) and image.flags.c_contiguous
    L.orc_backproject(nz, n, nrot, model, pw, angle, cfcr, drf.handle(real) if drf else None,
                      int(drf.fft) if drf else 0, esse, _p(atn), ratio.ctypes.data,
                      image.ctypes.data, primary, atnfac, pad, avg, atn_rotated, int(fast))
    return image


def subsets(nviews, per):
    L = lib("f64")
    s = L.orc_subsets(nviews, per, None)
    if s < 0:
        return s
    out = np.zeros(s * per, np.int32)
    L.orc_subsets(nviews, per, out.ctypes.data)
    return out.reshape(s, per)


def compute_orbit(img, pix, margin, angles, thresh, otype, margins=None, real="f32"):
    """ComputeOrbit (computeorbit.c:21-150); img (N x N float32) is
    thresholded in place.  Returns (rc, cfcr array)."""
    assert img.dtype == np.float32 and img.flags.c_contiguous
    angles = np.ascontiguousarray(angles, np.float32)
    out = np.zeros(len(angles), np.float32)
    m = None if margins is None else np.ascontiguousarray(margins, np.float32)
    rc = lib(real).orc_compute_orbit(img.ctypes.data, img.shape[0], len(angles), pix, margin,
                                     _p(m), thresh, otype, angles.ctypes.data, out.ctypes.data)
    return rc, out


def reorder(d0, d1, d2, a, real="f64"):
    a = _arr(a, real)
    out = np.empty((d1, d0, d2), _dt(real))
    lib(real).orc_reorder(d0, d1, d2, a.ctypes.data, out.ctypes.data)
    return out


def project_srf_columns(n, nz, angle, esse, atn, recon, cols, primary=1.0, real="f64", fast=0):
    """Project (prjbck.c:68-252) with attenuation and ESSE scatter, no DRF,
    of one view, on the detector columns cols = [(slice, bin), ...] only:
    rotation by 0.5 pi + angle (rot3dpar, or with 'fast' the three-shear
    rotation vShearsRotate), the effective source on those
    depth columns (Esse.effsrc_columns; esse None: attenuation only) and the
    half-pixel attenuated depth sum.  atn (per pixel) and recon are internal [y][z][x] volumes."""
    ang = np.float32(0.5 * np.pi + np.float64(np.float32(angle)))
    if fast:
        ra = shears_rotate(float(ang), atn, nz, n, real=real)
        rx = shears_rotate(float(ang), recon, nz, n, real=real)
    else:
        ra = rot3dpar(float(ang), atn, 0, nz, n, n, real=real)
        rx = rot3dpar(float(ang), recon, 0, nz, n, n, real=real)
    c = np.asarray(cols, np.int64).reshape(-1, 2)
    eff = 0.0 if esse is None else esse.effsrc_columns(rx, ra, c)   # [col][row]
    mu = ra[:, c[:, 0], c[:, 1]].T                           # [col][row]
    src = rx[:, c[:, 0], c[:, 1]].T * primary + eff
    S = np.cumsum(mu, axis=1) - 0.5 * mu                     # half-pixel running sum
    return (src * np.exp(-S)).sum(axis=1)


def measured_to_one(meas, left, right, pwid=1.0, nbins=None, real="f64"):
    """SetMeasuredPrjBinsToOne (osemrecon.c:92-138) of one view's [Nz][NRot]
    measured projection: the norm image's bin mask."""
    meas = _arr(meas, real)
    nz, nrot = meas.shape
    out = np.empty_like(meas)
    lib(real).orc_measured_to_one(out.ctypes.data, meas.ctypes.data, nrot,
                                  nrot if nbins is None else nbins, nz, pwid, left, right)
    return out


def nonmeasured_to_zero(prj, meas, left, pwid=1.0, nbins=None, real="f64"):
    """SetNonMeasuredBinsToZero (osemrecon.c:140-184): a copy of prj with the
    bins outside the measured range, or whose measured value is negative,
    set to zero."""
    out = _arr(prj, real).copy()
    meas = _arr(meas, real)
    nz, nrot = out.shape
    lib(real).orc_nonmeasured_to_zero(out.ctypes.data, meas.ctypes.data, nrot,
                                      nrot if nbins is None else nbins, nz, pwid, left)
    return out


def initial_estimate(n, nz, pw, views, prj, atn=None, contour=False, thresh=0.0):
    """pfComputeInitialEstimate (initial_est.c:29-180) with IrlOsem's inputs:
    prj [V][Nz][N] float32, atn the reordered but unscaled map (internal
    [y][z][x]) or None, views with .Angle and .CFCR.  Returns (internal-order
    estimate, warned)."""
    prj = np.ascontiguousarray(prj, np.float32)
    V = prj.shape[0]
    ang = np.array([v.Angle for v in views], np.float64)
    sn = np.sin(ang).astype(np.float32)
    cs = np.cos(ang).astype(np.float32)
    cf = np.array([v.CFCR for v in views], np.float32)
    a = None if atn is None else np.ascontiguousarray(atn, np.float32)
    out = np.empty((n, nz, n), np.float32)
    w = lib("f32").orc_initial_estimate(n, nz, prj.shape[2], V, pw, int(contour), thresh,
                                        sn.ctypes.data, cs.ctypes.data, cf.ctypes.data,
                                        prj.ctypes.data, _p(a), out.ctypes.data)
    return out, bool(w)


class Esse:
    """ESSE model of the oracle (orc_esse_init); None handle when scatFac is 0."""

    def __init__(self, krnl, krnl_mu, mu0, pixsize, n, nz, collapse=1, nterms=2, scatfac=1.0,
                 cntrpixfac=0.0, real="f64", xypad=2.0, zpad=2.0):
        self.real = real
        self.krnl = np.ascontiguousarray(krnl, np.float32)
        self.krnl_mu = np.ascontiguousarray(krnl_mu, np.float32)
        nd, no, _ = self.krnl.shape
        self.h = lib(real).orc_esse_init(self.krnl.ctypes.data, self.krnl_mu.ctypes.data, no, nd,
                                         mu0, scatfac, 1.0, cntrpixfac, pixsize, n, nz, collapse,
                                         nterms, xypad, zpad)

    def effsrc(self, act, mp):
        act = _arr(act, self.real)
        mp = _arr(mp, self.real)
        ny, nz, nx = act.shape
        out = np.zeros_like(act)
        lib(self.real).orc_esse_effsrc(self.h, act.ctypes.data, mp.ctypes.data, nx, ny, nz,
                                       out.ctypes.data)
        return out

    def info(self):
        """(grid dims (Ky, Kz, Kx), dict of the model's constants)."""
        d = np.zeros(3, np.int32)
        q = np.zeros(7, np.float64)
        lib(self.real).orc_esse_info(self.h, d.ctypes.data, q.ctypes.data)
        return tuple(int(v) for v in d), dict(mu0=q[0], muavg=q[1], pixsize=q[2], cntr=q[3],
                                               nterms=int(q[4]), collapse=int(q[5]),
                                               scatfac=q[6])

    def grid(self, k):
        """Dense [Ky][Kz][Kx] kernel term k (0: K, 1: K(mu - muavg), 2: K(mu - muavg)^2)."""
        dims, _ = self.info()
        g = np.zeros(dims, np.float32)
        lib(self.real).orc_esse_grid(self.h, k, g.ctypes.data)
        return g

    def effsrc_columns(self, act, mp, cols):
        """The effective source on the depth columns cols = [(slice, bin), ...]
        of [depth][slice][bin] volumes: [len(cols)][depth]."""
        act = _arr(act, self.real)
        mp = _arr(mp, self.real)
        ny, nz, nx = act.shape
        c = np.ascontiguousarray(cols, np.int32).reshape(-1, 2)
        out = np.zeros((len(c), ny), _dt(self.real))
        rc = lib(self.real).orc_esse_effsrc_cols(self.h, act.ctypes.data, mp.ctypes.data, nx, ny,
                                                 nz, len(c), c.ctypes.data, out.ctypes.data)
        if rc:
            raise ValueError("column evaluation needs an uncollapsed model")
        return out

    def __del__(self):
        try:
            lib(self.real).orc_esse_free(self.h)
        except Exception:
            pass


ITER_CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)


def osem(nz, n, views, per_subset, iterations, prj_model, bck_model, pw, drf, prj, atn, x0,
         scat=None, start_iteration=0, pad=0, avg=0, real="f64", iter_cb=None, esse=None,
         srf_iterations=0, fast=0, bck_drf=None):
    """vOsemRecon on internal-order inputs (atn already scaled to 1/pixel and
    clamped, prj clamped, x0 internal order).  views: list of dicts with
    angle, left, right, cfcr, viewtime, sensitivity.  Returns the estimate."""
    L = lib(real)
    nv = len(views)
    vv = (OrcView * nv)(*[OrcView(v["angle"], v["left"], v["right"], v["cfcr"],
                                  v.get("viewtime", 1.0), v.get("sensitivity", 1.0))
                          for v in views])
    x = _arr(x0, real).copy()
    prj = _arr(prj, real)
    atn = None if atn is None else _arr(atn, real)
    scat = None if scat is None else _arr(scat, real)
    M = x.size

    def _cb(it, ptr, ctx):
        if iter_cb is not None:
            buf = (ctypes.c_char * (M * x.itemsize)).from_address(ptr)
            iter_cb(it, np.frombuffer(buf, dtype=_dt(real)).copy())
    cb = ITER_CB(_cb)
    rc = L.orc_osem(nz, n, nv, per_subset, iterations, start_iteration, prj_model, bck_model,
                    pw, ctypes.addressof(vv), drf.handle(real) if drf else None,
                    int(drf.fft) if drf else 0, esse, srf_iterations, pad, avg,
                    prj.ctypes.data, _p(scat), _p(atn), x.ctypes.data,
                    ctypes.cast(cb, ctypes.c_void_p), None, int(fast),
                    bck_drf.handle(real) if bck_drf else None)
    if rc != 0:
        raise ValueError("orc_osem failed with %d" % rc)
    return x
