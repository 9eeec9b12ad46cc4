"""libirl_amd -- MI355X-native OSEM projector/backprojector with libirl's C API.

This module is the Python mirror of the C interface in include/: it loads the
in-tree libirl_amd.so (built for gfx950 by libirl_amd/build.py) and exposes
IrlOsem, the hook/option/detector calls and the device-resident
reconstruction API with numpy arrays.  Arrays are passed by pointer, so the
in-place side effects of IrlOsem on the caller's arrays (reordered and scaled
attenuation map, clamped projections, result in pfReconImage) are visible
exactly as for a C caller.

There is no fallback: if the shared library is missing, importing the
bindings raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libirl_amd.so")

MODEL_ATN = 1
MODEL_DRF = 2
MODEL_SRF = 4


class IrlParms(ctypes.Structure):
    """IrlParms_t (irl.h:39-79)."""
    _fields_ = [
        ("NumPixels", ctypes.c_int),
        ("NumSlices", ctypes.c_int),
        ("NumViews", ctypes.c_int),
        ("BinWidth", ctypes.c_float),
        ("fAtnScaleFac", ctypes.c_float),
        ("iNumSrfIterations", ctypes.c_int),
        ("fScatEstFac", ctypes.c_float),
        ("NumAngPerSubset", ctypes.c_int),
        ("NumIterations", ctypes.c_int),
        ("pchNormImageBase", ctypes.c_char_p),
        ("SrfCollapseFac", ctypes.c_int),
        ("fHoleDiam", ctypes.c_float),
        ("fHoleLen", ctypes.c_float),
        ("fBackToDet", ctypes.c_float),
        ("fIntrinsicFWHM", ctypes.c_float),
    ]


class Options(ctypes.Structure):
    """Options_t (irl.h:81-124)."""
    _fields_ = [
        ("bModelDrf", ctypes.c_int),
        ("bReconIsInitEst", ctypes.c_int),
        ("bUseGrfInBck", ctypes.c_int),
        ("bFFTConvolve", ctypes.c_int),
        ("bUseContourSupport", ctypes.c_int),
        ("fAtnMapThresh", ctypes.c_float),
        ("iMsgLevel", ctypes.c_int),
        ("iAxialPadLength", ctypes.c_int),
        ("iAxialAvgLength", ctypes.c_int),
    ]


class PrjView(ctypes.Structure):
    """PrjView_t (irl.h:127-143)."""
    _fields_ = [
        ("Angle", ctypes.c_float),
        ("Left", ctypes.c_float),
        ("Right", ctypes.c_float),
        ("CFCR", ctypes.c_float),
    ]


class OsemData(ctypes.Structure):
    """sOsemData (osemhooks.h:18-27)."""
    _fields_ = [
        ("iIterationIdx", ctypes.c_int),
        ("iSubsetIdx", ctypes.c_int),
        ("iAngleIdx", ctypes.c_int),
        ("pfEst", ctypes.POINTER(ctypes.c_float)),
        ("pfSct", ctypes.POINTER(ctypes.c_float)),
        ("pfNorm", ctypes.POINTER(ctypes.c_float)),
        ("pfUpdate", ctypes.POINTER(ctypes.c_float)),
    ]


class AdvOptions(ctypes.Structure):
    """sAdvOptions (osemhooks.h:29-33)."""
    _fields_ = [("bFastRotate", ctypes.c_int), ("iStartIteration", ctypes.c_int)]


class Drf(ctypes.Structure):
    """Drf_t (drftabio.h:7-12)."""
    _fields_ = [("MaxOffsX", ctypes.c_int), ("MaxOffsY", ctypes.c_int),
                ("bFFTdone", ctypes.c_int), ("Psf", ctypes.POINTER(ctypes.c_float))]


class DrfTab(ctypes.Structure):
    """DrfTab_t (drftabio.h:17-28)."""


DrfTab._fields_ = [
    ("iNumDistances", ctypes.c_int), ("fStartDistance", ctypes.c_float),
    ("fDrfSpacing", ctypes.c_float), ("fPixSize", ctypes.c_float), ("bIncKrnl", ctypes.c_int),
    ("bDrfFromFile", ctypes.c_int), ("psBckDrfTab", ctypes.POINTER(DrfTab)),
    ("fGap", ctypes.c_float), ("fHoleDiam", ctypes.c_float), ("fColThick", ctypes.c_float),
    ("fIntrinsicFWHM", ctypes.c_float), ("fMaxFracErr", ctypes.c_float),
    ("psDrfs", ctypes.POINTER(Drf)), ("bFFTconvolve", ctypes.c_int)]


def make_drftab(start, spacing, offs, psfs, fft=False, pixsize=None):
    """DrfTab_t holding the given PSFs (kept alive by the returned object)."""
    t = DrfTab()
    keep = [np.ascontiguousarray(p, np.float32) for p in psfs]
    drfs = (Drf * len(psfs))()
    for i, ((ox, oy), p) in enumerate(zip(offs, keep)):
        assert p.shape == (2 * oy - 1, 2 * ox - 1)
        drfs[i].MaxOffsX, drfs[i].MaxOffsY, drfs[i].bFFTdone = ox, oy, 0
        drfs[i].Psf = p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    t.iNumDistances = len(psfs)
    t.fStartDistance = start
    t.fDrfSpacing = spacing
    t.fPixSize = spacing if pixsize is None else pixsize
    t.psDrfs = drfs
    t.bFFTconvolve = 1 if fft else 0
    t._keep = (keep, drfs)
    return t


ITER_CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.POINTER(ctypes.c_float))
HOOK_CB = ctypes.CFUNCTYPE(None, ctypes.POINTER(OsemData), ctypes.c_void_p)
_FP = ctypes.POINTER(ctypes.c_float)

_lib = None


def load():
    """Loads libirl_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libirl_amd.so is not built (run libirl_amd/build.py): " + LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.IrlOsem.restype = ctypes.c_int
    lib.IrlOsem.argtypes = [ctypes.POINTER(IrlParms), ctypes.POINTER(Options),
                            ctypes.POINTER(PrjView), ctypes.c_char_p, ctypes.c_char_p,
                            ITER_CB, _FP, _FP, _FP, _FP, ctypes.c_char_p, ctypes.c_char_p]
    lib.IrlGenprj.restype = ctypes.c_int
    lib.IrlGenprj.argtypes = [ctypes.POINTER(IrlParms), ctypes.POINTER(Options),
                              ctypes.POINTER(PrjView), ctypes.c_char_p, ctypes.c_char_p,
                              ITER_CB, _FP, _FP, _FP, _FP, ctypes.c_float, ctypes.c_char_p,
                              ctypes.c_char_p]
    lib.pchIrlErrorString.restype = ctypes.c_char_p
    lib.OsemSetAdvOptions.argtypes = [ctypes.POINTER(AdvOptions)]
    for name in ("PrevIter", "PostIter", "PrevSubset", "PostSubset", "PrevAngle", "PostAngle"):
        f = getattr(lib, "OsemSet%sHook" % name)
        f.argtypes = [HOOK_CB, ctypes.c_void_p]
        f.restype = ctypes.c_int
    lib.iSetupExtraPrjViewData.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.iSetupDetectors.argtypes = [ctypes.c_int, ctypes.c_void_p]
    lib.pvStoreDetectorData.restype = ctypes.c_void_p
    lib.pvStoreDetectorData.argtypes = [ctypes.c_float] * 5 + [ctypes.c_int, ctypes.c_int,
                                                               ctypes.c_void_p]
    lib.irlamd_recon_create.restype = ctypes.c_void_p
    lib.irlamd_recon_create.argtypes = [ctypes.POINTER(IrlParms), ctypes.POINTER(Options),
                                        ctypes.POINTER(PrjView), ctypes.c_char_p,
                                        ctypes.c_char_p, _FP, _FP, _FP, _FP]
    lib.irlamd_recon_subsets.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]
    lib.irlamd_recon_subiteration.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.irlamd_recon_synchronize.argtypes = [ctypes.c_void_p]
    lib.irlamd_recon_prepare_graphs.argtypes = [ctypes.c_void_p]
    lib.irlamd_recon_get_estimate.argtypes = [ctypes.c_void_p, _FP]
    lib.irlamd_recon_get_norm.argtypes = [ctypes.c_void_p, ctypes.c_int, _FP]
    lib.irlamd_recon_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.irlamd_recon_reset_timing.argtypes = [ctypes.c_void_p]
    lib.irlamd_recon_stage_count.argtypes = [ctypes.c_void_p]
    lib.irlamd_recon_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_int)]
    lib.irlamd_recon_destroy.argtypes = [ctypes.c_void_p]
    lib.InitPrj.argtypes = [ctypes.c_int] * 4
    vp = ctypes.c_void_p
    lib.Project.argtypes = ([ctypes.c_int] * 4 + [ctypes.c_float, ctypes.c_int, ctypes.c_float,
                                                 ctypes.c_float, vp, vp, _FP, _FP, _FP,
                                                 ctypes.c_float, ctypes.c_float] +
                            [ctypes.c_int] * 5)
    lib.BackProject.argtypes = ([ctypes.c_int] * 4 + [ctypes.c_float, ctypes.c_int,
                                                     ctypes.c_float, ctypes.c_float, vp, vp, _FP,
                                                     _FP, _FP, ctypes.c_float, ctypes.c_float] +
                                [ctypes.c_int] * 4)
    lib.rot3dpar.argtypes = [ctypes.c_float, _FP, _FP] + [ctypes.c_int] * 4
    lib.vShearsRotate.argtypes = [ctypes.c_float, _FP, _FP] + [ctypes.c_int] * 3
    lib.ConvolveDrfYpad.argtypes = [_FP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Drf)] + \
        [ctypes.c_int] * 3 + [_FP]
    lib.irlamd_last_error.restype = ctypes.c_int
    lib.gendrf.argtypes = ([ctypes.c_int] + [ctypes.c_float] * 9 + [ctypes.c_int] * 2 +
                           [ctypes.c_void_p])
    lib.ComputeOrbit.argtypes = [_FP, ctypes.POINTER(IrlParms), ctypes.c_float, ctypes.c_void_p,
                                 ctypes.c_float, ctypes.c_int, ctypes.c_void_p]
    lib.ComputeOrbit.restype = ctypes.c_int
    lib.irlamd_set_precision.argtypes = [ctypes.c_int]
    lib.irlamd_set_precision.restype = ctypes.c_int
    lib.irlamd_recon_bf16_planes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _FP,
                                             ctypes.c_int]
    vp = ctypes.c_void_p
    lib.irlamd_cnn_train.restype = vp
    lib.irlamd_cnn_train.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_uint, _FP]
    lib.irlamd_cnn_create.restype = vp
    lib.irlamd_cnn_create.argtypes = [ctypes.c_float, ctypes.c_uint]
    lib.irlamd_cnn_save.argtypes = [vp, ctypes.c_char_p]
    lib.irlamd_cnn_load.restype = vp
    lib.irlamd_cnn_load.argtypes = [ctypes.c_char_p]
    lib.irlamd_cnn_free.argtypes = [vp]
    lib.irlamd_cnn_free.restype = None
    lib.irlamd_cnn_get_params.argtypes = [vp, _FP, _FP]
    lib.irlamd_cnn_set_params.argtypes = [vp, _FP]
    lib.irlamd_cnn_forward.argtypes = [vp, _FP] + [ctypes.c_int] * 4 + [_FP]
    lib.irlamd_cnn_set_row_length.argtypes = [vp, ctypes.c_int]
    lib.irlamd_cnn_loss_grad.argtypes = [vp, _FP, _FP] + [ctypes.c_int] * 3 + \
        [ctypes.POINTER(ctypes.c_double), _FP]
    lib.irlamd_cnn_sources.argtypes = [vp, _FP, _FP, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, _FP, _FP]
    lib.irlamd_set_scatter_cnn.argtypes = [vp, ctypes.c_int, ctypes.c_float]
    lib.irlamd_set_scatter_cnn_exact.argtypes = [ctypes.c_int]
    lib.irlamd_recon_cnn_status.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 3 + \
        [ctypes.POINTER(ctypes.c_float)] * 4
    lib.irlamd_set_graphs.argtypes = [ctypes.c_int]
    lib.irlamd_recon_cnn_row_length.argtypes = [vp]
    lib.irlamd_recon_prj_rows_per_chunk.argtypes = [vp]
    lib.irlamd_set_autotune.argtypes = [ctypes.c_int]
    lib.irlamd_recon_tiles.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.irlamd_recon_set_tiles.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.irlamd_set_drf_fft.argtypes = [ctypes.c_int]
    lib.irlamd_set_esse_fused.argtypes = [ctypes.c_int]
    lib.irlamd_recon_drf_algorithms.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 3
    lib.irlamd_batch_create.restype = vp
    lib.irlamd_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    lib.irlamd_batch_step.argtypes = [vp]
    lib.irlamd_batch_synchronize.argtypes = [vp]
    lib.irlamd_batch_stats.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 3
    lib.irlamd_batch_set_parent_graph.argtypes = [vp, ctypes.c_int]
    lib.irlamd_batch_destroy.argtypes = [vp]
    lib.irlamd_batch_destroy.restype = None
    _lib = lib
    return lib


def fptr(a):
    """float* of a C-contiguous float32 array, or NULL for None."""
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float32 array")
    return a.ctypes.data_as(_FP)


def error_string():
    return load().pchIrlErrorString().decode(errors="replace")


def _bytes(s):
    return None if s is None else os.fsencode(s)


def irl_osem(parms, options, views, drf_file, krnl_file, scat, atn, prj, recon,
             iter_cb=None, log_file=None, msg_file=None):
    """IrlOsem (irl.h:145-213).  numpy arrays are modified in place exactly as
    the C function modifies the caller's arrays.  iter_cb(it, est) receives a
    copy of the estimate in external order.  Returns the C return code."""
    lib = load()
    keep = []

    def _cb(it, p):
        if iter_cb is not None:
            n = parms.NumPixels * parms.NumPixels * parms.NumSlices
            iter_cb(it, np.ctypeslib.as_array(p, shape=(n,)).copy())
    cb = ITER_CB(_cb)
    keep.append(cb)
    vv = (PrjView * len(views))(*views)
    return lib.IrlOsem(ctypes.byref(parms), ctypes.byref(options), vv, _bytes(drf_file),
                       _bytes(krnl_file), cb, fptr(scat), fptr(atn), fptr(prj), fptr(recon),
                       _bytes(log_file), _bytes(msg_file))


def irl_genprj(parms, options, views, drf_file, krnl_file, scat, act, atn, prj,
               primary=1.0, view_cb=None, log_file=None, msg_file=None):
    """IrlGenprj (irl.h:215-288).  Writes the projections into 'prj'
    ([view][slice][bin] float32); act/atn/scat are modified in place as the C
    function modifies them."""
    lib = load()

    def _cb(v, p):
        if view_cb is not None:
            n = parms.NumPixels * parms.NumSlices
            view_cb(v, np.ctypeslib.as_array(p, shape=(n,)).copy())
    cb = ITER_CB(_cb)
    vv = (PrjView * len(views))(*views)
    return lib.IrlGenprj(ctypes.byref(parms), ctypes.byref(options), vv, _bytes(drf_file),
                         _bytes(krnl_file), cb, fptr(scat), fptr(act), fptr(atn), fptr(prj),
                         primary, _bytes(log_file), _bytes(msg_file))


def _tab(t):
    return None if t is None else ctypes.cast(ctypes.pointer(t), ctypes.c_void_p)


def init_prj(nz, n, nrot, model):
    load().InitPrj(nz, n, nrot, model)


def free_prj():
    load().FreePrj()


def project(nz, n, nrot, model, pw, angle, cfcr, drftab, atn, recon, primary=1.0, atnfac=1.0,
            pad=0, avg=0, atn_rotated=0, act_rotated=0, fast=0):
    """Project (prjbck.c:68-252) of internal-order [y][z][x] float32 volumes;
    returns the [slice][bin] projection.  fast selects the shear rotation."""
    out = np.zeros((nz, nrot), np.float32)
    load().Project(nz, n, nrot, model, pw, 0, angle, cfcr, _tab(drftab), None, fptr(atn),
                   fptr(recon), fptr(out), primary, atnfac, pad, avg, atn_rotated,
                   act_rotated, int(fast))
    return out


def backproject(nz, n, nrot, model, pw, angle, cfcr, drftab, atn, ratio, image, primary=1.0,
                atnfac=1.0, pad=0, avg=0, atn_rotated=0, fast=0):
    """BackProject (prjbck.c:254-427): adds into 'image' in place."""
    load().BackProject(nz, n, nrot, model, pw, 0, angle, cfcr, _tab(drftab), None, fptr(atn),
                       fptr(ratio), fptr(image), primary, atnfac, pad, avg, atn_rotated,
                       int(fast))
    return image


def shears_rotate(angle, vol, nz, n):
    """vShearsRotate (fastrotate.c:330-563) of a [y][z][x] float32 volume."""
    out = np.zeros((n, nz, n), np.float32)
    load().vShearsRotate(angle, fptr(np.ascontiguousarray(vol, np.float32)), fptr(out), 0, nz,
                         n)
    return out


def rot3dpar(angle, vol, pb_flag, nz, n, nrot):
    out = np.zeros((nrot, nz, nrot), np.float32)
    load().rot3dpar(angle, fptr(np.ascontiguousarray(vol, np.float32)), fptr(out), pb_flag, nz,
                    n, nrot)
    return out


def convolve_drf_ypad(plane, psf, fft=0, pad=0, avg=0):
    psf = np.ascontiguousarray(psf, np.float32)
    d = Drf((psf.shape[1] + 1) // 2, (psf.shape[0] + 1) // 2, 0,
            psf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    plane = np.ascontiguousarray(plane, np.float32)
    out = np.zeros_like(plane)
    load().ConvolveDrfYpad(fptr(plane), plane.shape[1], plane.shape[0], ctypes.byref(d), fft,
                           pad, avg, fptr(out))
    return out


def last_error():
    return load().irlamd_last_error()


PRECISION = {"fp32": 0, "bf16": 1, "bf16x3": 2, "bf16-plain": 3}


def set_precision(mode):
    """DRF convolution arithmetic of reconstructions created afterwards
    (irlamd_set_precision): 'fp32', 'bf16' (hi+lo split where needed),
    'bf16x3' (split always) or 'bf16-plain'.  Returns the previous mode."""
    m = PRECISION[mode] if isinstance(mode, str) else int(mode)
    old = load().irlamd_set_precision(m)
    if old < 0:
        raise ValueError("invalid precision mode %r" % (mode,))
    return old


def gendrf(col_thick, hole_diam, gap, intrinsic, max_err, bin_size, spacing, min_dist, ndist,
           noffs):
    """Computed parallel-hole GRF table (gendrf, gendrf.c:221-427): a list of
    (2*MaxOffsY-1, 2*MaxOffsX-1) float32 PSFs, one per distance."""
    lib = load()
    tab = (Drf * ndist)()
    lib.gendrf(0, 0.0, col_thick, hole_diam, gap, intrinsic, max_err, bin_size, spacing,
               min_dist, ndist, noffs, ctypes.addressof(tab))
    if ndist and not tab[0].Psf:
        raise RuntimeError("gendrf failed: " + error_string())
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    out = []
    for d in tab:
        h, w = 2 * d.MaxOffsY - 1, 2 * d.MaxOffsX - 1
        out.append(np.ctypeslib.as_array(d.Psf, shape=(h * w,)).reshape(h, w).copy())
        libc.free(ctypes.cast(d.Psf, ctypes.c_void_p))
    return out


def compute_orbit(img, parms, margin, views, thresh, orbit_type, margins=None):
    """ComputeOrbit (irl.h:291-316): fills views[v].CFCR; img (N x N float32,
    contiguous) is thresholded in place.  Returns the C return code."""
    vv = (PrjView * len(views))(*views)
    m = None if margins is None else np.ascontiguousarray(margins, np.float32)
    rc = load().ComputeOrbit(fptr(img), ctypes.byref(parms), margin,
                             None if m is None else m.ctypes.data, thresh, orbit_type,
                             ctypes.addressof(vv))
    for k in range(len(views)):
        views[k].CFCR = vv[k].CFCR
    return rc


def set_adv_options(fast_rotate=1, start_iteration=0):
    o = AdvOptions(fast_rotate, start_iteration)
    return load().OsemSetAdvOptions(ctypes.byref(o))


class Recon:
    """Device-resident reconstruction (include/irl_amd.h)."""

    def __init__(self, parms, options, views, drf_file=None, krnl_file=None, scat=None,
                 atn=None, prj=None, recon=None):
        lib = load()
        vv = (PrjView * len(views))(*views)
        self._lib = lib
        self.parms = parms
        h = lib.irlamd_recon_create(ctypes.byref(parms), ctypes.byref(options), vv,
                                    _bytes(drf_file), _bytes(krnl_file), fptr(scat), fptr(atn),
                                    fptr(prj), fptr(recon))
        if not h:
            raise RuntimeError("irlamd_recon_create failed: " + error_string())
        self._h = h
        s, vs = ctypes.c_int(), ctypes.c_int()
        lib.irlamd_recon_subsets(h, ctypes.byref(s), ctypes.byref(vs))
        self.num_subsets, self.views_per_subset = s.value, vs.value

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(error_string())

    def subiteration(self, s):
        self._check(self._lib.irlamd_recon_subiteration(self._h, s))

    def iteration(self):
        for s in range(self.num_subsets):
            self.subiteration(s)

    def synchronize(self):
        self._check(self._lib.irlamd_recon_synchronize(self._h))

    def prepare_graphs(self):
        """Captures every subset's sub-iteration graph now (graphs on)."""
        self._check(self._lib.irlamd_recon_prepare_graphs(self._h))

    def estimate(self):
        p = self.parms
        out = np.empty((p.NumSlices, p.NumPixels, p.NumPixels), np.float32)
        self._check(self._lib.irlamd_recon_get_estimate(self._h, fptr(out)))
        return out

    def norm(self, s):
        p = self.parms
        out = np.empty((p.NumPixels, p.NumSlices, p.NumPixels), np.float32)
        self._check(self._lib.irlamd_recon_get_norm(self._h, s, fptr(out)))
        return out

    def tiles(self):
        """DRF tile shapes in use (projector, backprojector); 0 = size heuristic."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._lib.irlamd_recon_tiles(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def set_tiles(self, prj, bck):
        """DRF tile shapes (0: size heuristic; BF16 2/4/8, FP32 4/8)."""
        self._check(self._lib.irlamd_recon_set_tiles(self._h, prj, bck))

    def cnn_row_length(self):
        """Slices per wave of the scatter network's MFMA layers in use (0: none)."""
        return self._lib.irlamd_recon_cnn_row_length(self._h)

    def prj_rows_per_chunk(self):
        """Rows per partial plane of the spatial projector in use."""
        return self._lib.irlamd_recon_prj_rows_per_chunk(self._h)

    def drf_algorithms(self):
        """FP32 DRF algorithm per direction ("spatial" or "fft") and the FFT length."""
        a, b, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._lib.irlamd_recon_drf_algorithms(self._h, ctypes.byref(a), ctypes.byref(b),
                                              ctypes.byref(n))
        name = ("spatial", "fft")
        return name[a.value], name[b.value], n.value

    def bf16_planes(self):
        """(modes, errors) of the BF16 DRF per table entry (delta entry last):
        modes 0 plain, 1 weights hi+lo, 2 weights and data hi+lo; errors
        [entry][plain, weight split] measured on a probe plane."""
        n = self._lib.irlamd_recon_bf16_planes(self._h, None, None, 0)
        m = np.zeros(max(n, 0), np.int32)
        e = np.zeros((max(n, 0), 2), np.float32)
        if n > 0:
            self._lib.irlamd_recon_bf16_planes(self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                               fptr(e), n)
        return m, e

    def set_timing(self, on=True):
        self._lib.irlamd_recon_set_timing(self._h, 1 if on else 0)

    def reset_timing(self):
        self._check(self._lib.irlamd_recon_reset_timing(self._h))

    def stage_times(self):
        """{kernel: (total_ms, launches)} from HIP events on the recon's stream."""
        n = self._lib.irlamd_recon_stage_count(self._h)
        out = {}
        for i in range(n):
            name = ctypes.create_string_buffer(64)
            ms = ctypes.c_double()
            cnt = ctypes.c_int()
            self._check(self._lib.irlamd_recon_stage(self._h, i, name, 64, ctypes.byref(ms),
                                                     ctypes.byref(cnt)))
            out[name.value.decode()] = (ms.value, cnt.value)
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.irlamd_recon_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ScatterCnn:
    """Handle of the scatter CNN (irlamd_cnn_*): create(), train(), load()."""

    def __init__(self, h):
        if not h:
            raise RuntimeError(error_string())
        self._lib = load()
        self._h = h

    @classmethod
    def create(cls, kappa=1.0, seed=0):
        return cls(load().irlamd_cnn_create(kappa, seed))

    @classmethod
    def train(cls, krnl_file, n, nz, pixsize, pairs, val, epochs, batch=4, lr=2e-3, seed=1):
        """Returns (network, stats) with stats = {loss, val_nrmse0, val_nrmse1, seconds, kappa}."""
        st = np.zeros(5, np.float32)
        h = load().irlamd_cnn_train(_bytes(krnl_file), n, nz, pixsize, pairs, val, epochs, batch,
                                    lr, seed, fptr(st))
        keys = ("loss", "val_nrmse0", "val_nrmse1", "seconds", "kappa")
        return cls(h), dict(zip(keys, map(float, st)))

    @classmethod
    def load_file(cls, fname):
        return cls(load().irlamd_cnn_load(_bytes(fname)))

    def save(self, fname):
        if self._lib.irlamd_cnn_save(self._h, _bytes(fname)):
            raise RuntimeError(error_string())

    @staticmethod
    def num_params():
        return load().irlamd_cnn_num_params()

    def params(self):
        p = np.zeros(self.num_params(), np.float32)
        k = np.zeros(1, np.float32)
        if self._lib.irlamd_cnn_get_params(self._h, fptr(p), fptr(k)):
            raise RuntimeError(error_string())
        return p, float(k[0])

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if self._lib.irlamd_cnn_set_params(self._h, fptr(p)):
            raise RuntimeError(error_string())

    def set_row_length(self, rowz):
        """Slices per wave of the inference path's MFMA layers (64, 32, 16)."""
        if self._lib.irlamd_cnn_set_row_length(self._h, rowz):
            raise RuntimeError(error_string())

    def forward(self, inp, bf16=False):
        inp = np.ascontiguousarray(inp, np.float32)
        Y, X, Z, _ = inp.shape
        out = np.zeros((Y, X, Z, 2), np.float32)
        if self._lib.irlamd_cnn_forward(self._h, fptr(inp), Y, X, Z, 1 if bf16 else 0, fptr(out)):
            raise RuntimeError(error_string())
        return out

    def loss_grad(self, inp, tgt):
        inp = np.ascontiguousarray(inp, np.float32)
        tgt = np.ascontiguousarray(tgt, np.float32)
        Y, X, Z, _ = inp.shape
        g = np.zeros(self.num_params(), np.float32)
        loss = ctypes.c_double()
        if self._lib.irlamd_cnn_loss_grad(self._h, fptr(inp), fptr(tgt), Y, X, Z,
                                          ctypes.byref(loss), fptr(g)):
            raise RuntimeError(error_string())
        return loss.value, g

    def sources(self, x, mu, dens_fac, cntr=0.0, primary=1.0):
        """g0, g1 of a volume pair in [y][x][z] (slices fastest) order."""
        x = np.ascontiguousarray(x, np.float32)
        mu = np.ascontiguousarray(mu, np.float32)
        N, _, Nz = x.shape
        g0 = np.zeros_like(x)
        g1 = np.zeros_like(x)
        if self._lib.irlamd_cnn_sources(self._h, fptr(x), fptr(mu), N, Nz, dens_fac, cntr, primary,
                                        fptr(g0), fptr(g1)):
            raise RuntimeError(error_string())
        return g0, g1

    def close(self):
        if getattr(self, "_h", None):
            self._lib.irlamd_cnn_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def unpack_cnn_params(p):
    """Packed parameters -> [(W[27][Cout][Cin], b[Cout])] per layer."""
    layers = ((2, 32), (32, 32), (32, 32), (32, 2))
    out, o = [], 0
    for ci, co in layers:
        W = np.asarray(p[o:o + 27 * co * ci]).reshape(27, co, ci)
        o += 27 * co * ci
        out.append((W, np.asarray(p[o:o + co])))
        o += co
    return out


def pack_cnn_params(params):
    return np.concatenate([np.concatenate([np.ravel(W), np.ravel(b)]) for W, b in params]).astype(
        np.float32)


def set_scatter_cnn(net, check_every=0, max_nrmse=0.01):
    """Network used by reconstructions created afterwards (None: ESSE only)."""
    h = net._h if net is not None else None
    if load().irlamd_set_scatter_cnn(h, check_every, max_nrmse):
        raise ValueError("invalid learned-scatter settings")


def set_scatter_cnn_exact(on):
    load().irlamd_set_scatter_cnn_exact(1 if on else 0)


def cnn_status(recon=None):
    """{active, checks, used, last_nrmse, worst_nrmse} of a Recon, or of the last IrlOsem."""
    a, c, u = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    last, worst = ctypes.c_float(), ctypes.c_float()
    lsc, wsc = ctypes.c_float(), ctypes.c_float()
    load().irlamd_recon_cnn_status(recon._h if recon is not None else None, ctypes.byref(a),
                                   ctypes.byref(c), ctypes.byref(u), ctypes.byref(last),
                                   ctypes.byref(worst), ctypes.byref(lsc), ctypes.byref(wsc))
    return {"active": bool(a.value), "checks": c.value, "used": u.value,
            "last_nrmse": last.value, "worst_nrmse": worst.value,
            "last_scatter_nrmse": lsc.value, "worst_scatter_nrmse": wsc.value}


def set_graphs(on):
    """Graph replay of sub-iterations for reconstructions created afterwards."""
    return load().irlamd_set_graphs(1 if on else 0)


def set_autotune(on):
    """DRF tile autotuning at setup for reconstructions created afterwards."""
    return load().irlamd_set_autotune(1 if on else 0)


def set_esse_fused(on):
    """Scatter model's plane transforms and fused source (the default) or the
    line passes and separate source kernel, for reconstructions created
    afterwards.  Returns the previous setting."""
    return bool(load().irlamd_set_esse_fused(1 if on else 0))


_DRF_FFT_MODES = {"reference": 0, "fft": 1, "auto": 2, "spatial": -1}


def set_drf_fft(mode):
    """FP32 DRF convolution algorithm for reconstructions created afterwards:
    "reference" (the FFT kernel for tables flagged bFFTconvolve, the default),
    "fft" (every table), "auto" (the faster per direction, timed at setup) or
    "spatial" (never the FFT kernel).  Returns the previous mode's name."""
    if mode not in _DRF_FFT_MODES:
        raise ValueError("DRF algorithm mode %r" % (mode,))
    old = load().irlamd_set_drf_fft(_DRF_FFT_MODES[mode])
    return {v: k for k, v in _DRF_FFT_MODES.items()}[old]


class Batch:
    """Reconstructions stepped together (irlamd_batch_*): each step runs the
    next sub-iteration of every member, the replayable ones in one graph launch."""

    def __init__(self, recons, parent_graph=2):
        self._lib = load()
        self._recons = list(recons)   # members must outlive the batch
        arr = (ctypes.c_void_p * len(self._recons))(*[r._h for r in self._recons])
        h = self._lib.irlamd_batch_create(arr, len(self._recons))
        if not h:
            raise RuntimeError("irlamd_batch_create failed: " + error_string())
        self._h = h
        self.set_parent_graph(parent_graph)

    def set_parent_graph(self, mode):
        """0: member graphs on their own streams; 1: child graphs of one parent
        graph; 2: one graph with the members' launches as parallel branches."""
        if self._lib.irlamd_batch_set_parent_graph(self._h, int(mode)):
            raise RuntimeError(error_string())

    def step(self):
        if self._lib.irlamd_batch_step(self._h):
            raise RuntimeError(error_string())

    def synchronize(self):
        if self._lib.irlamd_batch_synchronize(self._h):
            raise RuntimeError(error_string())

    def stats(self):
        g, r, e = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._lib.irlamd_batch_stats(self._h, ctypes.byref(g), ctypes.byref(r), ctypes.byref(e))
        return {"graph_launches": g.value, "replays": r.value, "eager_subiterations": e.value}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.irlamd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
