"""Seeded test problems shared by the CPU and GPU tests."""
import os
import tempfile

import numpy as np

from libirl_amd import IrlParms, Options, PrjView, synth
from oracle import oracle as O


class Problem:
    """A small reconstruction problem with its oracle-side equivalents."""

    def __init__(self, n=32, nz=16, nviews=16, per=4, iters=2, pix=0.5, seed=1, half=5,
                 atn=True, drf=True, cfcr=None, counts=2e5, noise=True, orbit="circular",
                 pad=0, avg=0, fft=False, inserts=(2, 3), rmax=None, views=None):
        self.n, self.nz, self.nviews, self.per, self.iters, self.pix = n, nz, nviews, per, iters, pix
        self.pad, self.avg = pad, avg
        act, mu = synth.phantom(n, nz, seed, n_inserts=inserts)
        self.act = act
        self.mu = mu if atn else None
        if cfcr is None:
            cfcr = n * pix / 2 + 2.0
        if views is not None:
            self.views = views
        elif orbit == "circular":
            self.views = synth.views_circular(nviews, cfcr, n)
        else:
            self.views = synth.views_ellipse(nviews, n, seed, cfcr,
                                             cfcr + 4.0 if rmax is None else rmax)
        self.tmpdir = tempfile.mkdtemp(prefix="irltest")
        self.drf_file = None
        self.drf = None
        if drf:
            self.drf_file = os.path.join(self.tmpdir, "drf.drft")
            offs, psfs, start = synth.drf_for_orbit(self.drf_file, pix, n, self.views, half)
            self.drf = O.DrfTable(start, pix, offs, psfs, fft=fft)
        self.prj_model = (1 if atn else 0) | (2 if drf else 0)
        # noise-free data from the FP64 oracle projector, then Poisson
        self.atn_int = None if not atn else O.reorder(nz, n, n, mu * pix)
        x_int = O.reorder(nz, n, n, act)
        O.init_prj(nz, n, n, 7)
        prj = np.stack([O.project(nz, n, n, self.prj_model, pix, v.Angle, v.CFCR, self.drf,
                                  self.atn_int, x_int, pad=pad, avg=avg)
                        for v in self.views]).astype(np.float32)
        self.prj_clean = prj
        self.prj = synth.poisson(prj, counts, seed) if noise else prj.copy()
        self.x0 = np.full((nz, n, n), float(self.prj.sum()) / (n * n * nz * np.pi / 4 * nviews),
                          np.float32)

    def parms(self):
        p = IrlParms()
        p.NumPixels, p.NumSlices, p.NumViews = self.n, self.nz, self.nviews
        p.BinWidth = self.pix
        p.fAtnScaleFac = self.pix
        p.iNumSrfIterations = 0
        p.fScatEstFac = 1.0
        p.NumAngPerSubset = self.per
        p.NumIterations = self.iters
        p.pchNormImageBase = None
        p.SrfCollapseFac = 1
        return p

    def options(self, init_est=True):
        o = Options()
        o.bModelDrf = 1 if self.drf is not None else 0
        o.bReconIsInitEst = 1 if init_est else 0
        o.bFFTConvolve = 1 if (self.drf is not None and self.drf.fft) else 0
        o.iMsgLevel = 1
        o.iAxialPadLength = self.pad
        o.iAxialAvgLength = self.avg
        return o

    def oracle_views(self):
        return [dict(angle=v.Angle, left=v.Left, right=v.Right, cfcr=v.CFCR) for v in self.views]

    def oracle_osem(self, real="f64", iters=None, iter_cb=None, fast=0):
        x0 = O.reorder(self.nz, self.n, self.n, self.x0)
        prj = np.clip(self.prj, 0, None)
        x = O.osem(self.nz, self.n, self.oracle_views(), self.per,
                   self.iters if iters is None else iters, self.prj_model, self.prj_model,
                   self.pix, self.drf, prj, self.atn_int, x0, pad=self.pad, avg=self.avg,
                   real=real, iter_cb=iter_cb, fast=fast)
        return O.reorder(self.n, self.nz, self.n, x.reshape(self.n, self.nz, self.n))


def config_problem(name, seed=1, **kw):
    """Problem of one of the benchmark configurations (libirl_amd.synth.CONFIGS:
    sizes, pixel, views and subsets, orbit, phantom class, DRF support,
    counts); keyword arguments override them."""
    c = dict(synth.CONFIGS[name])
    args = dict(n=c["n"], nz=c["nz"], nviews=c["views"], per=c["vs"], iters=c["iters"],
                pix=c["pix"], seed=seed, half=c["half"], counts=c["counts"],
                inserts=c["inserts"], orbit=c["orbit"])
    if c["orbit"] == "circular":
        args["cfcr"] = c["cfcr"]
    else:
        args["cfcr"], args["rmax"] = c.get("rmin", 20.0), c.get("rmax", 28.0)
    args.update(kw)
    return Problem(**args)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
