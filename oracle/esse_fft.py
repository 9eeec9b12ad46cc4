"""CalcEffScatterSource (krnlprj3d.c:676-954) with its products evaluated as
real 3-D FFT convolutions on the model's padded grid, as the reference
computes them (realft3d / multrealft3d / realinvft3d, krnlprj3d.c:783-861;
libmip's FFT is absent here, scipy.fft stands in, in the precision of the
input: float32 for the reference-semantics CPU baseline).

TEST INFRASTRUCTURE AND CPU BASELINE ONLY: used by tests/ and by bench.py's
cpu_baseline leg (the reference's per-view scatter cost); the product never
imports it.  Collapse factors other than 1 are not restated here (the
oracle's orc_esse_effsrc covers them).
"""
import numpy as np
import scipy.fft as sf


class EsseFFT:
    """The model of an oracle.Esse handle with its kernel spectra."""

    def __init__(self, esse, dtype=np.float32, workers=1):
        self.dims, self.p = esse.info()
        if self.p["collapse"] != 1:
            raise ValueError("the FFT restatement covers collapse factor 1 only")
        self.dtype = dtype
        self.workers = workers
        self.spec = [sf.rfftn(esse.grid(k).astype(dtype), workers=workers)
                     for k in range(min(self.p["nterms"], 3))]

    def _conv(self, F, k, shape):
        out = sf.irfftn(F * self.spec[k], s=self.dims, workers=self.workers)
        return out[:shape[0], :shape[1], :shape[2]]

    def effsrc(self, act, mp):
        """Effective source of [depth][slice][bin] volumes act and mp (the
        map in per-pixel units), as orc_esse_effsrc."""
        act = np.asarray(act, self.dtype)
        mp = np.asarray(mp, self.dtype)
        F = sf.rfftn(act, s=self.dims, workers=self.workers)
        eff = self._conv(F, 0, act.shape)
        mu0 = np.float32(self.p["mu0"]) if self.dtype == np.float32 else self.p["mu0"]
        # half-pixel running sum of the map along depth (the depth loops of
        # krnlprj3d.c:826-947), in units of 1/mu0
        half = mp * self.dtype(0.5)
        S = np.cumsum(mp, axis=0, dtype=self.dtype) - half
        d = S * self.dtype(1.0 / mu0)
        if self.p["nterms"] > 1:
            eff = eff - d * self._conv(F, 1, act.shape)
        if self.p["nterms"] > 2:
            eff = eff + self.dtype(0.5) * d * d * self._conv(F, 2, act.shape)
        fmu = self.p["mu0"] - self.p["muavg"]
        eff = eff * np.exp((d * fmu).astype(np.float64)).astype(self.dtype)
        if self.p["cntr"] > 0:
            eff = eff + act * self.dtype(self.p["cntr"])
        dens = self.dtype(1.0 / self.p["mu0"] / self.p["pixsize"])
        eff = eff * (mp * dens)
        return np.maximum(eff, 0)
