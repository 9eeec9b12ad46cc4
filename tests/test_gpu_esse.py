"""GPU ESSE (CalcEffScatterSource, Project with MODEL_SRF, IrlOsem with a
scatter kernel) against the oracle."""
import os
import tempfile

import numpy as np
import pytest

from libirl_amd import synth
from oracle import oracle as O
from tests.cases import Problem, rel_l2

pytestmark = pytest.mark.gpu
MU0 = 0.15


def _krnl_file(noffs, ndepths, pix, **kw):
    k, km = synth.gaussian_mixture_krnl(noffs, ndepths, pix, mu0=MU0)
    f = os.path.join(tempfile.mkdtemp(prefix="irlkrnl"), "k.krnl")
    synth.write_krnl(f, k, km, MU0, pix, **kw)
    return f, k, km


def _gpu_effsrc(fname, pix, n, nz, collapse, act, mp):
    import ctypes
    import libirl_amd as L
    lib = L.load()
    lib.pvInitESSEScatterModel.restype = ctypes.c_void_p
    lib.pvInitESSEScatterModel.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_float,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.CalcEffScatterSource.argtypes = [ctypes.c_void_p, L._FP, L._FP, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, L._FP]
    lib.pvDeinitScatterPrj.argtypes = [ctypes.c_void_p]
    lib.pvDeinitScatterPrj.restype = ctypes.c_void_p
    h = lib.pvInitESSEScatterModel(fname.encode(), b"2 2 2 3", pix, n, nz, collapse, 0)
    assert h
    a = np.ascontiguousarray(act, np.float32)
    m = np.ascontiguousarray(mp, np.float32)
    out = np.zeros_like(a)
    lib.CalcEffScatterSource(h, L.fptr(a), L.fptr(m), n, n, nz, L.fptr(out))
    lib.pvDeinitScatterPrj(h)
    assert L.last_error() == 0
    return out


# The last case needs FFT lines of 192 along x (128 + 48) and 160 along
# depth (128 + 9): the mixed-radix line lengths.
@pytest.mark.parametrize("n,nz,noffs,ndepths,collapse", [
    (32, 16, 5, 9, 1), (32, 16, 5, 9, 2), (16, 8, 9, 33, 1), (64, 32, 9, 17, 4),
    (128, 16, 49, 17, 1)])
def test_effsrc_matches_oracle(n, nz, noffs, ndepths, collapse):
    print("ESSE n=%d nz=%d krnl %dx%d c=%d" % (n, nz, noffs, ndepths, collapse), flush=True)
    pix = 0.5
    f, k, km = _krnl_file(noffs, ndepths, pix * collapse)
    rng = np.random.default_rng(7)
    act = rng.random((n, nz, n))
    mp = rng.random((n, nz, n)) * MU0 * pix
    ref = O.Esse(k, km, MU0, pix, n, nz, collapse=collapse).effsrc(act, mp)
    got = _gpu_effsrc(f, pix, n, nz, collapse, act, mp)
    e = rel_l2(got, ref)
    print("rel L2 %.3g" % e)
    assert e < 2e-5


def test_irlosem_with_esse_matches_oracle():
    # 3 iterations, scatter modelled in the first two: exercises the ESSE
    # projection, the stored scatter estimate and its reuse
    print("IrlOsem + ESSE N=32 Nz=16 V=8", flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False)
    f, k, km = _krnl_file(5, 9, pb.pix)
    esse = O.Esse(k, km, MU0, pb.pix, pb.n, pb.nz, real="f64")
    x0 = O.reorder(pb.nz, pb.n, pb.n, pb.x0)
    ref = O.osem(pb.nz, pb.n, pb.oracle_views(), pb.per, pb.iters, pb.prj_model | 4,
                 pb.prj_model, pb.pix, pb.drf, np.clip(pb.prj, 0, None), pb.atn_int, x0,
                 esse=esse.h, srf_iterations=2)
    ref = O.reorder(pb.n, pb.nz, pb.n, ref.reshape(pb.n, pb.nz, pb.n))
    parms = pb.parms()
    parms.iNumSrfIterations = 2
    atn, prj, recon = pb.mu.copy(), pb.prj.copy(), pb.x0.copy()
    L.set_adv_options(0, 0)
    rc = L.irl_osem(parms, pb.options(), pb.views, pb.drf_file, f, None, atn, prj, recon)
    assert rc == 0, L.error_string()
    e = rel_l2(recon, ref)
    print("rel L2 vs f64 %.3g" % e)
    assert e < 1e-4


@pytest.mark.parametrize("noffs", [17, 49])
def test_irlosem_with_esse_mixed_radix_slices(noffs):
    # slice-fastest engine layout: FFT lines of 128 + (noffs - 1) along the
    # slices, 160 and 192 points
    nz = 128
    print("IrlOsem + ESSE N=32 Nz=%d krnl offsets %d" % (nz, noffs), flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=nz, nviews=2, per=2, iters=1, half=3, noise=False)
    f, k, km = _krnl_file(noffs, 9, pb.pix)
    esse = O.Esse(k, km, MU0, pb.pix, pb.n, pb.nz, real="f64")
    x0 = O.reorder(pb.nz, pb.n, pb.n, pb.x0)
    ref = O.osem(pb.nz, pb.n, pb.oracle_views(), pb.per, pb.iters, pb.prj_model | 4,
                 pb.prj_model, pb.pix, pb.drf, np.clip(pb.prj, 0, None), pb.atn_int, x0,
                 esse=esse.h, srf_iterations=1)
    ref = O.reorder(pb.n, pb.nz, pb.n, ref.reshape(pb.n, pb.nz, pb.n))
    parms = pb.parms()
    parms.iNumSrfIterations = 1
    atn, prj, recon = pb.mu.copy(), pb.prj.copy(), pb.x0.copy()
    L.set_adv_options(0, 0)
    rc = L.irl_osem(parms, pb.options(), pb.views, pb.drf_file, f, None, atn, prj, recon)
    assert rc == 0, L.error_string()
    e = rel_l2(recon, ref)
    print("rel L2 vs f64 %.3g" % e)
    assert e < 1e-4


def test_project_with_srf_matches_oracle():
    print("Project MODEL_SRF", flush=True)
    import ctypes
    import libirl_amd as L
    n, nz, pix = 32, 16, 0.5
    f, k, km = _krnl_file(5, 9, pix)
    lib = L.load()
    lib.pvInitESSEScatterModel.restype = ctypes.c_void_p
    lib.pvInitESSEScatterModel.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_float,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    h = lib.pvInitESSEScatterModel(f.encode(), b"2 2 2 3", pix, n, nz, 1, 0)

    class Srf(ctypes.Structure):
        _fields_ = [("c", ctypes.c_int), ("l", ctypes.c_int), ("p", ctypes.c_void_p)]
    srf = Srf(1, 0, h)
    rng = np.random.default_rng(9)
    x = rng.random((n, nz, n)).astype(np.float32)
    mu = (rng.random((n, nz, n)) * 0.08).astype(np.float32)
    esse = O.Esse(k, km, MU0, pix, n, nz)
    L.init_prj(nz, n, n, 7)
    O.init_prj(nz, n, n, 7)
    for model, primary in ((5, 1.0), (7, 0.8)):
        offs, psfs = synth.gaussian_drf_table(pix, pix / 2, 80, pix, 3)
        tab = L.make_drftab(pix / 2, pix, offs, psfs)
        otab = O.DrfTable(pix / 2, pix, offs, psfs)
        out = np.zeros((nz, n), np.float32)
        lib.Project(nz, n, n, model, pix, 0, 0.4, 11.0, L._tab(tab), ctypes.addressof(srf),
                    L.fptr(mu), L.fptr(x), L.fptr(out), primary, 1.0, 0, 0, 0, 0, 0)
        ref = O.project(nz, n, n, model, pix, 0.4, 11.0, otab, mu, x, primary=primary,
                        esse=esse.h)
        e = rel_l2(out, ref)
        print("model %d rel L2 %.3g" % (model, e))
        assert e < 2e-5
    L.free_prj()


def _esse_recon(pb, f, fused):
    import libirl_amd as L
    L.set_adv_options(0, 0)
    prev = L.set_esse_fused(fused)
    try:
        parms = pb.parms()
        parms.iNumSrfIterations = 1
        r = L.Recon(parms, pb.options(), pb.views, pb.drf_file, f, None, pb.mu, pb.prj, pb.x0)
    finally:
        L.set_esse_fused(prev)
    for s in range(r.num_subsets):
        r.subiteration(s)
    x = r.estimate()
    r.close()
    return x


# In-plane FFT lengths 128 (96 + 32) and 160 (128 + 32) with at least 1024
# depth planes per batch (12 x 96, 8 x 128): the 2-D plane transforms and the
# source formed in their last pass, against the line passes and the separate
# source kernel.
@pytest.mark.parametrize("n,nviews,per", [(96, 24, 12), (128, 16, 8)])
def test_esse_plane_transforms_match_line_passes(n, nviews, per):
    print("ESSE plane transforms vs line passes N=Nz=%d views/subset %d" % (n, per), flush=True)
    pb = Problem(n=n, nz=n, nviews=nviews, per=per, iters=1, drf=False, noise=False,
                 pix=0.5 if n <= 96 else 0.4)
    f, k, km = _krnl_file(33, 17, pb.pix)
    a = _esse_recon(pb, f, False)
    b = _esse_recon(pb, f, True)
    e = rel_l2(b, a)
    print("rel L2 plane/fused vs line passes %.3g" % e)
    assert e <= 1e-5
