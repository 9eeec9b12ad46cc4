"""GPU parity of the projector-level entry points (InitPrj / Project /
BackProject / rot3dpar / ConvolveDrfYpad) against the oracle."""
import numpy as np
import pytest

from libirl_amd import synth
from oracle import oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def _lib():
    import libirl_amd as L
    return L


def _table(pw, half, n=90, fft=False):
    offs, psfs = synth.gaussian_drf_table(pw, pw / 2, n, pw, half)
    return offs, psfs


@pytest.mark.parametrize("angle", [0.0, 0.7, -2.1, 3.3])
@pytest.mark.parametrize("pb,n,nrot", [(0, 32, 32), (0, 30, 36), (1, 36, 36)])
def test_rot3dpar_bit_exact(angle, pb, n, nrot):
    print("rot3dpar angle=%g pb=%d n=%d nrot=%d" % (angle, pb, n, nrot), flush=True)
    L = _lib()
    rng = np.random.default_rng(3)
    nz = 5
    src = n if pb == 0 else nrot
    v = rng.random((src, nz, src)).astype(np.float32)
    got = L.rot3dpar(angle, v, pb, nz, n, nrot)
    ref = O.rot3dpar(angle, v, pb, nz, n, nrot, real="f32")
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("shape,psf_shape,pad,avg,fft", [
    ((20, 24), (7, 5), 0, 0, 0), ((128, 128), (31, 31), 0, 0, 0), ((33, 50), (9, 13), 3, 2, 0),
    ((16, 40), (15, 7), 9, 20, 0), ((16, 40), (15, 7), 3, 2, 1), ((70, 130), (63, 61), 5, 4, 0)])
def test_convolve_drf_ypad(shape, psf_shape, pad, avg, fft):
    print("ConvolveDrfYpad", shape, psf_shape, pad, avg, fft, flush=True)
    L = _lib()
    rng = np.random.default_rng(4)
    a = rng.random(shape).astype(np.float32)
    p = rng.random(psf_shape).astype(np.float32)
    got = L.convolve_drf_ypad(a, p, fft, pad, avg)
    ref = O.conv2d_ypad(a, p, 0 if fft else pad, 0 if fft else avg, real="f64")
    assert rel_l2(got, ref) < 1e-6


@pytest.mark.parametrize("model,nrot,pad,avg,primary,atnfac", [
    (3, 32, 0, 0, 1.0, 1.0), (0, 32, 0, 0, 1.0, 1.0), (1, 32, 0, 0, 0.7, 1.3),
    (2, 32, 4, 3, 1.0, 1.0), (3, 36, 2, 1, 2.0, 0.5)])
def test_project_backproject(model, nrot, pad, avg, primary, atnfac):
    print("Project/BackProject model=%d nrot=%d pad=%d avg=%d" % (model, nrot, pad, avg),
          flush=True)
    L = _lib()
    rng = np.random.default_rng(5)
    n, nz, pw = 32, 12, 0.5
    x = rng.random((n, nz, n)).astype(np.float32)
    mu = (rng.random((n, nz, n)) * 0.2).astype(np.float32)
    offs, psfs = _table(pw, 5)
    tab = L.make_drftab(pw / 2, pw, offs, psfs)
    otab = O.DrfTable(pw / 2, pw, offs, psfs)
    L.init_prj(nz, n, nrot, 7)
    O.init_prj(nz, n, nrot, 7, real="f32")
    O.init_prj(nz, n, nrot, 7, real="f64")
    for k, (ang, cfcr) in enumerate([(0.3, 9.0), (-1.7, 12.0), (2.9, 7.5)]):
        p = L.project(nz, n, nrot, model, pw, ang, cfcr, tab, mu, x, primary, atnfac, pad, avg)
        r = O.project(nz, n, nrot, model, pw, ang, cfcr, otab, mu, x, primary, atnfac, pad, avg,
                      real="f64")
        assert rel_l2(p, r) < 1e-5, k
        # reuse of the rotated estimate and attenuation map (bAct/bAtn rotated)
        p2 = L.project(nz, n, nrot, model, pw, ang, cfcr, tab, None, None, primary, atnfac, pad,
                       avg, atn_rotated=1, act_rotated=1)
        assert np.array_equal(p, p2)
        ratio = rng.random((nz, nrot)).astype(np.float32)
        img = (rng.random((n, nz, n)) * 3e-5).astype(np.float32)
        img64 = img.astype(np.float64)
        L.backproject(nz, n, nrot, model, pw, ang, cfcr, tab, mu, ratio, img, primary, atnfac,
                      pad, avg, atn_rotated=1)
        O.backproject(nz, n, nrot, model, pw, ang, cfcr, otab, mu, ratio, img64, primary,
                      atnfac, pad, avg, atn_rotated=0, real="f64")
        assert rel_l2(img, img64) < 1e-5, k
    L.free_prj()
    assert L.last_error() == 0
