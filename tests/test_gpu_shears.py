"""GPU parity of the three-shear rotation (bFastRotate, fastrotate.c:330-589):
vShearsRotate bit for bit against the FP32 oracle, Project/BackProject with
bFastRotate, and IrlOsem with the reference's default rotation."""
import numpy as np
import pytest

from libirl_amd import synth
from oracle import oracle as O
from tests.cases import Problem, rel_l2

pytestmark = pytest.mark.gpu

ANGLES = [0.0, 0.3, -0.6, 1.2, 2.6, 3.5, -2.5, 4.4, 0.5 * np.pi]


@pytest.mark.parametrize("angle", ANGLES)
@pytest.mark.parametrize("n", [32, 48])
def test_shears_rotate_bit_exact(angle, n):
    print("vShearsRotate angle=%g n=%d" % (angle, n), flush=True)
    import libirl_amd as L
    rng = np.random.default_rng(7)
    nz = 5
    v = rng.random((n, nz, n)).astype(np.float32)
    got = L.shears_rotate(angle, v, nz, n)
    ref = O.shears_rotate(angle, v, nz, n, real="f32")
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("model", [3, 0, 1])
def test_project_backproject_fast_rotate(model):
    print("Project/BackProject bFastRotate model=%d" % model, flush=True)
    import libirl_amd as L
    rng = np.random.default_rng(8)
    n, nz, pw = 32, 12, 0.5
    x = rng.random((n, nz, n)).astype(np.float32)
    mu = (rng.random((n, nz, n)) * 0.2).astype(np.float32)
    offs, psfs = synth.gaussian_drf_table(pw, pw / 2, 90, pw, 5)
    tab = L.make_drftab(pw / 2, pw, offs, psfs)
    otab = O.DrfTable(pw / 2, pw, offs, psfs)
    L.init_prj(nz, n, n, 7)
    O.init_prj(nz, n, n, 7, real="f32")
    for ang, cfcr in [(0.3, 9.0), (-1.7, 12.0), (2.9, 7.5)]:
        p = L.project(nz, n, n, model, pw, ang, cfcr, tab, mu, x, fast=1)
        r = O.project(nz, n, n, model, pw, ang, cfcr, otab, mu, x, real="f32", fast=1)
        assert rel_l2(p, r) < 1e-5
        ratio = rng.random((nz, n)).astype(np.float32)
        img = (rng.random((n, nz, n)) * 3e-5).astype(np.float32)
        ref = img.copy()
        L.backproject(nz, n, n, model, pw, ang, cfcr, tab, mu, ratio, img, atn_rotated=1,
                      fast=1)
        O.backproject(nz, n, n, model, pw, ang, cfcr, otab, mu, ratio, ref, atn_rotated=1,
                      real="f32", fast=1)
        assert rel_l2(img, ref) < 1e-5
    L.free_prj()


@pytest.mark.parametrize("atn,drf", [(True, True), (False, False)])
def test_irlosem_default_shear_rotation(atn, drf):
    print("IrlOsem bFastRotate=1 atn=%s drf=%s" % (atn, drf), flush=True)
    import libirl_amd as L
    pb = Problem(atn=atn, drf=drf)
    ref = pb.oracle_osem("f32", fast=1)
    atn_in = None if pb.mu is None else pb.mu.copy()
    prj = pb.prj.copy()
    recon = pb.x0.copy()
    L.set_adv_options(1, 0)
    rc = L.irl_osem(pb.parms(), pb.options(), pb.views, pb.drf_file, None, None, atn_in, prj,
                    recon)
    L.set_adv_options(0, 0)
    assert rc == 0, L.error_string()
    e = rel_l2(recon, ref)
    print("rel L2 vs f32 oracle (shear rotation) %.3g" % e)
    assert e <= 1e-4
    # the shear and bilinear reconstructions differ by interpolation only
    assert rel_l2(ref, pb.oracle_osem("f32", fast=0)) < 0.1
