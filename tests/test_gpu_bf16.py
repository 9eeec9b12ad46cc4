"""GPU checks of IrlGenprj and of the BF16 MFMA precision variant of the DRF
convolution, each against the FP64 oracle on the same seeded inputs.

Per stage: the forward projection (rotation and attenuation in FP32, DRF in
BF16) through IrlGenprj, and the full OSEM reconstruction (projector,
backprojector, sensitivity and update) through IrlOsem.  Tolerances: the
plain FP32 path <= 1e-5 per projection and <= 1e-4 after OSEM; BF16 with
hi+lo split operands <= 2e-5 / 1e-4; plain BF16 <= 2e-3 per projection and
<= 1e-2 (1 % NRMSE) after OSEM."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import Problem, rel_l2

pytestmark = pytest.mark.gpu


def _genprj(L, pb, mode):
    prev = L.set_precision(mode)
    try:
        act = pb.act.copy()
        atn = None if pb.mu is None else pb.mu.copy()
        prj = np.zeros((pb.nviews, pb.nz, pb.n), np.float32)
        rc = L.irl_genprj(pb.parms(), pb.options(), pb.views, pb.drf_file, None, None, act, atn,
                          prj)
        assert rc == 0, L.error_string()
    finally:
        L.set_precision(prev)
    return prj


def _oracle_genprj(pb):
    # irlgenprj.c:229-264: Project with the shear rotation, times
    # ViewTime*Sensitivity (1 here), scaled by 1/NumAngles
    x_int = O.reorder(pb.nz, pb.n, pb.n, pb.act)
    O.init_prj(pb.nz, pb.n, pb.n, 7)
    out = np.stack([O.project(pb.nz, pb.n, pb.n, pb.prj_model, pb.pix, v.Angle, v.CFCR, pb.drf,
                              pb.atn_int, x_int, fast=1) for v in pb.views])
    return out * np.float32(1.0 / np.float32(pb.nviews))


CASES = [dict(n=32, nz=16, nviews=8, per=2), dict(n=64, nz=128, nviews=4, per=2, half=7, pix=0.4)]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("mode,tol", [("fp32", 1e-5), ("bf16x3", 2e-5), ("bf16-plain", 2e-3),
                                      ("bf16", 2e-3)])
def test_genprj_projection_error(case, mode, tol):
    kw = CASES[case]
    print("IrlGenprj %s %s" % (mode, kw), flush=True)
    import libirl_amd as L
    pb = Problem(noise=False, **kw)
    got = _genprj(L, pb, mode)
    ref = _oracle_genprj(pb)
    errs = [rel_l2(got[v], ref[v]) for v in range(pb.nviews)]
    print("per-view rel L2 vs FP64 oracle: max %.3g" % max(errs))
    assert max(errs) <= tol


@pytest.mark.parametrize("mode,tol", [("bf16x3", 1e-4), ("bf16-plain", 1e-2), ("bf16", 1e-2)])
def test_osem_bf16_nrmse(mode, tol):
    print("IrlOsem %s" % mode, flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=16, per=4, iters=2)
    ref = pb.oracle_osem("f64")
    prev = L.set_precision(mode)
    try:
        atn = pb.mu.copy()
        prj = pb.prj.copy()
        recon = pb.x0.copy()
        L.set_adv_options(0, 0)
        rc = L.irl_osem(pb.parms(), pb.options(), pb.views, pb.drf_file, None, None, atn, prj,
                        recon)
        assert rc == 0, L.error_string()
    finally:
        L.set_precision(prev)
    e = rel_l2(recon, ref)
    print("OSEM rel L2 (NRMSE) vs FP64 oracle: %.3g" % e)
    assert e <= tol


@pytest.mark.parametrize("otype", [1, 2])
def test_compute_orbit_matches_oracle(otype):
    print("ComputeOrbit type %d" % otype, flush=True)
    import libirl_amd as L
    from libirl_amd import IrlParms, synth
    n = 64
    act, _ = synth.phantom(n, 8, 3)
    img = np.ascontiguousarray(act[4])
    views = synth.views_circular(30, 20.0, n)
    p = IrlParms()
    p.NumPixels, p.NumSlices, p.NumViews, p.BinWidth = n, 8, 30, 0.4
    g_img = img.copy()
    rc = L.compute_orbit(g_img, p, 2.5, views, 0.2, otype)
    assert rc == 0, L.error_string()
    o_img = img.copy()
    orc, ref = O.compute_orbit(o_img, 0.4, 2.5, [v.Angle for v in views], 0.2, otype)
    assert orc == 0
    assert np.array_equal(np.array([v.CFCR for v in views], np.float32), ref)
    assert np.array_equal(g_img, o_img)


# drf "fft": the 2-D FFT kernels, the backprojector's with the spectra of
# its own table
@pytest.mark.parametrize("grf_in_bck,drf", [(0, "reference"), (1, "reference"), (1, "fft")])
def test_irlosem_with_computed_grf(grf_in_bck, drf):
    """IrlOsem without a DRF table file (computed GRF, IrlSetup.c:243-334),
    or with a file plus the GRF in the backprojector (bUseGrfInBck), against
    the oracle run with the same tables."""
    print("IrlOsem computed GRF, GRF in backprojector=%d, drf=%s" % (grf_in_bck, drf), flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=16, per=4, iters=2, pix=0.4)
    parms = pb.parms()
    parms.fHoleDiam, parms.fHoleLen, parms.fBackToDet, parms.fIntrinsicFWHM = 0.15, 2.4, 0.5, 0.4
    # the table vComputeDrfs builds: distances of the orbit, half size NRot/2
    cf = [v.CFCR for v in pb.views]
    mn = min(cf) - pb.n // 2 * pb.pix + pb.pix / 2
    mn = pb.pix / 2 if mn < 0 else mn
    mx = max(cf) + pb.n // 2 * pb.pix - pb.pix / 2
    nd = int((np.float32(mx) - np.float32(mn)) / np.float32(pb.pix) + 1 + 0.5)
    psfs = L.gendrf(2.4, 0.15, 0.5, 0.4, 0.05, pb.pix, pb.pix, mn, nd, pb.n // 2)
    grf = O.DrfTable(mn, pb.pix, [((p.shape[1] + 1) // 2, (p.shape[0] + 1) // 2) for p in psfs],
                     psfs)
    x0 = O.reorder(pb.nz, pb.n, pb.n, pb.x0)
    prj_tab = pb.drf if grf_in_bck else grf
    ref = O.osem(pb.nz, pb.n, pb.oracle_views(), pb.per, pb.iters, pb.prj_model, pb.prj_model,
                 pb.pix, prj_tab, np.clip(pb.prj, 0, None), pb.atn_int, x0, real="f32",
                 bck_drf=grf if grf_in_bck else None)
    ref = O.reorder(pb.n, pb.nz, pb.n, ref.reshape(pb.n, pb.nz, pb.n))
    opts = pb.options()
    opts.bUseGrfInBck = grf_in_bck
    atn, prj, recon = pb.mu.copy(), pb.prj.copy(), pb.x0.copy()
    L.set_adv_options(0, 0)
    prev = L.set_drf_fft(drf)
    try:
        if drf == "fft":
            # the same setup as a device-resident reconstruction: both
            # directions on the FFT kernels
            r = L.Recon(parms, opts, pb.views, pb.drf_file if grf_in_bck else None, None, None,
                        pb.mu, pb.prj, pb.x0)
            assert r.drf_algorithms()[:2] == ("fft", "fft")
            r.close()
        rc = L.irl_osem(parms, opts, pb.views, pb.drf_file if grf_in_bck else None, None, None,
                        atn, prj, recon)
    finally:
        L.set_drf_fft(prev)
    assert rc == 0, L.error_string()
    e = rel_l2(recon, ref)
    print("rel L2 vs f32 oracle with the same tables: %.3g" % e)
    assert e <= 1e-4


def test_per_plane_modes_follow_the_measured_errors():
    # mode 'bf16': each DRF table entry takes the cheapest arithmetic whose
    # error, measured at setup against the FP32 kernel, is at most 1e-3; the
    # projections of the resulting table stay within that bound
    print("BF16 per-plane modes from measured errors", flush=True)
    import libirl_amd as L
    pb = Problem(n=64, nz=64, nviews=4, per=2, half=7, pix=0.4, noise=False)
    prev = L.set_precision("bf16")
    try:
        r = L.Recon(pb.parms(), pb.options(), pb.views, pb.drf_file, None, None, pb.mu, pb.prj,
                    pb.x0)
    finally:
        L.set_precision(prev)
    modes, err = r.bf16_planes()
    r.close()
    print("entries %d, modes %s, plain errors %.3g..%.3g" %
          (len(modes), np.bincount(modes, minlength=3).tolist(), err[:, 0].min(), err[:, 0].max()))
    assert len(modes) == len(pb.drf.offs) // 2 + 1
    assert np.all(err[:, 0] > 0) and np.all(err[:, 0] < 0.05)
    assert np.all((modes == 0) == (err[:, 0] <= 1e-3))
    assert np.all((modes == 1) == ((err[:, 0] > 1e-3) & (err[:, 1] <= 1e-3)))
    assert np.all(err[modes == 1, 1] < err[modes == 1, 0])
    got = _genprj(L, pb, "bf16")
    ref = _oracle_genprj(pb)
    e = max(rel_l2(got[v], ref[v]) for v in range(pb.nviews))
    print("projection rel L2 vs FP64 oracle %.3g" % e)
    assert e <= 1e-3
