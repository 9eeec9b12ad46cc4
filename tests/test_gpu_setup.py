"""IrlOsem's setup-side behaviours on the GPU through the C-ABI, each against
the oracle: the computed initial estimate (initial_est.c:29-180), measured-bin
masks from PrjView_t.Left/Right (osemrecon.c:92-184), sensitivity images on
disk (pchNormImageBase, osemrecon.c:266-368), resuming at iStartIteration
(osemrecon.c:628-636), and the learned scatter's fallback to ESSE."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import Problem, rel_l2

pytestmark = pytest.mark.gpu
MU0 = 0.15


def _irlosem(pb, parms=None, opts=None, krnl=None, recon=None, iter_cb=None, views=None):
    import libirl_amd as L
    atn, prj = pb.mu.copy(), pb.prj.copy()
    recon = pb.x0.copy() if recon is None else recon
    rc = L.irl_osem(parms or pb.parms(), opts or pb.options(), views or pb.views, pb.drf_file,
                    krnl, None, atn, prj, recon, iter_cb=iter_cb)
    assert rc == 0, L.error_string()
    return recon


@pytest.mark.parametrize("contour,thresh", [(False, 0.0), (True, 0.0), (True, 0.1),
                                            (False, 0.12)])
def test_initial_estimate_matches_oracle(contour, thresh):
    print("initial estimate contour=%s thresh=%s" % (contour, thresh), flush=True)
    import libirl_amd as L
    # an orbit close to the body so the detector faces cut the square
    pb = Problem(n=32, nz=8, nviews=8, per=2, iters=1, half=3, orbit="ellipse", cfcr=5.0)
    opts = pb.options(init_est=False)
    opts.bUseContourSupport = int(contour)
    opts.fAtnMapThresh = thresh
    p = pb.parms()
    p.NumIterations = 0
    seen = {}
    L.set_adv_options(0, 0)
    _irlosem(pb, p, opts, iter_cb=lambda it, x: seen.setdefault(it, x))
    got = seen[0].reshape(pb.nz, pb.n, pb.n)
    # IrlOsem hands the reordered, not yet scaled map to the estimate
    atn_int = O.reorder(pb.nz, pb.n, pb.n, pb.mu).astype(np.float32)
    want, _ = O.initial_estimate(pb.n, pb.nz, pb.pix, pb.views, np.clip(pb.prj, 0, None),
                                 atn=atn_int, contour=contour, thresh=thresh)
    want = O.reorder(pb.n, pb.nz, pb.n, want).astype(np.float32)
    zeros = int((want == 0).sum())
    print("zeroed voxels %d of %d" % (zeros, want.size))
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
    assert np.array_equal(got == 0, want == 0)
    if contour or thresh:
        full, _ = O.initial_estimate(pb.n, pb.nz, pb.pix, pb.views, np.clip(pb.prj, 0, None))
        assert zeros > int((full == 0).sum())   # the support removed something


def test_measured_bin_masks_match_oracle():
    print("IrlOsem with Left/Right masks N=32 V=8", flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=2, half=3)
    for k, v in enumerate(pb.views):
        v.Left = -4.5 + 1.25 * k          # the first bins unmeasured in early views
        v.Right = 31.0 - 2.0 * k          # the norm mask cut at the right in later ones
    L.set_adv_options(0, 0)
    got = _irlosem(pb)
    ref32, ref64 = pb.oracle_osem("f32"), pb.oracle_osem("f64")
    full = Problem(n=32, nz=16, nviews=8, per=2, iters=2, half=3)
    unmasked = _irlosem(full)
    e32, e64, d = rel_l2(got, ref32), rel_l2(got, ref64), rel_l2(unmasked, ref64)
    print("rel L2 vs f32 %.3g vs f64 %.3g; unmasked run %.3g away" % (e32, e64, d))
    assert e32 <= 1e-4 and e64 <= 1e-4
    assert d > 1e-3                       # the masks matter in this case


def test_norm_images_on_disk():
    print("IrlOsem with pchNormImageBase", flush=True)
    import libirl_amd as L
    lib = L.load()
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=1, half=3)
    L.set_adv_options(0, 0)
    mem = _irlosem(pb)
    base = os.path.join(tempfile.mkdtemp(prefix="irlnorm"), "recon")
    p = pb.parms()
    p.pchNormImageBase = base.encode()
    n = pb.n * pb.n * pb.nz
    checked = []

    def post_subset(d, ctx):
        s = d.contents.iSubsetIdx
        fname = "%s.norm.%d.dat" % (base, s)
        disk = np.fromfile(fname, np.float32)
        norm = np.ctypeslib.as_array(d.contents.pfNorm, shape=(n,))
        checked.append((s, disk.size == n and np.array_equal(disk, norm)))
    cb = L.HOOK_CB(post_subset)
    lib.OsemSetPostSubsetHook(cb, None)
    try:
        got = _irlosem(pb, p)
    finally:
        lib.OsemSetPostSubsetHook(ctypes.cast(None, L.HOOK_CB), None)
    assert checked == [(s, True) for s in range(4)]
    assert np.array_equal(got, mem)
    # the files are removed at the end (osemrecon.c:376-391)
    assert not any(os.path.exists("%s.norm.%d.dat" % (base, s)) for s in range(4))


def _krnl(pix):
    from libirl_amd import synth
    k, km = synth.gaussian_mixture_krnl(5, 9, pix, mu0=MU0)
    k = synth.scale_krnl_to_fraction(k, 0.3)
    f = os.path.join(tempfile.mkdtemp(prefix="irlkrnl"), "k.krnl")
    synth.write_krnl(f, k, km, MU0, pix)
    return f, k, km


@pytest.mark.parametrize("scatter", [False, True])
def test_resume_at_start_iteration_matches_oracle(scatter):
    # iterations start+1..NumIterations from the given estimate; with the
    # scatter model the first resumed iteration's decision reads a start of
    # 0 (osemrecon.c:628 before :636)
    print("IrlOsem resume at iteration 2 of 4, scatter=%s" % scatter, flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=4, half=3, noise=False)
    p = pb.parms()
    krnl, esse, model, srf = None, None, pb.prj_model, 0
    if scatter:
        krnl, k, km = _krnl(pb.pix)
        esse = O.Esse(k, km, MU0, pb.pix, pb.n, pb.nz, real="f64")
        p.iNumSrfIterations = srf = 3
        model |= 4
    L.set_adv_options(0, 2)
    try:
        got = _irlosem(pb, p, krnl=krnl)
    finally:
        L.set_adv_options(0, 0)
    x0 = O.reorder(pb.nz, pb.n, pb.n, pb.x0)
    ref = O.osem(pb.nz, pb.n, pb.oracle_views(), pb.per, 4, model, pb.prj_model, pb.pix, pb.drf,
                 np.clip(pb.prj, 0, None), pb.atn_int, x0, start_iteration=2,
                 esse=None if esse is None else esse.h, srf_iterations=srf)
    ref = O.reorder(pb.n, pb.nz, pb.n, ref.reshape(pb.n, pb.nz, pb.n))
    def fresh(iters):
        x = O.osem(pb.nz, pb.n, pb.oracle_views(), pb.per, iters, model, pb.prj_model, pb.pix,
                   pb.drf, np.clip(pb.prj, 0, None), pb.atn_int, x0,
                   esse=None if esse is None else esse.h, srf_iterations=srf)
        return O.reorder(pb.n, pb.nz, pb.n, x.reshape(pb.n, pb.nz, pb.n))
    e = rel_l2(got, ref)
    d4, d2 = rel_l2(fresh(4), ref), rel_l2(fresh(2), ref)
    print("rel L2 vs f64 %.3g (4 fresh iterations %.3g, 2 fresh %.3g away)" % (e, d4, d2))
    assert e <= 1e-4
    assert d4 > 1e-3            # two iterations ran, from x0
    if scatter:
        # the resumed iterations follow the scatter schedule of iterations
        # 3-4 (model and store, then the stored estimate), not of 1-2
        assert d2 > 1e-6
    else:
        assert d2 < 1e-6        # without scatter the iteration number plays no part


@pytest.mark.parametrize("graphs,precision", [(False, "fp32"), (True, "fp32"), (True, "bf16")])
def test_learned_scatter_falls_back_to_esse(graphs, precision):
    # a network that is far from ESSE (random weights, large output scale):
    # the first check exceeds 1 % and the reconstruction continues on ESSE,
    # with the same floats as a reconstruction that never had a network.
    # The check runs beside the learned projection and gates the update on
    # the device; with the BF16 DRF it reads the projector's bf16 rows.
    print("learned-scatter fallback, graphs=%s precision=%s" % (graphs, precision), flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False)
    krnl, _, _ = _krnl(pb.pix)
    p = pb.parms()
    p.iNumSrfIterations = 3
    net = L.ScatterCnn.create(kappa=50.0, seed=11)
    L.set_adv_options(0, 0)
    prev = L.set_graphs(graphs)
    prev_p = L.set_precision(precision)
    try:
        plain = _irlosem(pb, p, krnl=krnl)
        L.set_scatter_cnn(net, check_every=2, max_nrmse=0.01)
        got = _irlosem(pb, p, krnl=krnl)
        st = L.cnn_status()
    finally:
        L.set_scatter_cnn(None, check_every=0, max_nrmse=0.01)
        L.set_graphs(prev)
        L.set_precision(prev_p)
    print("status", st)
    assert st["checks"] == 1 and not st["active"] and st["used"] == 0
    assert st["last_nrmse"] > 0.01
    assert np.array_equal(got, plain)
