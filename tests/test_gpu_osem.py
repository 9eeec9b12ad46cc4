"""GPU parity of the OSEM path (IrlOsem through the C ABI) against the oracle."""
import numpy as np
import pytest

from tests.cases import Problem, rel_l2

pytestmark = pytest.mark.gpu


def _gpu_osem(pb, **kw):
    import libirl_amd as L
    atn = None if pb.mu is None else pb.mu.copy()
    prj = pb.prj.copy()
    recon = pb.x0.copy()
    L.set_adv_options(0, 0)
    rc = L.irl_osem(pb.parms(), pb.options(), pb.views, pb.drf_file, None, None, atn, prj,
                    recon, **kw)
    assert rc == 0, L.error_string()
    return recon


@pytest.mark.parametrize("atn,drf", [(True, True), (False, False), (True, False), (False, True)])
def test_irlosem_small_matches_reference_semantics(atn, drf):
    print("case: N=32 Nz=16 V=16 atn=%s drf=%s" % (atn, drf), flush=True)
    pb = Problem(atn=atn, drf=drf)
    ref32 = pb.oracle_osem("f32")
    ref64 = pb.oracle_osem("f64")
    got = _gpu_osem(pb)
    e32, e64 = rel_l2(got, ref32), rel_l2(got, ref64)
    print("rel L2 vs f32 %.3g vs f64 %.3g (f32 vs f64 %.3g)" % (e32, e64, rel_l2(ref32, ref64)))
    assert e32 <= 1e-4 and e64 <= 1e-4


def test_irlosem_tall_volume_uses_wide_tiles():
    # Nz >= 128 selects 32 slices per thread in the DRF kernels
    print("case: N=64 Nz=128 V=8 atn+drf", flush=True)
    pb = Problem(n=64, nz=128, nviews=8, per=2, iters=1, half=7, pix=0.4)
    ref32 = pb.oracle_osem("f32")
    got = _gpu_osem(pb)
    e = rel_l2(got, ref32)
    print("rel L2 vs f32 %.3g" % e)
    assert e <= 1e-4


def test_irlosem_more_views_per_subset_than_a_wave_and_ragged_slices():
    # 72 views in one subset (the back rotation takes the views' angles in
    # groups of 64 lanes) and 70 slices (the last wave of slices is partly idle)
    print("case: N=16 Nz=70 V=72 one subset atn+drf", flush=True)
    pb = Problem(n=16, nz=70, nviews=72, per=72, iters=2, half=3, pix=0.5)
    ref32 = pb.oracle_osem("f32")
    ref64 = pb.oracle_osem("f64")
    got = _gpu_osem(pb)
    e32, e64 = rel_l2(got, ref32), rel_l2(got, ref64)
    print("rel L2 vs f32 %.3g vs f64 %.3g" % (e32, e64))
    assert e32 <= 1e-4 and e64 <= 1e-4
