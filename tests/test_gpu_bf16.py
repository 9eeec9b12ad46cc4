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
