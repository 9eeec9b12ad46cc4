"""GPU scatter CNN (libirl_amd/csrc/cnn.hip) against the FP64 oracle
(oracle/cnn.py): forward in FP32 and in the BF16-MFMA inference arithmetic,
the loss gradients of every layer, the source volumes, save/load, and the
learned scatter inside IrlOsem against the FP64 ESSE reconstruction."""
import os
import tempfile

import numpy as np
import pytest

from oracle import cnn as C
from oracle import oracle as O
from tests.cases import Problem, rel_l2

pytestmark = pytest.mark.gpu
MU0 = 0.15


def _net(seed=3, bias=0.05):
    import libirl_amd as L
    net = L.ScatterCnn.create(kappa=1.7, seed=seed)
    rng = np.random.default_rng(seed)
    params = [(W, rng.normal(0, bias, size=b.shape)) for W, b in C.init_params(seed)]
    net.set_params(L.pack_cnn_params(params))
    return net, params


def _grid_input(shape, seed):
    rng = np.random.default_rng(seed)
    inp = np.stack([rng.random(shape) * 2.0, rng.random(shape)], axis=-1)
    return inp.astype(np.float32)


# 12 x 10 x 14 = 1680 voxels: not a multiple of the 64-voxel MFMA wave tile
@pytest.mark.parametrize("shape", [(12, 10, 14), (16, 16, 16)])
def test_forward_fp32_matches_oracle(shape):
    print("CNN forward FP32", shape, flush=True)
    net, params = _net()
    inp = _grid_input(shape, 1)
    ref = C.forward(params, inp.astype(np.float64))[-1]
    got = net.forward(inp, bf16=False)
    e = rel_l2(got, ref)
    print("rel L2 %.3g" % e)
    assert e < 1e-5


@pytest.mark.parametrize("shape", [(12, 10, 14), (16, 16, 32)])
def test_forward_bf16_matches_oracle(shape):
    print("CNN forward BF16 MFMA", shape, flush=True)
    net, params = _net()
    inp = _grid_input(shape, 2)
    ref = C.forward(params, inp.astype(np.float64))[-1]
    got = net.forward(inp, bf16=True)
    e = rel_l2(got, ref)
    print("rel L2 %.3g" % e)
    assert e < 2e-2


# Z = 40: rows of 64 slices cover it with one partial row, of 32 and 16
# with a partial last row (16: none)
def test_row_lengths_give_identical_outputs():
    print("CNN forward BF16 MFMA, rows of 64, 32 and 16 slices", flush=True)
    net, _ = _net()
    inp = _grid_input((12, 10, 40), 5)
    outs = []
    try:
        for rz in (64, 32, 16):
            net.set_row_length(rz)
            outs.append(net.forward(inp, bf16=True))
    finally:
        net.set_row_length(64)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    with pytest.raises(RuntimeError):
        net.set_row_length(48)


def test_loss_gradients_match_oracle():
    print("CNN loss gradients", flush=True)
    import libirl_amd as L
    net, params = _net(seed=7)
    shape = (10, 12, 8)
    inp = _grid_input(shape, 3)
    tgt = np.random.default_rng(4).normal(size=shape + (2,)).astype(np.float32)
    acts = C.forward(params, inp.astype(np.float64))
    # FP32 and FP64 take the same ReLU branches: no pre-activation of this
    # case lies within FP32 rounding of zero (a flipped branch moves a
    # gradient by a whole term)
    a = inp.astype(np.float64)
    for (W, b), (_, _, dil, relu) in zip(params, C.LAYERS):
        pre = C.conv(a, W, b, dil)
        if relu:
            assert np.abs(pre).min() > 2e-5
        a = np.maximum(pre, 0) if relu else pre
    ref_loss, d = C.mse(acts[-1], tgt.astype(np.float64))
    ref = C.backward(params, acts, d)
    loss, g = net.loss_grad(inp, tgt)
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    for l, ((gW, gb), (rW, rb)) in enumerate(zip(L.unpack_cnn_params(g), ref)):
        eW, eb = rel_l2(gW, rW), rel_l2(gb, rb)
        print("layer %d: dW rel L2 %.3g, db rel L2 %.3g" % (l, eW, eb))
        assert eW < 1e-4 and eb < 1e-4


def test_save_load_round_trip():
    import libirl_amd as L
    net, _ = _net(seed=5)
    f = os.path.join(tempfile.mkdtemp(prefix="irlcnn"), "net.bin")
    net.save(f)
    back = L.ScatterCnn.load_file(f)
    p0, k0 = net.params()
    p1, k1 = back.params()
    assert k0 == k1 and np.array_equal(p0, p1)


def test_sources_match_oracle():
    print("CNN sources N=32 Nz=24", flush=True)
    net, params = _net(seed=6, bias=0.2)
    N, Nz = 32, 24
    rng = np.random.default_rng(6)
    x = (rng.random((N, N, Nz)) * 3.0).astype(np.float32)
    mu = (rng.random((N, N, Nz)) * 0.05).astype(np.float32)
    fdens, cntr = 1.0 / (MU0 * 0.33), 0.1
    g0, g1 = net.sources(x, mu, fdens, cntr=cntr, primary=0.9)
    x64, dens = x.astype(np.float64), mu.astype(np.float64) * fdens
    sx = x64.mean()
    out = C.forward(params, C.network_input(x64, dens, sx))[-1]
    r0, r1 = C.sources(x64, dens, out, sx, 1.7, primary=0.9, cntr=cntr)
    e0 = rel_l2(g0 - 0.9 * x, r0 - 0.9 * x64)
    e1 = rel_l2(g1, r1)
    print("scatter part of g0 rel L2 %.3g, g1 rel L2 %.3g" % (e0, e1))
    assert e0 < 2e-2 and e1 < 2e-2


def _krnl_file(noffs, ndepths, pix):
    from libirl_amd import synth
    k, km = synth.gaussian_mixture_krnl(noffs, ndepths, pix, mu0=MU0)
    k = synth.scale_krnl_to_fraction(k, 0.3)
    f = os.path.join(tempfile.mkdtemp(prefix="irlkrnl"), "k.krnl")
    synth.write_krnl(f, k, km, MU0, pix)
    return f, k, km


def _osem_pair(pb, f, k, km, srf_iters):
    """(FP64 oracle ESSE reconstruction, IrlOsem arguments) of a problem."""
    esse = O.Esse(k, km, MU0, pb.pix, pb.n, pb.nz, real="f64")
    x0 = O.reorder(pb.nz, pb.n, pb.n, pb.x0)
    ref = O.osem(pb.nz, pb.n, pb.oracle_views(), pb.per, pb.iters, pb.prj_model | 4,
                 pb.prj_model, pb.pix, pb.drf, np.clip(pb.prj, 0, None), pb.atn_int, x0,
                 esse=esse.h, srf_iterations=srf_iters)
    return O.reorder(pb.n, pb.nz, pb.n, ref.reshape(pb.n, pb.nz, pb.n))


def _gpu_osem(pb, f, srf_iters, model_srf=True):
    import libirl_amd as L
    parms = pb.parms()
    parms.iNumSrfIterations = srf_iters
    atn, prj, recon = pb.mu.copy(), pb.prj.copy(), pb.x0.copy()
    L.set_adv_options(0, 0)
    rc = L.irl_osem(parms, pb.options(), pb.views, pb.drf_file, f if model_srf else None, None,
                    atn, prj, recon)
    assert rc == 0, L.error_string()
    return recon


def test_irlosem_first_order_source_with_exact_terms():
    # The learned path's source with the scatter model's own FFT terms in
    # the unrotated frame: its distance to the per-view ESSE reconstruction
    # is the floor any network can reach (first order in d fMu, one frame).
    # This kernel's support (+-4 lateral, -5..3 in depth pixels) is far
    # narrower than its widest Gaussian, so its truncation is anisotropic and
    # the single-frame terms differ from the per-view ones by ~1%; at the
    # 128^3 workload (+-16 pixels) the floor is 0.4%.
    print("IrlOsem + first-order scatter source (exact terms) N=32 Nz=16 V=8", flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False)
    f, k, km = _krnl_file(5, 9, pb.pix)
    ref = _osem_pair(pb, f, k, km, 3)
    L.set_scatter_cnn_exact(True)
    L.set_scatter_cnn(None, check_every=0, max_nrmse=1.0)
    try:
        got = _gpu_osem(pb, f, 3)
        st = L.cnn_status()
    finally:
        L.set_scatter_cnn_exact(False)
        L.set_scatter_cnn(None, check_every=0, max_nrmse=0.01)   # the library defaults
    e = rel_l2(got, ref)
    print("rel L2 vs FP64 ESSE %.3g, status %s" % (e, st))
    assert st["checks"] == 1 and st["active"] and st["used"] == 12
    assert st["last_nrmse"] < 0.01
    assert e < 0.02


def test_trained_network_gains_over_no_scatter_model():
    # A short training run on ESSE-generated pairs of this problem's size and
    # scatter kernel; the reconstruction with the network (kept in use) lands
    # much closer to the FP64 ESSE reconstruction than one without any
    # scatter model.
    print("train CNN N=32 Nz=16, IrlOsem with it", flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False)
    f, k, km = _krnl_file(5, 9, pb.pix)
    ref = _osem_pair(pb, f, k, km, 3)
    net, st = L.ScatterCnn.train(f, 32, 16, pb.pix, pairs=160, val=16, epochs=12, batch=4,
                                 lr=3e-3, seed=2)
    print("training", st, flush=True)
    assert st["val_nrmse0"] < 0.2
    L.set_scatter_cnn(net, check_every=0, max_nrmse=1.0)
    try:
        got = _gpu_osem(pb, f, 3)
        cs = L.cnn_status()
    finally:
        L.set_scatter_cnn(None)
    plain = _gpu_osem(pb, f, 0, model_srf=False)
    e_cnn, e_none = rel_l2(got, ref), rel_l2(plain, ref)
    print("rel L2 vs FP64 ESSE: CNN %.3g, no scatter model %.3g, status %s" % (e_cnn, e_none, cs))
    assert cs["used"] == 12
    assert e_cnn < 0.25 * e_none
