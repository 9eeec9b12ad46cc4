"""Graph replay of sub-iterations, the batch of reconstructions of uneven
sizes stepped in one parent-graph launch, and the DRF tile autotuner: each
against the eager path of the same reconstruction, bit for bit (same kernels,
arguments and order), with eager work and host synchronisation between the
replays."""
import os
import tempfile

import numpy as np
import pytest

from tests.cases import Problem

pytestmark = pytest.mark.gpu
MU0 = 0.15


def _krnl_file(pix):
    from libirl_amd import synth
    k, km = synth.gaussian_mixture_krnl(5, 9, pix, mu0=MU0)
    k = synth.scale_krnl_to_fraction(k, 0.3)
    f = os.path.join(tempfile.mkdtemp(prefix="irlkrnl"), "k.krnl")
    synth.write_krnl(f, k, km, MU0, pix)
    return f


def _recon(pb, krnl=None, srf_iters=0, graphs=False, autotune=False, precision="fp32",
           drf="reference"):
    import libirl_amd as L
    L.set_adv_options(1, 0)
    p = pb.parms()
    p.iNumSrfIterations = srf_iters
    prev_g = L.set_graphs(graphs)
    prev_a = L.set_autotune(autotune)
    prev_p = L.set_precision(precision)
    prev_d = L.set_drf_fft(drf)
    try:
        return L.Recon(p, pb.options(), pb.views, pb.drf_file, krnl, None, pb.mu, pb.prj, pb.x0)
    finally:
        L.set_graphs(prev_g)
        L.set_autotune(prev_a)
        L.set_precision(prev_p)
        L.set_drf_fft(prev_d)


# the spatial DRF kernel, and the 2-D FFT kernels (transform length 64)
@pytest.mark.parametrize("scatter,drf", [(False, "reference"), (True, "reference"),
                                         (False, "fft"), (True, "fft")])
def test_graph_replay_equals_eager(scatter, drf):
    print("graph replay vs eager, scatter=%s, drf=%s" % (scatter, drf), flush=True)
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False)
    krnl = _krnl_file(pb.pix) if scatter else None
    srf = 2 if scatter else 0
    a = _recon(pb, krnl, srf, graphs=False, drf=drf)
    b = _recon(pb, krnl, srf, graphs=True, drf=drf)
    if drf == "fft":
        assert a.drf_algorithms()[:2] == ("fft", "fft") == b.drf_algorithms()[:2]
    for it in range(3):
        for s in range(a.num_subsets):
            a.subiteration(s)
            b.subiteration(s)
            # every replay (from the second iteration on) against the eager result
            assert np.array_equal(a.estimate(), b.estimate()), (it, s)


@pytest.mark.parametrize("drf", ["reference", "fft"])
def test_graph_replay_with_learned_scatter_equals_eager(drf):
    print("graph replay vs eager, learned scatter (exact terms), drf=%s" % drf, flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False)
    krnl = _krnl_file(pb.pix)
    L.set_scatter_cnn_exact(True)
    L.set_scatter_cnn(None, check_every=3, max_nrmse=1.0)
    try:
        a = _recon(pb, krnl, 3, graphs=False, drf=drf)
        b = _recon(pb, krnl, 3, graphs=True, drf=drf)
    finally:
        L.set_scatter_cnn_exact(False)
        L.set_scatter_cnn(None, check_every=0, max_nrmse=0.01)   # the library defaults
    if drf == "fft":
        assert a.drf_algorithms()[:2] == ("fft", "fft") == b.drf_algorithms()[:2]
    for it in range(3):
        for s in range(a.num_subsets):
            a.subiteration(s)
            b.subiteration(s)
            assert np.array_equal(a.estimate(), b.estimate()), (it, s)
    sa, sb = L.cnn_status(a), L.cnn_status(b)
    assert sa == sb and sa["checks"] == 4 and sa["used"] == 12


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_autotuned_tiles_give_the_same_result(precision):
    print("autotune %s" % precision, flush=True)
    pb = Problem(n=32, nz=128, nviews=8, per=2, iters=1, half=5, noise=False)
    a = _recon(pb, precision=precision)
    b = _recon(pb, autotune=True, precision=precision)
    tp, tb = b.tiles()
    print("tiles", tp, tb)
    allowed = (2, 4, 8) if precision == "bf16" else (4, 8)
    assert tp in allowed and tb in allowed and a.tiles() == (0, 0)
    for s in range(a.num_subsets):
        a.subiteration(s)
        b.subiteration(s)
    assert np.array_equal(a.estimate(), b.estimate())


@pytest.mark.parametrize("precision,tiles", [("fp32", (4, 8)), ("bf16", (2, 4, 8))])
def test_every_drf_tile_gives_the_same_result(precision, tiles):
    """Each DRF tile shape set by hand (the autotuner's candidates) against the
    size heuristic, bit for bit, with graphs on (set_tiles drops the captured
    graphs)."""
    print("DRF tiles %s %s" % (precision, tiles), flush=True)
    pb = Problem(n=32, nz=128, nviews=8, per=2, iters=1, half=5, noise=False)
    ref = _recon(pb, precision=precision)
    for s in range(ref.num_subsets):
        ref.subiteration(s)
    want = ref.estimate()
    ref.close()
    for tp in tiles:
        for tb in tiles:
            r = _recon(pb, precision=precision, graphs=True)
            r.set_tiles(tp, tb)
            assert r.tiles() == (tp, tb)
            for s in range(r.num_subsets):
                r.subiteration(s)
            assert np.array_equal(r.estimate(), want), (tp, tb)
            r.close()
    r = _recon(pb, precision=precision)
    with pytest.raises(RuntimeError):
        r.set_tiles(3, 4)
    r.close()


@pytest.mark.parametrize("parent", [0, 1, 2])
def test_batch_of_uneven_reconstructions_equals_each_alone(parent):
    print("batch of 3 uneven reconstructions, launch mode %d" % parent, flush=True)
    import libirl_amd as L
    probs = [Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False, seed=1),
             Problem(n=48, nz=24, nviews=12, per=3, iters=3, half=3, noise=False, seed=2),
             Problem(n=40, nz=16, nviews=16, per=8, iters=3, half=3, noise=False, seed=3)]
    krnl = [_krnl_file(p.pix) for p in probs]
    alone = [_recon(p, k, 2) for p, k in zip(probs, krnl)]
    batched = [_recon(p, k, 2) for p, k in zip(probs, krnl)]
    S = [r.num_subsets for r in alone]
    assert S == [4, 4, 2]
    batch = L.Batch(batched, parent_graph=parent)
    steps = 3 * max(S)
    for k in range(steps):
        batch.step()
        for r, n in zip(alone, S):
            r.subiteration(k % n)
        if k % 3 == 2:
            batch.synchronize()
            for r, b in zip(alone, batched):
                assert np.array_equal(r.estimate(), b.estimate()), k
    batch.synchronize()
    for r, b in zip(alone, batched):
        assert np.array_equal(r.estimate(), b.estimate())
    st = batch.stats()
    print("batch stats", st)
    assert st["replays"] > 0 and st["eager_subiterations"] > 0
    assert (st["graph_launches"] > 0) == (parent > 0)
    with pytest.raises(RuntimeError):
        batch.set_parent_graph(3)
    batch.close()


def test_batch_with_a_shared_network_equals_each_alone():
    # every member shares one network's weights; each has its own inference
    # buffers, so members overlapping on their streams do not interfere
    print("batch of 3 with a shared scatter network", flush=True)
    import libirl_amd as L
    net = L.ScatterCnn.create(kappa=1.0, seed=3)
    p, _ = net.params()
    layers = L.unpack_cnn_params(p)
    layers[-1] = (layers[-1][0] * 0.01, layers[-1][1])
    net.set_params(L.pack_cnn_params(layers))
    probs = [Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False, seed=1),
             Problem(n=48, nz=24, nviews=12, per=3, iters=3, half=3, noise=False, seed=2),
             Problem(n=40, nz=16, nviews=16, per=8, iters=3, half=3, noise=False, seed=3)]
    krnl = [_krnl_file(p.pix) for p in probs]
    L.set_scatter_cnn(net, check_every=5, max_nrmse=1e9)
    try:
        alone = [_recon(p, k, 3) for p, k in zip(probs, krnl)]
        batched = [_recon(p, k, 3) for p, k in zip(probs, krnl)]
    finally:
        L.set_scatter_cnn(None)
    S = [r.num_subsets for r in alone]
    batch = L.Batch(batched)
    for k in range(3 * max(S)):
        batch.step()
        for r, n in zip(alone, S):
            r.subiteration(k % n)
            r.synchronize()      # alone: one at a time
    batch.synchronize()
    for r, b in zip(alone, batched):
        assert np.array_equal(r.estimate(), b.estimate())
        assert L.cnn_status(r) == L.cnn_status(b) and L.cnn_status(b)["active"]
    batch.close()


def test_genprj_with_set_up_timings_equals_plain():
    # IrlGenprj's forward projection has no subset table; the set-up timings
    # (tile autotuning, the timed FP32 DRF algorithm choice) run on its first
    # views and leave the projections as without them (up to the FFT
    # kernel's rounding when the timing picks it)
    print("IrlGenprj with autotune and the timed DRF choice", flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=1, half=3, noise=False)

    def genprj():
        out = np.zeros((pb.nviews, pb.nz, pb.n), np.float32)
        rc = L.irl_genprj(pb.parms(), pb.options(), pb.views, pb.drf_file, None, None,
                          pb.act.copy(), pb.mu.copy(), out)
        assert rc == 0, L.error_string()
        return out
    plain = genprj()
    prev = (L.set_autotune(True), L.set_drf_fft("auto"))
    try:
        timed = genprj()
    finally:
        L.set_autotune(prev[0])
        L.set_drf_fft(prev[1])
    assert np.abs(timed - plain).max() <= 1e-5 * np.abs(plain).max()


@pytest.mark.parametrize("learned", [False, True])
def test_graphs_prepared_at_setup_equal_eager(learned):
    """Every subset's graph captured before the first sub-iteration (for
    every scatter mode of the remaining iterations: modelled, modelled and
    stored, stored added), then replayed from the first occurrence on."""
    print("graphs prepared at setup vs eager, learned=%s" % learned, flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=3, half=3, noise=False)
    krnl = _krnl_file(pb.pix)
    if learned:
        L.set_scatter_cnn_exact(True)
        L.set_scatter_cnn(None, check_every=3, max_nrmse=1.0)
    try:
        a = _recon(pb, krnl, 2, graphs=False)
        b = _recon(pb, krnl, 2, graphs=True)
    finally:
        L.set_scatter_cnn_exact(False)
        L.set_scatter_cnn(None, check_every=0, max_nrmse=0.01)   # the library defaults
    b.prepare_graphs()
    for it in range(3):
        for s in range(a.num_subsets):
            a.subiteration(s)
            b.subiteration(s)
            assert np.array_equal(a.estimate(), b.estimate()), (it, s)
    if learned:
        assert L.cnn_status(a) == L.cnn_status(b)


def test_autotuned_network_row_length_gives_the_same_result():
    """The learned scatter's row length (slices per wave of the network's
    MFMA layers) timed at setup: any choice gives the same sources, so the
    reconstruction equals the untuned one bit for bit."""
    print("autotune with learned scatter", flush=True)
    import libirl_amd as L
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=2, half=3, noise=False)
    krnl = _krnl_file(pb.pix)
    net = L.ScatterCnn.create(kappa=1.0, seed=4)
    L.set_scatter_cnn(net, check_every=0, max_nrmse=1e9)
    try:
        a = _recon(pb, krnl, 2)
        b = _recon(pb, krnl, 2, autotune=True)
    finally:
        L.set_scatter_cnn(None, check_every=0, max_nrmse=0.01)   # the library defaults
    assert a.cnn_row_length() == 64 and b.cnn_row_length() in (64, 32, 16)
    print("row length", b.cnn_row_length())
    for it in range(2):
        for s in range(a.num_subsets):
            a.subiteration(s)
            b.subiteration(s)
    assert np.array_equal(a.estimate(), b.estimate())
    assert L.cnn_status(a)["active"] and L.cnn_status(b)["active"]
