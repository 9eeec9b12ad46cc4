"""OSEM sub-iteration throughput of libirl_amd on one MI355X.

Workload (the 128^3 x 120-view configuration): 128^3 fp32 voxels of 0.33 cm,
120 views over 360 deg on a seeded non-circular (ellipse) orbit with radius
20-28 cm, seeded ellipsoid phantom with 6-10 inserts and a lung region,
attenuation modelled, Gaussian distance-dependent DRF tabulated per plane with
31x31 kernels, data forward-projected on the GPU (IrlGenprj) and Poisson
sampled at 2M counts, 8 subsets x 15 views.

A step is one OSEM sub-iteration (15 views of projection, ratio,
backprojection, then the multiplicative update) with every input resident in
HBM.  The timed region cycles through the subsets in order.  The headline is
the FP32 path with its DRF convolution algorithm chosen by timing at setup
(the reference's spatial convolution or its 2-D FFT form); the spatial FP32
path is reported beside it (fp32_spatial) and is the FP32 reference of the
other lines' errors.

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "OSEM subiters/s, 128³×120 views: ESSE vs CNN-scatter at ≤1% NRMSE"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0


def build_workload(tmpdir, seed=2, name="c2", cfg=None):
    import libirl_amd as L
    from libirl_amd import synth
    if cfg is None:
        cfg = synth.CONFIGS[name]
    n, nz, pix, V = cfg["n"], cfg["nz"], cfg["pix"], cfg["views"]
    act, mu = synth.phantom(n, nz, seed, n_inserts=cfg["inserts"])
    views = synth.views_ellipse(V, n, seed, cfg.get("rmin", 20.0), cfg.get("rmax", 28.0))
    drf_file = os.path.join(tmpdir, name + ".drft")
    offs, psfs, start = synth.drf_for_orbit(drf_file, pix, n, views, cfg["half"])
    parms = L.IrlParms()
    parms.NumPixels, parms.NumSlices, parms.NumViews = n, nz, V
    parms.BinWidth = pix
    parms.fAtnScaleFac = pix
    parms.fScatEstFac = 1.0
    parms.NumAngPerSubset = cfg["vs"]
    parms.NumIterations = cfg["iters"]
    parms.SrfCollapseFac = 1
    opts = L.Options()
    opts.bModelDrf = 1
    opts.bReconIsInitEst = 0
    opts.iMsgLevel = 1
    krnl_file = None
    if "krnl_offs" in cfg:
        # ESSE kernel (33^3 support): Gaussian mixture, mu0 = 0.15/cm, scaled
        # to a 30% scatter fraction
        krnl_file = synth.esse_kernel_file(os.path.join(tmpdir, name + ".krnl"), cfg)
        parms.iNumSrfIterations = cfg["srf_iters"]
    prj = np.zeros((V, nz, n), np.float32)
    rc = L.irl_genprj(parms, opts, views, drf_file, krnl_file, None, act.copy(), mu.copy(), prj)
    if rc != 0:
        raise RuntimeError("IrlGenprj failed: " + L.error_string())
    data = synth.poisson(prj, cfg["counts"], seed)
    return dict(parms=parms, opts=opts, views=views, drf_file=drf_file, mu=mu, prj=data,
                act=act, offs=offs, psfs=psfs, start=start, cfg=cfg, krnl_file=krnl_file)


def run_parms(w, subiterations):
    """The workload's parameters with enough iterations for 'subiterations'
    sub-iterations, the scatter modelled in all of them where the workload
    models it in all of its own (each sub-iteration does the work of the
    workload's own iterations)."""
    parms = type(w["parms"]).from_buffer_copy(w["parms"])
    S0 = max(1, parms.NumViews // parms.NumAngPerSubset)
    iters = max(parms.NumIterations, -(-subiterations // S0))
    if parms.iNumSrfIterations >= parms.NumIterations:
        parms.iNumSrfIterations = iters
    parms.NumIterations = iters
    return parms


def time_subiterations(w, steps, warmup, precision="fp32", cnn=None, check_every=0,
                       drf="spatial", graphs=True, stage_steps=None):
    """Device-resident reconstruction of the workload: 'warmup' untimed
    sub-iterations, then exactly 'steps' timed ones, replayed from the
    library's captured sub-iteration graphs when 'graphs' (each subset's graph
    is captured at its second occurrence, so a warmup of two passes over the
    subsets keeps capture out of the timed steps), with no kernel timing.
    Then 'stage_steps' (default 'steps') more sub-iterations run eagerly with
    per-kernel HIP events on the engine's stream for the per-kernel times.
    With 'cnn' the scatter network replaces the per-view ESSE convolutions
    (checked against ESSE every 'check_every' subsets).  'drf' is the FP32
    DRF algorithm mode (set_drf_fft: "spatial" the reference's spatial
    convolution, "auto" the faster of it and the 2-D FFT, timed at setup).
    Returns (recon, seconds of the timed steps, {kernel: (ms_total,
    launches)} of the event pass, estimate after warmup + steps)."""
    import libirl_amd as L
    L.set_adv_options(0, 0)
    n = steps if stage_steps is None else stage_steps
    parms = run_parms(w, warmup + steps + n)
    prev = L.set_precision(precision)
    prev_drf = L.set_drf_fft(drf)
    prev_g = L.set_graphs(graphs)
    L.set_scatter_cnn(cnn, check_every=check_every, max_nrmse=0.01)
    try:
        rec = L.Recon(parms, w["opts"], w["views"], w["drf_file"], w["krnl_file"], None,
                      w["mu"], w["prj"], None)
    finally:
        L.set_precision(prev)
        L.set_drf_fft(prev_drf)
        L.set_graphs(prev_g)
        L.set_scatter_cnn(None)
    S = rec.num_subsets
    for k in range(warmup):
        rec.subiteration(k % S)
    rec.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        rec.subiteration((warmup + k) % S)
    rec.synchronize()
    dt = time.perf_counter() - t0
    est = rec.estimate()
    rec.reset_timing()
    rec.set_timing(True)
    for k in range(n):
        rec.subiteration((warmup + steps + k) % S)
    rec.synchronize()
    stages = rec.stage_times()
    rec.set_timing(False)
    return rec, dt, stages, est


def drf_flops_per_subset(w, recon):
    """Algorithmic flops of one subset's projector (or backprojector) DRF
    stage: 2 * Nz * N * taps for every (view, row) whose plane is convolved
    (distance > 0, prjbck.c:233)."""
    p = w["parms"]
    n, nz, pw = p.NumPixels, p.NumSlices, p.BinWidth
    taps = [(2 * ox - 1) * (2 * oy - 1) for ox, oy in w["offs"]]
    per_view = []
    for v in w["views"]:
        f = 0
        for r in range(n):
            dist = v.CFCR - pw * n / 2 + 0.5 * pw + r * pw
            if dist > 0:
                k = int(np.floor((dist - w["start"]) / pw + 0.5))
                k = min(max(k, 0), len(taps) - 1)
                f += 2 * nz * n * taps[k]
        per_view.append(f)
    S, Vs = recon.num_subsets, recon.views_per_subset
    sub = [[s // 2 + (s % 2) * S // 2 + m * S for m in range(Vs)] for s in range(S)]
    return [sum(per_view[a] for a in subset) for subset in sub]


def cpu_baseline(w, nviews=15):
    """Reference-semantics CPU path (oracle, FP32, 1 thread): Project +
    BackProject of the 15 views of one sub-iteration of the same workload
    (the update itself is negligible)."""
    from oracle import oracle as O
    p = w["parms"]
    n, nz, pw = p.NumPixels, p.NumSlices, p.BinWidth
    drf = O.DrfTable(w["start"], pw, w["offs"], w["psfs"])
    atn = O.reorder(nz, n, n, w["mu"] * pw, real="f32")
    x = np.full((n, nz, n), 1.0, np.float32)
    img = np.zeros((n, nz, n), np.float32)
    O.init_prj(nz, n, n, 3, real="f32")
    t0 = time.perf_counter()
    for v in w["views"][:nviews]:
        prj = O.project(nz, n, n, 3, pw, v.Angle, v.CFCR, drf, atn, x, real="f32")
        O.backproject(nz, n, n, 3, pw, v.Angle, v.CFCR, drf, atn, prj, img, atn_rotated=1,
                      real="f32")
    dt = time.perf_counter() - t0
    per_sub = dt / nviews * p.NumAngPerSubset
    return dict(value=1.0 / per_sub, unit="subiters/s", cores=1, kind="port",
                sample="Project+BackProject of %d of the 15 views of one sub-iteration "
                       "(%.1f s), scaled x%.2f" % (nviews, dt, p.NumAngPerSubset / nviews))


def fft_drf_bytes(w, rec, stage):
    """Algorithmic HBM bytes of one launch of the FFT DRF kernel of a stage:
    PRJ reads the subset's attenuated rotated planes and writes one partial
    projection per chunk of rows (the engine's FFT chunking: max(NRot/32,
    NRot/16) rows); BCK reads the ratio planes and, with attenuation, every
    row's factors, and writes every row's plane.  The PSF spectra are read
    from the on-chip caches (a few MB per table) and not counted."""
    p = w["parms"]
    n, nz, vs = p.NumPixels, p.NumSlices, rec.views_per_subset
    plane = 4 * n * nz
    if stage.startswith("drf_prj"):
        rchunk = max(max(1, n // 32), n // 16)
        nchunk = -(-n // rchunk)
        return vs * n * plane + vs * nchunk * plane
    atn = w["mu"] is not None
    return vs * plane + (vs * n * plane if atn else 0) + vs * n * plane


def drf_roofline(w, rec, stages, steps, first):
    """Dominant DRF stage and its roofline entry.  Spatial kernels: achieved
    algorithmic TFLOP/s (flops of the timed subsets / average launch time from
    the HIP events) against the FP32 or BF16 peak; FFT kernels: achieved
    algorithmic HBM GB/s (fft_drf_bytes / average launch time) against the
    HBM peak.  Returns (stage, roofline dict without peak/frac for spatial
    kernels' precision, which the caller fills)."""
    S = rec.num_subsets
    cands = [k for k in stages if k.startswith("drf_prj") or k.startswith("drf_bck")]
    dom = max(cands, key=lambda k: stages[k][0])
    dom_ms, dom_n = stages[dom]
    avg_s = dom_ms / dom_n * 1e-3
    if dom.endswith("_fft"):
        gbs = fft_drf_bytes(w, rec, dom) / avg_s / 1e9
        return dom, {"kernel": dom, "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS}
    flops = drf_flops_per_subset(w, rec)
    avg_flops = float(np.mean([flops[(first + k) % S] for k in range(steps)]))
    return dom, {"kernel": dom, "bound": "mfma", "achieved": avg_flops / avg_s / 1e12,
                 "unit": "TFLOP/s"}


DRF_NAMES = {"spatial": "spatial FP32 DRF", "auto": "FP32 DRF, faster of spatial and 2-D FFT"}


def roofline_line(w, rec, stages, steps, first, precision):
    """Roofline entry of the dominant DRF stage; the event pass covered
    'steps' sub-iterations from subset index 'first'."""
    dom, roof = drf_roofline(w, rec, stages, steps, first)
    if roof["bound"] == "mfma":
        roof["peak"] = PEAK_BF16_TFLOPS if precision != "fp32" else PEAK_FP32_TFLOPS
        roof["frac"] = roof["achieved"] / roof["peak"]
    return dom, roof


def workload_line(w, steps, warmup, precision, ref_estimate=None, cnn=None, check_every=0,
                  drf="spatial"):
    """Sub-object of the JSON line for one workload and DRF arithmetic."""
    import libirl_amd as L
    rec, dt, stages, est = time_subiterations(w, steps, warmup, precision, cnn, check_every, drf)
    dom, roof = roofline_line(w, rec, stages, steps, warmup + steps, precision)
    esse = w["krnl_file"] is not None
    name = "osem-128^3x120v-atn-drf31-%s8subsets" % (
        ("cnn-scatter-" if cnn is not None else "esse33-") if esse else "")
    bf16 = precision != "fp32"
    algo = rec.drf_algorithms()
    line = {"workload": name + (" (BF16 MFMA DRF, FP32 accumulation)" if bf16 else
                                " (%s)" % DRF_NAMES[drf]),
            "value": steps / dt, "unit": "subiters/s", "ms_per_step": dt * 1e3 / steps,
            "drf_arithmetic": (dom.split("_", 2)[2] if dom.count("_") >= 2 else "fp32")
            if bf16 else "fp32",
            "drf_algorithm": {"prj": algo[0], "bck": algo[1], "fft_length": algo[2]}
            if not bf16 else "bf16 mfma",
            "roofline": roof,
            "stages_ms_per_launch": {k: v[0] / max(v[1], 1) for k, v in stages.items()}}
    if ref_estimate is not None:
        line["errors"] = {"osem_rel_l2_vs_fp32_gpu_after_%d_subiterations" % (warmup + steps):
                          float(np.linalg.norm(est - ref_estimate) /
                                np.linalg.norm(ref_estimate))}
    else:
        line["_estimate"] = est
    if cnn is not None:
        line["cnn"] = L.cnn_status(rec)
        line["cnn"]["check_every_subsets"] = check_every
    rec.close()
    return line


def bf16_projection_error(w, tmpdir, nviews=2):
    """Forward projection (IrlGenprj) of two views of the workload with the
    BF16 DRF against the FP64 oracle: the largest per-view relative L2."""
    import libirl_amd as L
    from oracle import oracle as O
    p = w["parms"]
    n, nz, pw = p.NumPixels, p.NumSlices, p.BinWidth
    views = [w["views"][k] for k in (0, len(w["views"]) // 3)][:nviews]
    q = L.IrlParms()
    ctypes_copy(q, p)
    q.NumViews = len(views)
    q.NumAngPerSubset = len(views)
    q.iNumSrfIterations = 0
    prev = L.set_precision("bf16")
    try:
        prj = np.zeros((len(views), nz, n), np.float32)
        rc = L.irl_genprj(q, w["opts"], views, w["drf_file"], None, None, w["act"].copy(),
                          w["mu"].copy(), prj)
        if rc != 0:
            raise RuntimeError("IrlGenprj failed: " + L.error_string())
    finally:
        L.set_precision(prev)
    drf = O.DrfTable(w["start"], pw, w["offs"], w["psfs"])
    atn = O.reorder(nz, n, n, w["mu"] * pw)
    x = O.reorder(nz, n, n, w["act"])
    O.init_prj(nz, n, n, 3)
    errs = []
    for k, v in enumerate(views):
        ref = O.project(nz, n, n, 3, pw, v.Angle, v.CFCR, drf, atn, x, fast=1)
        ref = ref * np.float32(1.0 / np.float32(len(views)))
        errs.append(float(np.linalg.norm(prj[k] - ref) / np.linalg.norm(ref)))
    return max(errs)


def ctypes_copy(dst, src):
    for f, _ in src._fields_:
        setattr(dst, f, getattr(src, f))


def measured_traffic(stage):
    """HBM bytes per launch of a stage's kernel, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload (counters
    scaled by the copy-kernel calibration), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "traffic_c2.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][t["stages"][stage]]["hbm_bytes_per_dispatch"]
    except (OSError, KeyError, ValueError):
        return None


def batch_configs(count=16, seed=5):
    """Uneven reconstructions of the batched workload: N in {64, ..., 160}
    (even multiples of 16), V a multiple of 8 in [64, 128] with 8 subsets,
    the ESSE scatter configuration's phantom class, DRF and kernel at 0.33 cm,
    orbits just outside each field of view."""
    from libirl_amd import synth
    rng = np.random.default_rng(seed)
    base = synth.CONFIGS["c3"]
    out = []
    for i in range(count):
        n = int(rng.choice(np.arange(64, 161, 16)))
        V = int(rng.choice(np.arange(64, 129, 8)))
        r0 = 0.5 * n * base["pix"] + 2.0
        out.append(dict(base, n=n, nz=n, views=V, vs=V // 8, rmin=r0, rmax=r0 + 8.0,
                        counts=2e6 * (n / 128.0) ** 2))
    return out


def batch_line(tmpdir, net, steps=16, warmup=24):
    """16 concurrent reconstructions of uneven size: learned scatter + BF16
    DRF, graph replay and autotuned tiles, stepped through one batch (each
    member's graph on its own stream, and one parent-graph launch per step)
    and, for comparison, one after another; then each estimate after 5
    iterations against its FP32 ESSE reconstruction."""
    import libirl_amd as L
    cfgs = batch_configs()
    works = [build_workload(tmpdir, seed=100 + i, name="b%d" % i, cfg=c)
             for i, c in enumerate(cfgs)]
    L.set_adv_options(0, 0)
    prev = (L.set_precision("bf16"), L.set_graphs(True), L.set_autotune(True))
    L.set_scatter_cnn(net, check_every=8, max_nrmse=0.01)
    try:
        # iterations for every step below (warmup, the timed runs, the
        # parent-graph warmup, one at a time), the scatter modelled in all
        total = warmup + 5 * steps + max(c["views"] // c["vs"] for c in cfgs)
        recs = [L.Recon(run_parms(w, total), w["opts"], w["views"], w["drf_file"],
                        w["krnl_file"], None, w["mu"], w["prj"], None) for w in works]
    finally:
        L.set_precision(prev[0])
        L.set_graphs(prev[1])
        L.set_autotune(prev[2])
        L.set_scatter_cnn(None)
    batch = L.Batch(recs)

    def timed(parent):
        batch.set_parent_graph(parent)
        batch.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.step()
        batch.synchronize()
        return time.perf_counter() - t0

    for _ in range(warmup):
        batch.step()
    dts = [timed(False)]
    st = batch.stats()
    # warmup: every member's subsets run eagerly once, then are captured and
    # replayed once, so the timed steps replay warm graph executables
    est = [r.estimate() for r in recs]          # after warmup + steps = 5 iterations
    tiles = [r.tiles() for r in recs]
    status = [L.cnn_status(r) for r in recs]
    # the same steps as one parent-graph launch per step (its executables
    # instantiated first), each mode timed twice, alternating
    S = [r.num_subsets for r in recs]
    batch.set_parent_graph(True)
    for _ in range(max(S)):
        batch.step()
    dtp = [timed(True)]
    dts.append(timed(False))
    dtp.append(timed(True))
    dt_batch, dt_parent = min(dts), min(dtp)
    pst = batch.stats()
    batch.set_parent_graph(False)
    # one reconstruction at a time, each sub-iteration finished before the next
    t0 = time.perf_counter()
    for k in range(steps):
        for r, n in zip(recs, S):
            r.subiteration(k % n)
            r.synchronize()
    dt_seq = time.perf_counter() - t0
    batch.close()
    for r in recs:
        r.close()
    errs = []
    for w, e in zip(works, est):
        rec, _, _, ref = time_subiterations(w, warmup + steps, 0, "fp32", stage_steps=0)
        rec.close()
        errs.append(float(np.linalg.norm(e - ref) / np.linalg.norm(ref)))
    n = len(recs)
    return {"workload": "16 reconstructions, N in 64..160, 64..128 views, 8 subsets, ATN + DRF + "
                        "learned scatter, BF16 DRF",
            "value": n * steps / dt_batch, "unit": "subiters/s (all reconstructions)",
            "ms_per_step": dt_batch * 1e3 / steps,
            "mode": "each member's sub-iteration graph launched on its own stream (best of 2 timed runs)",
            "member_streams_runs": [n * steps / t for t in dts],
            "parent_graph_value": n * steps / dt_parent,
            "parent_graph_runs": [n * steps / t for t in dtp],
            "parent_graph_launches": pst["graph_launches"],
            "one_at_a_time_value": n * steps / dt_seq,
            "batch_speedup_vs_one_at_a_time": dt_seq / dt_batch,
            "sizes": [[c["n"], c["views"]] for c in cfgs],
            "tiles_prj_bck": tiles,
            "replays": st["replays"],
            "eager_subiterations": st["eager_subiterations"],
            "learned_scatter_active": sum(s["active"] for s in status),
            "worst_check_nrmse": max(s["worst_nrmse"] for s in status),
            "rel_l2_vs_fp32_esse_after_5_iterations": {"max": max(errs),
                                                        "mean": float(np.mean(errs))}}


CNN_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libirl_amd", "weights",
                           "scatter_cnn_c3.bin")


def load_scatter_cnn():
    """The trained scatter network of the 128^3 ESSE workload, or None."""
    import libirl_amd as L
    if not os.path.exists(CNN_WEIGHTS):
        return None
    return L.ScatterCnn.load_file(CNN_WEIGHTS)


def cnn_training_record():
    try:
        with open(CNN_WEIGHTS + ".json") as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-esse", action="store_true", help="skip the ESSE workload line")
    ap.add_argument("--only-esse", action="store_true",
                    help="time only the ESSE workload (for profiling runs)")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16 precision variant")
    ap.add_argument("--no-cnn", action="store_true", help="skip the learned-scatter lines")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batch of 16 uneven reconstructions")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batch of 16 uneven reconstructions")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the FP64-oracle projection error of the BF16 variant")
    ap.add_argument("--drf", default="auto", choices=["auto", "spatial", "fft"],
                    help="FP32 DRF algorithm of the --only-esse line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of this build "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    tmpdir = tempfile.mkdtemp(prefix="irlbench")
    if args.only_batch:
        print(json.dumps(batch_line(tmpdir, load_scatter_cnn())))
        return
    if args.only_esse:
        we = build_workload(tmpdir, name="c3")
        line = workload_line(we, args.steps, args.warmup, "fp32", drf=args.drf)
        line.pop("_estimate", None)
        print(json.dumps(line))
        return
    w = build_workload(tmpdir)
    # the reference's spatial DRF convolution first: its estimate is the
    # FP32 reference of the other lines
    spatial = workload_line(w, args.steps, args.warmup, "fp32", drf="spatial")
    est32 = spatial.pop("_estimate")
    spatial["roofline"]["traffic"] = measured_traffic(spatial["roofline"]["kernel"])
    # headline: the library's FP32 path with the DRF algorithm timed at setup
    rec, dt, stages, est = time_subiterations(w, args.steps, args.warmup, drf="auto")
    S = rec.num_subsets
    ms_step = dt * 1e3 / args.steps
    dom, roof = roofline_line(w, rec, stages, args.steps, args.warmup + args.steps, "fp32")
    roof["traffic"] = measured_traffic(dom)
    algo = rec.drf_algorithms()
    out = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "subiters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded phantom, GPU forward projection, Poisson 2M counts)",
        "config": {"workload": "osem-128^3x120v-atn-drf31-8subsets (plain FP32, no scatter)",
                   "N": 128, "Nz": 128, "views": 120, "subsets": S,
                   "views_per_subset": rec.views_per_subset, "drf": "31x31 Gaussian",
                   "drf_algorithm": {"prj": algo[0], "bck": algo[1], "fft_length": algo[2]},
                   "parallelism": "single GPU"},
        "roofline": roof,
        "stages_ms_per_launch": {k: v[0] / max(v[1], 1) for k, v in stages.items()},
        "errors": {"osem_rel_l2_vs_spatial_fp32_after_%d_subiterations" % (
            args.warmup + args.steps): float(np.linalg.norm(est - est32) / np.linalg.norm(est32))},
        "fp32_spatial": spatial,
    }
    rec.close()
    if not args.no_bf16:
        # Precision variant of the same workload: BF16 MFMA DRF convolution.
        line = workload_line(w, args.steps, args.warmup, "bf16", ref_estimate=est32)
        if not args.no_errors:
            line["errors"]["projection_rel_l2_vs_fp64_oracle"] = bf16_projection_error(w, tmpdir)
        # the two fixed BF16 modes beside the per-plane choice
        line["modes"] = {}
        for mode in ("bf16-plain", "bf16x3"):
            alt = workload_line(w, args.steps, args.warmup, mode, ref_estimate=est32)
            line["modes"][mode] = {"value": alt["value"], "drf_ms_per_launch": {
                k: v for k, v in alt["stages_ms_per_launch"].items() if k.startswith("drf")},
                "errors": alt["errors"]}
        out["bf16"] = line
    if not args.no_esse:
        # Same geometry with ESSE scatter modelled in the projector (the
        # plain-method scatter configuration), FP32 and with the BF16 DRF.
        we = build_workload(tmpdir, name="c3")
        K, W = args.steps, args.warmup
        # the plain method with the reference's spatial DRF (the FP32
        # reference of the other lines), then with the timed DRF choice
        out["esse_spatial"] = workload_line(we, K, W, "fp32", drf="spatial")
        ref = out["esse_spatial"].pop("_estimate")
        out["esse"] = workload_line(we, K, W, "fp32", ref_estimate=ref, drf="auto")
        if not args.no_bf16:
            out["esse_bf16"] = workload_line(we, K, W, "bf16", ref_estimate=ref)
        net = load_scatter_cnn()
        if net is not None and not args.no_cnn:
            # Learned scatter + BF16 DRF (and + FP32 DRF), checked against
            # ESSE every 8 subsets
            for prec, key in (("bf16", "cnn_bf16"), ("fp32", "cnn_fp32")):
                line = workload_line(we, K, W, prec, ref_estimate=ref, cnn=net, check_every=8,
                                     drf="auto")
                line["speedup_vs_esse_fp32"] = line["value"] / out["esse"]["value"]
                line["speedup_vs_esse_fp32_spatial"] = (line["value"] /
                                                        out["esse_spatial"]["value"])
                line["training"] = cnn_training_record()
                out[key] = line
            if not args.no_batch:
                out["batch16"] = batch_line(tmpdir, net)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
