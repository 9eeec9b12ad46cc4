"""OSEM sub-iteration throughput of libirl_amd on one MI355X.

Workload (128^3 x 120 views, attenuation + DRF + ESSE scatter, 8 subsets):
128^3 fp32 voxels of 0.33 cm, 120 views over 360 deg on a seeded
non-circular (ellipse) orbit with radius 20-28 cm, seeded ellipsoid phantom
with 6-10 inserts and a lung region, attenuation modelled, Gaussian
distance-dependent DRF tabulated per plane with 31x31 kernels, ESSE scatter
from a 17 x 33 Gaussian-mixture kernel (mu0 = 0.15/cm, 30% scatter
fraction), data = primary + scatter forward-projected on the GPU (IrlGenprj)
and Poisson sampled at 2M counts, 8 subsets x 15 views.

A step is one OSEM sub-iteration (15 views of scatter source, attenuated
projection, ratio, backprojection, then the multiplicative update) with every
input resident in HBM.  The timed region cycles through the subsets in
order, each replayed from its graph (every graph is captured at setup).

value: the learned-scatter form (a 3-D CNN gives the scatter source terms
once per sub-iteration, checked against ESSE once per pass over the subsets)
with the BF16 MFMA DRF.  plain_fp32_esse: the library's fastest plain FP32
ESSE path (ESSE through the 2-D plane FFT, FP32 DRF with its algorithm timed
at setup) on the same workload; speedup_vs_plain_fp32_esse = value / its
value.  Other keys: the spatial FP32 ESSE path, the learned form with the
FP32 DRF, ESSE with the BF16 DRF, the workload without scatter, and the batch
of 16 uneven reconstructions.

Prints one JSON line.
"""
import argparse
import csv
import json
import os
import shutil
import subprocess
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_c3_cnn_bf16.json")


def build_workload(tmpdir, seed=2, name="c3", cfg=None):
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
    """Device-resident reconstruction of the workload (tile shapes of the
    DRF kernels and the network's row length timed at setup by the library's
    autotuner): every subset's sub-iteration graph captured at setup (graphs
    on), 'warmup' untimed
    sub-iterations, then exactly 'steps' timed ones replayed from the graphs,
    with no kernel timing.  Then 'stage_steps' (default 'steps') more
    sub-iterations run eagerly with per-kernel HIP events on the engine's
    stream for the per-kernel times.  With 'cnn' the scatter network replaces
    the per-view ESSE convolutions (checked against ESSE every 'check_every'
    subsets; a checked subset runs eagerly).  'drf' is the FP32 DRF algorithm
    mode (set_drf_fft: "spatial" the reference's spatial convolution, "auto"
    the faster of it and the 2-D FFT, timed at setup).  Returns (recon,
    seconds of the timed steps, {kernel: (ms_total, launches)} of the event
    pass, estimate after warmup + steps)."""
    import libirl_amd as L
    L.set_adv_options(0, 0)
    n = steps if stage_steps is None else stage_steps
    parms = run_parms(w, warmup + steps + n)
    prev = L.set_precision(precision)
    prev_drf = L.set_drf_fft(drf)
    prev_g = L.set_graphs(graphs)
    prev_t = L.set_autotune(True)
    L.set_scatter_cnn(cnn, check_every=check_every, max_nrmse=0.01)
    try:
        rec = L.Recon(parms, w["opts"], w["views"], w["drf_file"], w["krnl_file"], None,
                      w["mu"], w["prj"], None)
    finally:
        L.set_precision(prev)
        L.set_drf_fft(prev_drf)
        L.set_graphs(prev_g)
        L.set_autotune(prev_t)
        L.set_scatter_cnn(None)
    if graphs:
        rec.prepare_graphs()
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


# ---------------------------------------------------------------------------
# Algorithmic work per launch of the single-kernel stages
# ---------------------------------------------------------------------------

def row_entries(w):
    """[view][row] DRF table entry of every row of the rotated volume, -1
    where the row is not convolved (distance <= 0, prjbck.c:233): the entry
    nearest the row's distance from the collimator face."""
    p = w["parms"]
    n, pw = p.NumPixels, p.BinWidth
    last = len(w["offs"]) - 1
    out = np.full((len(w["views"]), n), -1, np.int64)
    for a, v in enumerate(w["views"]):
        for r in range(n):
            dist = v.CFCR - pw * n / 2 + 0.5 * pw + r * pw
            if dist > 0:
                k = int(np.floor((dist - w["start"]) / pw + 0.5))
                out[a, r] = min(max(k, 0), last)
    return out


def subset_views(rec):
    """The views of every subset, in the order the reconstruction takes them
    (the reference's interleaved subset order)."""
    S, Vs = rec.num_subsets, rec.views_per_subset
    return [[s // 2 + (s % 2) * S // 2 + m * S for m in range(Vs)] for s in range(S)]


def drf_flops_per_subset(w, rec):
    """Algorithmic flops of one subset's projector (or backprojector) DRF
    stage: 2 * Nz * N * taps for every (view, row) whose plane is convolved."""
    p = w["parms"]
    n, nz = p.NumPixels, p.NumSlices
    taps = np.array([(2 * ox - 1) * (2 * oy - 1) for ox, oy in w["offs"]], np.int64)
    ent = row_entries(w)
    per_view = [int(sum(2 * nz * n * taps[k] for k in row if k >= 0)) for row in ent]
    return [sum(per_view[a] for a in sub) for sub in subset_views(rec)]


def att_write_bytes(w, rec, bf16):
    """Bytes one subset's projector-input planes take: 4 per voxel in FP32;
    with the BF16 DRF, 2 for rows whose DRF entry runs plain or weight-split
    BF16 and 4 for data-split and unconvolved rows."""
    p = w["parms"]
    plane = p.NumPixels * p.NumSlices
    ent = row_entries(w)
    if not bf16:
        per_view = [4 * plane * p.NumPixels] * len(w["views"])
    else:
        modes = rec.bf16_planes()[0]
        per_view = [sum(plane * (2 if (k >= 0 and modes[k] <= 1) else 4) for k in row)
                    for row in ent]
    return [sum(per_view[a] for a in sub) for sub in subset_views(rec)]


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
    return vs * plane + vs * n * plane + vs * n * plane


# Stages that launch one kernel each, and that kernel.  The others (cnn_src:
# pooling, mean, four convolution layers and the source build; esse_conv*:
# the plane transforms) launch several and have no single roofline.
SINGLE_KERNEL = {"cnn_att": "k_att_src", "att": "k_att", "rot_x": "k_att", "ratio": "k_ratio",
                 "bck_update": "k_bck"}


def stage_kernel(stage):
    if stage in SINGLE_KERNEL:
        return SINGLE_KERNEL[stage]
    if stage.startswith("drf_prj") or stage.startswith("drf_bck"):
        if "_bf16" in stage:
            return "k_drf_bf16"
        return "k_drf_fft" if stage.endswith("_fft") else "k_drf"
    return None


def stage_work(w, rec, stage, subsets, bf16):
    """(bound, algorithmic amount per launch averaged over 'subsets', unit)
    of a single-kernel stage: bytes for the HBM-bound kernels, flops for the
    spatial DRF convolutions.  M = N^2 Nz voxels, Vs views per subset."""
    p = w["parms"]
    M = p.NumPixels * p.NumPixels * p.NumSlices
    Vs = rec.views_per_subset
    P = p.NumPixels * p.NumSlices
    mean = lambda per: float(np.mean([per[s] for s in subsets]))
    if stage == "cnn_att":
        # afac of the subset's views, the two source terms once (FP32, or
        # with the BF16 DRF one bf16 pair per voxel), the planes
        src = 4 * M if bf16 else 2 * 4 * M
        return "hbm", 4 * Vs * M + src + mean(att_write_bytes(w, rec, bf16)), "B"
    if stage == "att":
        return "hbm", 4 * Vs * M + 4 * M + mean(att_write_bytes(w, rec, bf16)), "B"
    if stage == "rot_x":
        return "hbm", 4 * M + 4 * Vs * M, "B"
    if stage == "ratio":
        # data, the projection partials, the ratio planes
        return "hbm", 3 * 4 * Vs * P, "B"
    if stage == "bck_update":
        # the subset's backprojected planes, the estimate, the sensitivity
        # inverse; the estimate written
        return "hbm", 4 * Vs * M + 3 * 4 * M, "B"
    if stage.endswith("_fft"):
        return "hbm", fft_drf_bytes(w, rec, stage), "B"
    return "mfma", mean(drf_flops_per_subset(w, rec)), "flop"


def roofline(w, rec, stages, steps, first, precision, traffic=None):
    """Roofline entry of the line's dominant single-kernel stage (largest
    total event time of the event pass, which covered 'steps' sub-iterations
    from subset index 'first'): achieved = algorithmic bytes (or flops) per
    launch / average launch time."""
    S = rec.num_subsets
    subsets = [(first + k) % S for k in range(steps)]
    cands = [k for k in stages if stage_kernel(k) is not None]
    dom = max(cands, key=lambda k: stages[k][0])
    ms, nl = stages[dom]
    avg_s = ms / nl * 1e-3
    bf16 = precision != "fp32"
    bound, amount, unit = stage_work(w, rec, dom, subsets, bf16)
    if bound == "hbm":
        ach, peak, u = amount / avg_s / 1e9, PEAK_HBM_GBS, "GB/s"
    else:
        ach = amount / avg_s / 1e12
        peak, u = (PEAK_BF16_TFLOPS if bf16 else PEAK_FP32_TFLOPS), "TFLOP/s"
    out = {"kernel": stage_kernel(dom), "stage": dom, "bound": bound, "achieved": ach,
           "peak": peak, "unit": u, "frac": ach / peak, "avg_launch_ms": ms / nl,
           "algorithmic_per_launch": amount, "algorithmic_unit": unit,
           "traffic": traffic_for(traffic, dom, rec.tiles())}
    return out


def traffic_for(traffic, stage, tiles=None):
    """Measured HBM bytes per dispatch of a stage's kernel from a traffic
    summary (kernel names as rocprofv3 prints them: k_drf_bf16<KY, NZT, BCK>,
    k_bck<NORM, SHEAR>, k_att_src<SHEAR, PK>), or None.  For the DRF kernels
    the instantiation of the tile in use ('tiles' = (projector,
    backprojector)); the others in the summary are the autotuner's timings."""
    if traffic is None:
        return None
    kern = stage_kernel(stage)
    names = [k for k in traffic["kernels"] if k == kern or k.startswith(kern + "<")]
    if kern in ("k_drf_bf16", "k_drf"):
        bck = stage.startswith("drf_bck")
        want = "true" if bck else "false"
        names = [k for k in names if k.split("<")[1].split(",")[2].strip(" >") == want]
        if tiles is not None and len(names) > 1:
            t = str(tiles[1] if bck else tiles[0])
            names = [k for k in names if k.split("<")[1].split(",")[1].strip() == t]
    elif kern == "k_bck":
        names = [k for k in names if k.startswith("k_bck<false")]
    if len(names) != 1:
        return None
    return traffic["kernels"][names[0]]["hbm_bytes_per_dispatch"]


def load_traffic():
    """HBM bytes per launch of the headline's kernels from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (counters scaled by the probe
    calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


DRF_NAMES = {"spatial": "spatial FP32 DRF", "auto": "FP32 DRF, faster of spatial and 2-D FFT"}


def workload_line(w, steps, warmup, precision, ref_estimate=None, cnn=None, check_every=0,
                  drf="spatial", traffic=None):
    """Sub-object of the JSON line for one workload, scatter form and DRF
    arithmetic.  The estimate goes under '_estimate' (for the caller) unless
    'ref_estimate' is given (then its relative L2 against it is reported)."""
    import libirl_amd as L
    rec, dt, stages, est = time_subiterations(w, steps, warmup, precision, cnn, check_every, drf)
    roof = roofline(w, rec, stages, steps, warmup + steps, precision, traffic)
    esse = w["krnl_file"] is not None
    name = "osem-%d^3x%dv-atn-drf31-%s8subsets" % (
        w["parms"].NumPixels, w["parms"].NumViews,
        ("cnn-scatter-" if cnn is not None else "esse33-") if esse else "")
    bf16 = precision != "fp32"
    algo = rec.drf_algorithms()
    line = {"workload": name + (" (BF16 MFMA DRF, FP32 accumulation)" if bf16 else
                                " (%s)" % DRF_NAMES[drf]),
            "tiles_prj_bck_cnnrow": list(rec.tiles()) + [rec.cnn_row_length()],
            "prj_rows_per_chunk": rec.prj_rows_per_chunk(),
            "value": steps / dt, "unit": "subiters/s", "ms_per_step": dt * 1e3 / steps,
            "drf_algorithm": {"prj": algo[0], "bck": algo[1], "fft_length": algo[2]}
            if not bf16 else "bf16 mfma",
            "roofline": roof,
            "stages_ms_per_launch": {k: v[0] / max(v[1], 1) for k, v in stages.items()},
            "stages_ms_per_subiteration": {k: v[0] / steps for k, v in stages.items()}}
    if bf16:
        modes = rec.bf16_planes()[0]
        line["bf16_entries_by_mode"] = {"plain": int((modes[:-1] == 0).sum()),
                                        "weights_hi_lo": int((modes[:-1] == 1).sum()),
                                        "weights_and_data_hi_lo": int((modes[:-1] == 2).sum())}
    if ref_estimate is not None:
        line["errors"] = {"recon_rel_l2_vs_plain_fp32_after_%d_subiterations" % (warmup + steps):
                          float(np.linalg.norm(est - ref_estimate) /
                                np.linalg.norm(ref_estimate))}
    else:
        line["_estimate"] = est
    if cnn is not None:
        line["cnn"] = L.cnn_status(rec)
        line["cnn"]["check_every_subsets"] = check_every
    rec.close()
    return line


def bf16_projection_error(w, nviews=2):
    """Forward projection (IrlGenprj, attenuation + DRF, no scatter) of two
    views of the workload with the BF16 DRF against the FP64 oracle: the
    largest per-view relative L2."""
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


# ---------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------

def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(views=2, nproc=8):
    """Reference-semantics CPU path on the same workload, one thread per
    process: per view the estimate and map rotated, the ESSE effective
    source (FFT convolutions on the model's padded grid, float32), the FP32
    attenuated DRF Project and BackProject (the oracle's restatement of
    prjbck.c).  'views' views of subset 0 are timed and scaled to the
    subset's 15 (the update is negligible).  One process on one core pinned
    with taskset, then 'nproc' processes at once, each pinned to its own
    core, for the host's throughput."""
    cores = sorted(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1",
               MKL_NUM_THREADS="1")
    taskset = shutil.which("taskset")

    def cmd(core):
        base = [sys.executable, "-m", "oracle.cpu_baseline", "--views", str(views)]
        return ["taskset", "-c", str(core)] + base if taskset else base + ["--core", str(core)]

    def result(out):
        return json.loads(out.strip().splitlines()[-1])

    one = result(subprocess.run(cmd(cores[-1]), cwd=ROOT, env=env, capture_output=True,
                                text=True, check=True).stdout)
    n = max(1, min(nproc, len(cores)))
    procs = [subprocess.Popen(cmd(c), cwd=ROOT, env=env, stdout=subprocess.PIPE, text=True)
             for c in cores[-n:]]
    outs = [p.communicate()[0] for p in procs]
    if any(p.returncode != 0 for p in procs):
        raise RuntimeError("CPU baseline process failed")
    per = [result(o)["s_per_view"] for o in outs]
    vs = 15
    return {"value": 1.0 / (one["s_per_view"] * vs), "unit": "subiters/s", "cores": 1,
            "kind": "port", "cpu_model": cpu_model(),
            "pinning": "taskset -c %d" % cores[-1] if taskset else "sched_setaffinity",
            "sample": "%d views of subset 0 of the same 128^3 x 120-view workload "
                      "(rotation of estimate and map, ESSE effective source by float32 FFT "
                      "convolution, FP32 attenuated DRF Project + BackProject, one thread), "
                      "%.1f s per view (%.1f s Project+BackProject, %.1f s scatter model), "
                      "scaled x%.1f to a sub-iteration" % (
                          views, one["s_per_view"], one["s_per_view_project_backproject"],
                          one["s_per_view_scatter_model"], vs / views),
            "n_process": {"processes": n, "cores": n, "value": sum(1.0 / (s * vs) for s in per),
                          "unit": "subiters/s (all processes)",
                          "s_per_view": [round(s, 3) for s in per]}}


# ---------------------------------------------------------------------------
# Batch of uneven reconstructions
# ---------------------------------------------------------------------------

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
    member's graph on its own stream, one parent-graph launch per step, and
    one launch per step of a graph with the members' launches as branches)
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
        total = warmup + 7 * steps + 2 * max(c["views"] // c["vs"] for c in cfgs)
        recs = [L.Recon(run_parms(w, total), w["opts"], w["views"], w["drf_file"],
                        w["krnl_file"], None, w["mu"], w["prj"], None) for w in works]
    finally:
        L.set_precision(prev[0])
        L.set_graphs(prev[1])
        L.set_autotune(prev[2])
        L.set_scatter_cnn(None)
    # warmup and the first timed run on member streams (executables of each
    # mode are made before that mode is timed)
    batch = L.Batch(recs, parent_graph=0)

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
    dts = [timed(0)]
    st = batch.stats()
    # warmup: every member's subsets run eagerly once, then are captured and
    # replayed once, so the timed steps replay warm graph executables
    est = [r.estimate() for r in recs]          # after warmup + steps = 5 iterations
    tiles = [r.tiles() + (r.cnn_row_length(),) for r in recs]
    rows_per_chunk = [r.prj_rows_per_chunk() for r in recs]
    status = [L.cnn_status(r) for r in recs]
    # the same steps as one parent-graph launch per step (its executables
    # instantiated first), each mode timed twice, alternating
    S = [r.num_subsets for r in recs]
    batch.set_parent_graph(1)
    for _ in range(max(S)):
        batch.step()
    dtp = [timed(1)]
    dts.append(timed(0))
    dtp.append(timed(1))
    pst = batch.stats()
    # one graph per step with the members' launches as parallel branches
    batch.set_parent_graph(2)
    for _ in range(max(S)):
        batch.step()
    dtf = [timed(2), timed(2)]
    batch_launches = batch.stats()["graph_launches"]
    dt_streams, dt_parent, dt_flat = min(dts), min(dtp), min(dtf)
    batch.set_parent_graph(0)
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
            "value": n * steps / dt_flat, "unit": "subiters/s (all reconstructions)",
            "ms_per_step": dt_flat * 1e3 / steps,
            "mode": "one graph launch per step, the members' launches as parallel branches "
                    "(the library's default; best of 2 timed runs)",
            "one_graph_runs": [n * steps / t for t in dtf],
            "member_streams_value": n * steps / dt_streams,
            "member_streams_runs": [n * steps / t for t in dts],
            "parent_graph_value": n * steps / dt_parent,
            "parent_graph_runs": [n * steps / t for t in dtp],
            "graph_launches": batch_launches,
            "one_at_a_time_value": n * steps / dt_seq,
            "batch_speedup_vs_one_at_a_time": dt_seq / dt_flat,
            "sizes": [[c["n"], c["views"]] for c in cfgs],
            "tiles_prj_bck_cnnrow": tiles,
            "prj_rows_per_chunk": rows_per_chunk,
            "replays": st["replays"],
            "eager_subiterations": st["eager_subiterations"],
            "learned_scatter_active": sum(s["active"] for s in status),
            "worst_check_nrmse": max(s["worst_nrmse"] for s in status),
            "rel_l2_vs_fp32_esse_after_5_iterations": {"max": max(errs),
                                                        "mean": float(np.mean(errs))}}


CNN_WEIGHTS = os.path.join(ROOT, "libirl_amd", "weights", "scatter_cnn_c3.bin")


def load_scatter_cnn():
    """The trained scatter network of the 128^3 ESSE workload."""
    import libirl_amd as L
    if not os.path.exists(CNN_WEIGHTS):
        raise RuntimeError("scatter network weights missing: " + CNN_WEIGHTS)
    return L.ScatterCnn.load_file(CNN_WEIGHTS)


def cnn_training_record():
    try:
        with open(CNN_WEIGHTS + ".json") as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


# ---------------------------------------------------------------------------
# Counter calibration probes and profile summaries
# ---------------------------------------------------------------------------

def probe(kind, mib):
    """A plain device copy ('copy': reads and writes 'mib' MiB) or a
    read-only sum ('read': reads 'mib' MiB, writes a few bytes), three
    launches each, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda")
    if kind == "copy":
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
    else:
        for _ in range(3):
            s = a.sum()
        s.item()
    torch.cuda.synchronize()
    print(json.dumps({"probe": kind, "bytes": n * 4}))


def read_counter_csv(path, counter):
    """{kernel: [value per dispatch]} of one counter from a rocprofv3
    counter_collection.csv."""
    out = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            name = row["Kernel_Name"].split("(")[0]
            name = name.replace("void ", "").replace("irlamd::", "")
            out.setdefault(name, []).append(float(row["Counter_Value"]))
    return out


def traffic_summary(fetch_csv, write_csv, probe_fetch_csv, probe_write_csv, probe_bytes):
    """HBM bytes per dispatch of every kernel from FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch), each scaled by what that counter reads for a probe of
    known size (the read-only sum for FETCH_SIZE, the copy for WRITE_SIZE)."""
    pf = read_counter_csv(probe_fetch_csv, "FETCH_SIZE")
    pw = read_counter_csv(probe_write_csv, "WRITE_SIZE")
    # the probe's largest-counter kernel is the probe itself
    f_kib = max(np.mean(v) for v in pf.values())
    w_kib = max(np.mean(v) for v in pw.values())
    fs = probe_bytes / (f_kib * 1024.0)
    ws = probe_bytes / (w_kib * 1024.0)
    fe = read_counter_csv(fetch_csv, "FETCH_SIZE")
    wr = read_counter_csv(write_csv, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(fe) & set(wr)):
        f, wv = float(np.mean(fe[k])), float(np.mean(wr[k]))
        kernels[k] = {"fetch_kib": f, "write_kib": wv,
                      "hbm_bytes_per_dispatch": f * 1024.0 * fs + wv * 1024.0 * ws}
    return {"calibration": {"fetch_scale": fs, "write_scale": ws,
                            "probe_fetch_kib": f_kib, "probe_write_kib": w_kib,
                            "probe_bytes": probe_bytes},
            "kernels": kernels}


def kernel_summary(stats_csv, top=25):
    """Markdown table of a rocprofv3 --stats kernel_stats.csv (kernel,
    calls, total and average ms, share)."""
    rows = []
    with open(stats_csv) as f:
        for r in csv.DictReader(f):
            name = r["Name"].split("(")[0].replace("void ", "").replace("irlamd::", "")
            rows.append((name, int(r["Calls"]), float(r["TotalDurationNs"]) * 1e-6,
                         float(r["AverageNs"]) * 1e-6, float(r["Percentage"])))
    rows.sort(key=lambda t: -t[2])
    out = ["| kernel | calls | total ms | avg ms | % |", "|---|---|---|---|---|"]
    for name, n, tot, avg, pct in rows[:top]:
        out.append("| %s | %d | %.3f | %.4f | %.1f |" % (name, n, tot, avg, pct))
    return "\n".join(out)


def counter_summary(counter_csv, top=25):
    """Markdown table of the per-dispatch means of every counter of a
    rocprofv3 counter_collection.csv, per kernel."""
    vals, names = {}, []
    with open(counter_csv) as f:
        for r in csv.DictReader(f):
            k = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("irlamd::", "")
            c = r["Counter_Name"]
            if c not in names:
                names.append(c)
            vals.setdefault(k, {}).setdefault(c, []).append(float(r["Counter_Value"]))
    out = ["| kernel | dispatches | " + " | ".join(names) + " |",
           "|---|---|" + "---|" * len(names)]
    ks = sorted(vals, key=lambda k: -len(next(iter(vals[k].values()))))[:top]
    for k in ks:
        n = len(next(iter(vals[k].values())))
        out.append("| %s | %d | " % (k, n) + " | ".join(
            "%.4g" % np.mean(vals[k][c]) if c in vals[k] else "" for c in names) + " |")
    return "\n".join(out)


# ---------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=2,
                    help="views of the CPU baseline's sample")
    ap.add_argument("--cpu-procs", type=int, default=8,
                    help="processes of the CPU baseline's throughput figure")
    ap.add_argument("--only-headline", action="store_true",
                    help="time only the learned-scatter + BF16 line (for profiling runs)")
    ap.add_argument("--only-esse", action="store_true",
                    help="time only the plain FP32 ESSE line (for profiling runs)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the lines beside the headline and the plain FP32 ESSE line")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batch of 16 uneven reconstructions")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batch of 16 uneven reconstructions")
    ap.add_argument("--probe", choices=["copy", "read"],
                    help="only run a counter calibration probe")
    ap.add_argument("--probe-mib", type=int, default=256)
    ap.add_argument("--traffic-summary", nargs=4,
                    metavar=("FETCH_CSV", "WRITE_CSV", "PROBE_FETCH_CSV", "PROBE_WRITE_CSV"),
                    help="print the HBM traffic summary of rocprofv3 counter runs")
    ap.add_argument("--kernel-summary", metavar="KERNEL_STATS_CSV",
                    help="print a markdown table of a rocprofv3 kernel_stats.csv")
    ap.add_argument("--counter-summary", metavar="COUNTER_COLLECTION_CSV",
                    help="print a markdown table of a rocprofv3 counter_collection.csv")
    args = ap.parse_args()
    if args.kernel_summary:
        print(kernel_summary(args.kernel_summary))
        return
    if args.counter_summary:
        print(counter_summary(args.counter_summary))
        return
    if args.traffic_summary:
        print(json.dumps(traffic_summary(*args.traffic_summary,
                                         probe_bytes=args.probe_mib * (1 << 20)), indent=1))
        return
    if args.probe:
        probe(args.probe, args.probe_mib)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of this build "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    tmpdir = tempfile.mkdtemp(prefix="irlbench")
    K, W = args.steps, args.warmup
    if args.only_batch:
        print(json.dumps(batch_line(tmpdir, load_scatter_cnn())))
        return
    traffic = load_traffic()
    w = build_workload(tmpdir, name="c3")
    net = load_scatter_cnn()
    if args.only_esse:
        line = workload_line(w, K, W, "fp32", drf="auto")
        line.pop("_estimate", None)
        print(json.dumps(line))
        return
    plain = None
    if not args.only_headline:
        # the plain FP32 ESSE path, DRF algorithm timed at setup: the
        # denominator, and its estimate the reference of the other lines
        plain = workload_line(w, K, W, "fp32", drf="auto")
        ref = plain.pop("_estimate")
    head = workload_line(w, K, W, "bf16", ref_estimate=None if plain is None else ref,
                         cnn=net, check_every=8, drf="auto", traffic=traffic)
    head.pop("_estimate", None)
    cfg = w["cfg"]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "subiters/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 DRF / fp32",
        "data": "synthetic (seeded phantom, GPU forward projection of primary + ESSE scatter, "
                "Poisson 2M counts; CNN weights trained on the GPU from ESSE pairs)",
        "config": {"workload": "osem-128^3x120v-atn-drf31-esse33-8subsets, learned scatter "
                               "(CNN) + BF16 MFMA DRF",
                   "N": cfg["n"], "Nz": cfg["nz"], "views": cfg["views"], "subsets": 8,
                   "views_per_subset": cfg["vs"], "drf": "31x31 Gaussian, per plane",
                   "scatter": "ESSE 17x33 kernel (plain form); CNN 64^3, 4 dilated 3x3x3 "
                              "layers, 32 channels (learned form), checked against ESSE "
                              "once per 8 subsets, ESSE reinstated above 1% NRMSE",
                   "parallelism": "single GPU"},
        "roofline": head["roofline"],
        "stages_ms_per_launch": head["stages_ms_per_launch"],
        "stages_ms_per_subiteration": head["stages_ms_per_subiteration"],
        "bf16_entries_by_mode": head.get("bf16_entries_by_mode"),
        "cnn": head["cnn"],
        "cnn_training": cnn_training_record(),
    }
    if plain is not None:
        out["speedup_vs_plain_fp32_esse"] = head["value"] / plain["value"]
        out["errors"] = head["errors"]
        out["plain_fp32_esse"] = plain
    if not args.only_headline and not args.no_extras:
        # the reference's spatial DRF convolution with ESSE, the learned form
        # with the FP32 DRF, ESSE with the BF16 DRF
        out["esse_fp32_spatial"] = workload_line(w, K, W, "fp32", ref_estimate=ref, drf="spatial")
        out["cnn_fp32"] = workload_line(w, K, W, "fp32", ref_estimate=ref, cnn=net,
                                        check_every=8, drf="auto")
        out["esse_bf16"] = workload_line(w, K, W, "bf16", ref_estimate=ref)
        # the same geometry without scatter (plain FP32 and BF16 DRF)
        w2 = build_workload(tmpdir, name="c2")
        ns = workload_line(w2, K, W, "fp32", drf="auto")
        ref2 = ns.pop("_estimate")
        out["no_scatter_fp32"] = ns
        nb = workload_line(w2, K, W, "bf16", ref_estimate=ref2)
        nb["errors"]["projection_rel_l2_vs_fp64_oracle"] = bf16_projection_error(w2)
        out["no_scatter_bf16"] = nb
        if not args.no_batch:
            out["batch16"] = batch_line(tmpdir, net)
    if not args.no_cpu_baseline and not args.only_headline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_views, args.cpu_procs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
