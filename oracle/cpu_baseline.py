"""CPU baseline of one OSEM sub-iteration of the scatter workload (128^3 x
120 views, attenuation, 31x31 Gaussian DRF, ESSE with the 17 x 33 kernel):
the oracle's reference-semantics FP32 Project and BackProject (prjbck.c, one
thread, the reference's loop order) plus the reference's per-view ESSE
effective source as FFT convolutions (oracle.esse_fft, float32, one thread),
timed on a sample of the sub-iteration's views.

CPU BASELINE / TEST INFRASTRUCTURE ONLY.  Run as a child process by
bench.py (python -m oracle.cpu_baseline --views K --core C), pinned to one
core, printing one JSON line: seconds per view and the split between the
projector pair and the scatter model.  Inputs are built from the same seeds
as the benchmark workload, in numpy (no GPU).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def run(nviews, seed=2):
    from libirl_amd import synth
    from oracle import oracle as O
    from oracle.esse_fft import EsseFFT
    c = synth.CONFIGS["c3"]
    n, nz, pix = c["n"], c["nz"], c["pix"]
    act, mu = synth.phantom(n, nz, seed, n_inserts=c["inserts"])
    views = synth.views_ellipse(c["views"], n, seed, 20.0, 28.0)
    tmp = tempfile.mkdtemp(prefix="irlcpu")
    offs, psfs, start = synth.drf_for_orbit(os.path.join(tmp, "d.drft"), pix, n, views,
                                            c["half"])
    drf = O.DrfTable(start, pix, offs, psfs)
    k, km = synth.gaussian_mixture_krnl(c["krnl_offs"], c["krnl_depths"], pix)
    k = synth.scale_krnl_to_fraction(k, 0.3)
    esse = EsseFFT(O.Esse(k, km, 0.15, pix, n, nz, real="f32"))
    atn = O.reorder(nz, n, n, mu * pix, real="f32")
    x = O.reorder(nz, n, n, act, real="f32")
    img = np.zeros((n, nz, n), np.float32)
    O.init_prj(nz, n, n, 3, real="f32")
    # the views of subset 0 (every 8th), as a sub-iteration takes them
    sample = [views[8 * m] for m in range(nviews)]
    t_pb = t_esse = 0.0
    for v in sample:
        t0 = time.perf_counter()
        ang = float(np.float32(0.5 * np.pi + np.float64(np.float32(v.Angle))))
        xr = O.rot3dpar(ang, x, 0, nz, n, n, real="f32")
        mr = O.rot3dpar(ang, atn, 0, nz, n, n, real="f32")
        esse.effsrc(xr, mr)
        t1 = time.perf_counter()
        prj = O.project(nz, n, n, 3, pix, v.Angle, v.CFCR, drf, atn, x, real="f32")
        O.backproject(nz, n, n, 3, pix, v.Angle, v.CFCR, drf, atn, prj, img, atn_rotated=1,
                      real="f32")
        t2 = time.perf_counter()
        t_esse += t1 - t0
        t_pb += t2 - t1
    return {"views": nviews, "s_per_view": (t_pb + t_esse) / nviews,
            "s_per_view_project_backproject": t_pb / nviews,
            "s_per_view_scatter_model": t_esse / nviews}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--views", type=int, default=2)
    ap.add_argument("--core", type=int, default=-1)
    a = ap.parse_args()
    if a.core >= 0:
        os.sched_setaffinity(0, {a.core})
    r = run(a.views)
    r["core"] = a.core
    print(json.dumps(r))


if __name__ == "__main__":
    main()
