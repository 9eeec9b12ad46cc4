"""Host logic of bench.py: the stage-to-kernel map, the algorithmic work of
the single-kernel stages, and the HBM traffic summary of rocprofv3 counter
files (synthetic CSVs in rocprofv3's counter_collection layout)."""
import csv
import os

import numpy as np

import bench


def test_stage_kernels():
    assert bench.stage_kernel("cnn_att") == "k_att_src"
    assert bench.stage_kernel("drf_prj_bf16mix") == "k_drf_bf16"
    assert bench.stage_kernel("drf_bck_fft") == "k_drf_fft"
    assert bench.stage_kernel("drf_prj") == "k_drf"
    assert bench.stage_kernel("bck_update") == "k_bck"
    # multi-kernel stages have no single roofline
    for s in ("cnn_src", "esse_conv_source", "esse_conv", "esse_source"):
        assert bench.stage_kernel(s) is None


def _write_counters(path, rows):
    with open(path, "w", newline="") as f:
        wr = csv.DictWriter(f, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        wr.writeheader()
        for i, (k, c, v) in enumerate(rows):
            wr.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": c,
                         "Counter_Value": v})


def test_traffic_summary_scales_by_the_probes(tmp_path):
    mib = 256
    nbytes = mib << 20
    prj = "void irlamd::k_drf_bf16<31, 8, false>(irlamd::DrfArgs, HIP_vector_type<unsigned int, 4u> const*)"
    bck = "void irlamd::k_drf_bf16<31, 8, true>(irlamd::DrfArgs, HIP_vector_type<unsigned int, 4u> const*)"
    att = "void irlamd::k_att_src<false>(float const*, float const*)"
    upd = "void irlamd::k_bck<false, false>(irlamd::BckArgs)"
    nrm = "void irlamd::k_bck<true, false>(irlamd::BckArgs)"
    p = {n: str(tmp_path / (n + ".csv")) for n in ("f", "w", "pf", "pw")}
    _write_counters(p["f"], [(prj, "FETCH_SIZE", 1000), (prj, "FETCH_SIZE", 3000),
                             (bck, "FETCH_SIZE", 500), (att, "FETCH_SIZE", 100),
                             (upd, "FETCH_SIZE", 7), (nrm, "FETCH_SIZE", 9)])
    _write_counters(p["w"], [(prj, "WRITE_SIZE", 10), (prj, "WRITE_SIZE", 30),
                             (bck, "WRITE_SIZE", 50), (att, "WRITE_SIZE", 40),
                             (upd, "WRITE_SIZE", 1), (nrm, "WRITE_SIZE", 2)])
    # the read probe's counter reads twice its bytes, the copy's write exactly
    _write_counters(p["pf"], [("at::reduce_kernel<...>", "FETCH_SIZE", 2 * nbytes / 1024),
                              ("tiny", "FETCH_SIZE", 1)])
    _write_counters(p["pw"], [("at::copy_kernel<...>", "WRITE_SIZE", nbytes / 1024)])
    t = bench.traffic_summary(p["f"], p["w"], p["pf"], p["pw"], nbytes)
    assert abs(t["calibration"]["fetch_scale"] - 0.5) < 1e-12
    assert abs(t["calibration"]["write_scale"] - 1.0) < 1e-12
    k = t["kernels"]["k_drf_bf16<31, 8, false>"]
    assert k["fetch_kib"] == 2000 and k["write_kib"] == 20
    assert k["hbm_bytes_per_dispatch"] == 2000 * 1024 * 0.5 + 20 * 1024
    assert bench.traffic_for(t, "drf_prj_bf16mix") == k["hbm_bytes_per_dispatch"]
    assert bench.traffic_for(t, "drf_bck_bf16mix") == 500 * 512 + 50 * 1024
    assert bench.traffic_for(t, "cnn_att") == 100 * 512 + 40 * 1024
    assert bench.traffic_for(t, "bck_update") == 7 * 512 + 1024
    assert bench.traffic_for(t, "ratio") is None
    assert bench.traffic_for(None, "cnn_att") is None


class _View:
    def __init__(self, cfcr):
        self.CFCR = cfcr


class _Parms:
    NumPixels, NumSlices, BinWidth = 4, 2, 1.0


class _Rec:
    num_subsets, views_per_subset = 2, 1

    def bf16_planes(self):
        return np.array([0, 2, 1, 3]), None


def test_row_entries_flops_and_plane_bytes():
    # rows at distances cfcr - 1.5 + r from the face; table from 0.5 with
    # spacing 1 and entries of 1x1, 3x3, 3x5 taps (the last one clamps)
    w = {"parms": _Parms(), "views": [_View(1.0), _View(2.0)], "start": 0.5,
         "offs": [(1, 1), (2, 2), (2, 3)]}
    ent = bench.row_entries(w)
    assert ent.tolist() == [[-1, 0, 1, 2], [0, 1, 2, 2]]
    rec = _Rec()
    # subset 0 = view 0, subset 1 = view 1
    assert bench.subset_views(rec) == [[0], [1]]
    per = 2 * 2 * 4
    assert bench.drf_flops_per_subset(w, rec) == [per * (1 + 9 + 15), per * (1 + 9 + 15 + 15)]
    plane = 4 * 2
    # modes: entry 0 plain, 1 data split, 2 weight split; unconvolved rows FP32
    assert bench.att_write_bytes(w, rec, True) == [plane * (4 + 2 + 4 + 2),
                                                   plane * (2 + 4 + 2 + 2)]
    assert bench.att_write_bytes(w, rec, False) == [4 * plane * 4] * 2
