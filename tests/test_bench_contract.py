"""bench.py's command-line contract, checked without a GPU: the default
workload and step counts, the single-GPU guard, the workload table, and the
profile-derived roofline inputs it reads."""
import json
import os
import subprocess
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_defaults(monkeypatch):
    monkeypatch.setattr(sys, "argv", ["bench.py"])
    a = bench.parse()
    assert a.gpus == 1 and a.workload == "duct128-fp64"
    assert a.steps == 500 and a.warmup == 10


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_workloads_have_their_sizes():
    w = bench.WORKLOADS
    assert (w["duct128-fp64"]["nx"], w["duct128-fp64"]["ny"], w["duct128-fp64"]["nz"]) == (128, 128, 128)
    assert w["duct128-fp64"]["dtype"] == "f64" and w["duct128-fp32"]["dtype"] == "f32"
    assert (w["large256-fp32"]["nx"], w["large256-fp32"]["ny"], w["large256-fp32"]["nz"]) == (256, 192, 192)
    # one of 8 x-slabs of the large run: 32 planes + two 4-plane halos, 1 mm cells
    slab = w["slab32-fp32"]
    assert (slab["nx"], slab["ny"], slab["nz"]) == (256 // 8 + 8, 192, 192)
    assert abs(slab["over"]["lx"] - slab["nx"] * 1e-3) < 1e-12
    # algorithmic bytes per cell and RK3 stage: 13.33 fields of 8 (4) bytes
    assert abs(w["duct128-fp64"]["bytes_per_cell_stage"] - 106.7) < 0.05
    assert abs(w["duct128-fp32"]["bytes_per_cell_stage"] - 53.3) < 0.05


def test_profile_inputs_are_present_and_consistent():
    with open(os.path.join(ROOT, "profiles", "rhs_traffic.json")) as f:
        traffic = json.load(f)
    with open(os.path.join(ROOT, "profiles", "rhs_valu.json")) as f:
        valu = json.load(f)
    for wl in ("duct128-fp64", "duct128-fp32"):
        # HBM bytes per stage of the RHS launch for each tile width and XCD
        # mapping recorded ("tile_k:xcd_map")
        assert traffic[wl]
        for key, val in traffic[wl].items():
            tk, xm = (int(x) for x in key.split(":"))
            assert tk in (0, 16, 32, 64, 128) and 0 <= xm <= 4
            assert bench.traffic_from_profiles(wl, xm, tk) == val > 0
        v = bench.valu_from_profiles(wl)
        assert v == valu[wl] and v["insts_per_stage"] > 0
        # physical shares: a quad-cycle issues at most once without dual issue
        assert 0 < v["busy_frac"] <= 1 and 0 <= v["dual_issue_share"] <= 1 and 0 < v["wait_any_frac"] < 1
        assert abs(v["insts_per_simd_quad"] - v["busy_frac"] * (1 + v["dual_issue_share"])) < 0.05
    # FP32 moves fewer bytes than FP64 for the same stage and launch shape
    for key in set(traffic["duct128-fp32"]) & set(traffic["duct128-fp64"]):
        assert traffic["duct128-fp32"][key] < traffic["duct128-fp64"][key]
