"""bench.py's contract on the CPU: it refuses more than one GPU, turns PMC
counter CSVs into calibrated per-kernel HBM bytes, and names the dominant
kernel as the profiles do."""
import csv
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def _write_counters(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, kib in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": kib})


def test_pmc_summary_scales_by_the_copy_calibration(tmp_path):
    """The copy kernel moves 64 MiB each way; FETCH_SIZE reads half of that and
    WRITE_SIZE all of it, so fetch values are doubled and writes kept."""
    copy = 64 << 20
    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    # KiB per dispatch; two dispatches of k_raster with different readings
    _write_counters(fetch, "FETCH_SIZE", [("void sft::k_copy(uint4 const*, uint4*, unsigned long)", copy / 2 / 1024),
                                         ("sft::k_raster(sft::FramePtrs)", 1000.0),
                                         ("sft::k_raster(sft::FramePtrs)", 3000.0),
                                         ("void sft::k_shade<true>(sft::FramePtrs)", 500.0),
                                         ("__amd_rocclr_copyBuffer", 7.0)])
    _write_counters(write, "WRITE_SIZE", [("void sft::k_copy(uint4 const*, uint4*, unsigned long)", copy / 1024),
                                         ("sft::k_raster(sft::FramePtrs)", 4000.0),
                                         ("sft::k_raster(sft::FramePtrs)", 4000.0),
                                         ("void sft::k_shade<true>(sft::FramePtrs)", 250.0),
                                         ("__amd_rocclr_copyBuffer", 7.0)])
    s = _bench().summarize_pmc(str(fetch), str(write), copy, "w")
    assert s["fetch_factor"] == pytest.approx(2.0) and s["write_factor"] == pytest.approx(1.0)
    k = s["kernels"]["k_raster"]
    assert k["dispatches"] == 2
    assert k["fetch_bytes"] == round(2000.0 * 1024 * 2.0)
    assert k["write_bytes"] == round(4000.0 * 1024)
    assert k["hbm_bytes"] == k["fetch_bytes"] + k["write_bytes"]
    assert s["kernels"]["k_shade<true>"]["fetch_bytes"] == round(500.0 * 1024 * 2.0)
    assert "__amd_rocclr_copyBuffer" not in s["kernels"]


def test_dominant_kernel_names_match_the_profiles():
    b = _bench()
    assert b.pass_kernel("visibility", "box_room") == "k_raster<1>"
    assert b.pass_kernel("shading", "box_room") == "k_shade<false>"
    assert b.pass_kernel("shading", "crowd_shadows") == "k_shade<true>"


def test_measured_pass_bandwidth_sums_the_pass_kernels():
    """pass_measured_gbs adds the committed PMC bytes of every kernel of a
    pass and divides by the pass time."""
    import json
    b = _bench()
    ks = json.load(open(os.path.join(ROOT, "profiles", "traffic_box_room_1920x1080.json")))["kernels"]
    got = b.measured_pass_gbs("box_room_1920x1080", {"geometry": 0.02, "binning": 0.04, "visibility": 0.1, "shading": 0.1})
    want = sum(ks[k]["hbm_bytes"] for k in ("k_raster_plan", "k_raster<1>", "k_raster<2>", "k_tile_lights") if k in ks) / 0.1e-3 / 1e9
    assert abs(got["visibility"] - want) < 0.1
    assert b.measured_pass_gbs("no_such_workload", {"geometry": 1, "binning": 1, "visibility": 1, "shading": 1}) is None
