"""bench.py's contract pieces that run without a GPU: the one-GPU refusal of
--gpus N > 1, the traffic summary's arithmetic (per-dispatch counter means
in KiB, scaled by what the same counter reads for a 1 GiB copy), the CPU
baseline record, and the default workload."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "only --gpus 1" in r.stderr


def _csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, counter, value in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value})


def test_traffic_summary_scales_by_the_copy_calibration(tmp_path):
    copy = "__amd_rocclr_copyBuffer"
    apply_, rup = "void k_lz_apply<float>(lz_args, ...)", "void k_lz_rupdate<float>(lz_args, ...)"
    # a 1 GiB copy reads as 512 MiB of FETCH_SIZE (factor 2) and 1 GiB of WRITE_SIZE (factor 1)
    _csv(tmp_path / "cf.csv", [(copy, "FETCH_SIZE", 524288.0)] * 3)
    _csv(tmp_path / "cw.csv", [(copy, "WRITE_SIZE", 1048576.0)] * 3)
    _csv(tmp_path / "f.csv", [(apply_, "FETCH_SIZE", 1000.0), (apply_, "FETCH_SIZE", 3000.0),
                              (rup, "FETCH_SIZE", 100.0)])
    _csv(tmp_path / "w.csv", [(apply_, "WRITE_SIZE", 500.0), (apply_, "WRITE_SIZE", 500.0),
                              (rup, "WRITE_SIZE", 50.0)])
    out = bench.summarize_traffic(str(tmp_path / "t.json"), str(tmp_path / "f.csv"),
                                  str(tmp_path / "w.csv"), str(tmp_path / "cf.csv"),
                                  str(tmp_path / "cw.csv"), "demo", ["cmd"])
    assert out["scale_fetch"] == pytest.approx(2.0) and out["scale_write"] == pytest.approx(1.0)
    assert out["launches"] == 2
    assert out["fetch_bytes_per_launch"] == pytest.approx(2000.0 * 1024 * 2.0)
    assert out["write_bytes_per_launch"] == pytest.approx(500.0 * 1024)
    assert out["bytes_per_launch"] == pytest.approx(2000.0 * 1024 * 2.0 + 500.0 * 1024)
    assert out["rupdate"]["bytes_per_launch"] == pytest.approx(100.0 * 1024 * 2.0 + 50.0 * 1024)
    assert json.load(open(tmp_path / "t.json"))["workload"] == "demo"


def test_cpu_baseline_record():
    r = bench.cpu_baseline("hex16")
    assert set(r) == {"value", "unit", "cores", "kind", "sample", "seconds", "threads",
                      "cpu_model", "nproc"}
    assert r["kind"] == "port" and r["unit"] == "Melem/s"
    # every core the process may use, capped at a one-GPU box's 16
    assert r["cores"] == min(len(os.sched_getaffinity(0)), 16)
    assert r["value"] > 0 and r["sample"] == "the full workload"


def test_threaded_oracle_partitions_like_the_serial_one():
    """The CPU baseline's OpenMP build: on a mesh below its threading
    threshold only whole segments run in parallel, so the partition equals
    the serial oracle's element for element."""
    from oracle import pyoracle as O
    from parrsb_amd import meshgen as M
    vtx, xyz, nv = M.box_hex(24, 20, 16, jitter=0.1, seed=4)
    a, ra = O.part_mesh(vtx, xyz, nv, 24)
    b, rb = O.part_mesh(vtx, xyz, nv, 24, threads=4)
    assert (a == b).all() and ra == rb


def test_default_workload_is_the_256_cubed_box():
    gen, kw, parts, desc = bench.CONFIGS["warped256"]
    assert gen == "warped_box_hex" and kw["n"] == 256 and parts == 65536
    assert desc == "warped_hex_256^3_65536_parts"
