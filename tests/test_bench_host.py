"""Host-side parts of bench.py that need no GPU: the committed per-workload HBM traffic and counter
summaries are found for the workload they were measured on and for no other, and a run on more
than one GPU is refused with a message (the build is single-GPU)."""
import json
import os
import subprocess
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_traffic_lookup_is_per_workload():
    for workload in ("heavy_cluster", "ragged_molecules"):
        path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.txt")
        assert os.path.exists(path), path
        for kernel in ("k_screen", "k_radial_mfma", "k_contract"):
            v = bench.traffic_bytes(kernel, workload)
            assert isinstance(v, int) and v > 0, (workload, kernel, v)
    # the two workloads were profiled separately: the ragged screening moves far more bytes
    assert bench.traffic_bytes("k_screen", "ragged_molecules") > 10 * bench.traffic_bytes("k_screen", "heavy_cluster")
    # no summary for this workload or kernel: no traffic figure rather than another workload's
    assert bench.traffic_bytes("k_screen", "small_pair") is None
    assert bench.traffic_bytes("k_no_such_kernel", "heavy_cluster") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "single-GPU" in msg["error"]


def test_valu_issue_from_the_committed_counters():
    """the screening's VALU issue load: its wave instructions (rocprofv3 summary) at 4 cycles each over
    the issue capacity of the duration; absent for a workload without a summary"""
    v = bench.valu_issue("ragged_molecules", "k_screen", 0.175)
    assert v is not None and v["valu_wave_insts"] > 1e10 and 0.0 < v["frac"] < 1.5
    assert abs(v["frac"] - v["valu_wave_insts"] * 4 / (0.175 * 1024 * 2.4e9)) < 1e-12
    assert bench.valu_issue("small_pair", "k_screen", 0.001) is None
    assert bench.valu_issue("ragged_molecules", "k_no_such_kernel", 0.1) is None
    assert bench.fp64_counters("ragged_molecules", 0.41)["issued_flops_per_build"] > 1e12
