"""bench.py pieces that run without a GPU: the one-GPU rule, the CPU
baseline record (the FP64 oracle's MD loop with the reference's phase
split, cabanamd_impl.h:404-425), the drift helper and the committed HBM
traffic record the roofline line carries."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_cpu_baseline_record():
    b = bench._cpu_baseline((6, 6, 6), 20)
    assert b["unit"] == "atom-steps/s" and b["kind"] == "port"
    assert b["value"] > 0 and b["cores"] >= 1
    assert "restatement of the reference" in b["sample"]
    split = b["split_s"]
    assert set(split) == {"force", "neigh", "comm", "int", "other", "total"}
    parts = sum(split[k] for k in ("force", "neigh", "comm", "int", "other"))
    assert all(v >= 0 for v in split.values())
    # the phases cover the loop (the total is the loop's own clock)
    assert abs(parts - split["total"]) <= 0.05 * split["total"] + 1e-4
    assert split["force"] > 0 and split["neigh"] > 0


def test_drift_helper():
    assert bench._drift(np.array([1.0])) is None
    assert abs(bench._drift(np.array([-2.0, -2.002, -1.999])) - 0.001) < 1e-12


def test_traffic_record_is_committed():
    with open(os.path.join(ROOT, "profiles", "lj_full_fused_2M_traffic.json")) as fh:
        t = json.load(fh)
    # bytes per launch of the fused kernel at 2,048,000 atoms: between the
    # positions alone and the 436 B/atom algorithmic model
    assert 32 * 2048000 < t["bytes_per_launch"] < 436 * 2048000
