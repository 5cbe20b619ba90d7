"""The bench's contract on the host: one GPU only, the counter summary it quotes as
`roofline.traffic`, and the CPU baseline leg (the oracle timed on the host cores)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("key", [
    # the dense rows: schedule and workgroup shape as the bench forms them for the 256^3 channel
    "fp64 256^3 flag-word-first (4 bricks/workgroup)",
    "fp64 256^3 loads-with-flag-word (2 bricks/workgroup)",
    "fp32s 256^3 flag-word-first (8 rows/workgroup)",
    "fp32s 256^3 loads-with-flag-word (2 rows/workgroup)",
    "fp64 256^3 aa-in-place (4 bricks/workgroup)",
    "fp32s 256^3 aa-in-place (2 rows/workgroup)",
    "patches fp32s step",
    "patches fp64 step",
])
def test_profile_summary_has_the_traffic_rows_the_bench_reads(key):
    t = bench.read_traffic(key)
    assert t is not None, key
    # measured bytes per launch exceed the algorithmic ones (whole lines, ghost records) but stay
    # within a factor two of them
    alg = 13417967 * 38 * (8 if key.startswith("fp64") else 4)
    if key.startswith("patches"):
        alg = 52166656 * 38 * (8 if "fp64" in key else 4)
    assert alg <= t <= 2 * alg, (key, t, alg)


def test_cpu_baseline_leg_on_a_small_sample():
    from madpac_amd import workloads as wl

    r = bench.cpu_baseline(wl, n=16, budget_s=0.2)
    assert r["kind"] == "port" and r["unit"] == "MLUPS"
    assert r["value"] > 0 and r["cores"] >= 1
    assert "16^3" in r["sample"]
