"""Host-side pieces of bench.py: the CPU baseline leg and the per-leg guard.

A run with `bench.py --steps 20 --warmup 5` builds the GPU leg's model with 25 steps of source
samples while the CPU window can be longer: the baseline must build its own series for its window."""
import json
import subprocess
import sys

import numpy as np
import pytest

import bench
from rfnetwork_amd import workloads


def test_cpu_baseline_builds_its_own_series():
    res = bench.cpu_baseline("microstrip", budget_s=0.6, max_window=40, runs=3)
    assert res["unit"] == "Mcell-updates/s" and res["kind"] == "port"
    assert res["value"] > 0 and res["cores"] >= 1
    assert res["cpu_model"] and res["host_cores"] >= res["cores"]
    by = res["by_threads"]
    assert str(res["cores"]) in by
    for v in by.values():
        assert 4 <= v["window_steps"] <= 40 and len(v["runs_s"]) == 3
        assert v["runs_s"] == sorted(v["runs_s"]) and v["median_s"] == v["runs_s"][1]


def test_oracle_inputs_refuse_a_short_series():
    """A window longer than the model's series is refused with a clear message (not the oracle's)."""
    with pytest.raises(ValueError, match="steps of source samples"):
        bench._oracle_inputs(workloads.microstrip(Nt=25), 1, 200)
    c, probes, monitors, *_ = bench._oracle_inputs(workloads.microstrip(Nt=60), 1, 40)
    assert all(len(p["values"]) == 40 for p in probes)
    assert all(m["dtft_phase"].shape[0] == 41 for m in monitors if "dtft_phase" in m)


def test_thread_counts_follow_the_reference_cap():
    cores, top, counts = bench.cpu_thread_counts()
    assert top == min(cores, 20, bench.HOST_CPU_SHARE)
    assert counts[0] == top and min(counts) == min(4, top)


def test_guard_records_errors():
    def boom():
        raise ValueError("Invalid source array.")

    assert bench.guarded("x", boom) == {"error": "ValueError: Invalid source array."}
    assert bench.guarded("y", lambda a, b=0: a + b, 1, b=2) == 3


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], capture_output=True, text=True,
                       cwd=str(bench.ROOT), timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
