"""bench.py's output contract: one JSON line with the metric, the whole-job
value, the step timing, and the roofline of the dominant kernel.

The CPU test checks that a run asking for more than one GPU is refused with
a message; the GPU test runs a small batched solve (16 shots, one timed
step, no CPU legs) in a child process and checks the line it prints.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _last_json(stdout):
    lines = [ln for ln in stdout.splitlines() if ln.startswith("{")]
    assert lines, stdout[-2000:]
    return json.loads(lines[-1])


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in _last_json(r.stdout)["error"]


@pytest.mark.gpu
def test_bench_line_contract():
    cmd = [sys.executable, BENCH, "--shots", "16", "--batch", "16", "--steps", "1", "--warmup", "1",
           "--block", "8", "--grid", "2048", "--max-inner", "24", "--band", "0.2", "--no-cpu",
           "--no-plain", "--no-cold-leg", "--no-extra"]
    print("running:", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=110, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    d = _last_json(r.stdout)
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config", "roofline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 1 and d["warmup"] == 1
    assert d["higher_is_better"] is True and d["value"] > 0 and d["ms_per_step"] > 0
    # whole-job throughput: 16 shots per timed step
    assert d["value"] == pytest.approx(16 * 1000.0 / d["ms_per_step"], rel=1e-6)
    assert d["dtype"] == "f32" and isinstance(d["config"]["workload"], str)
    roof = d["roofline"]
    assert roof["bound"] in ("hbm", "mfma") and roof["unit"] == "GB/s"
    assert roof["peak"] == 8000.0
    assert roof["frac"] == pytest.approx(roof["achieved"] / roof["peak"], rel=1e-9)
    assert 0.0 < roof["frac"] < 1.0
