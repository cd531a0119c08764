"""bench.py's command-line contract: one JSON line with the fields the
measurement needs (checked on the GPU with a short run), and a clear refusal
of more than one GPU (checked on the CPU: it exits before touching torch)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr and not r.stdout.strip()


@pytest.mark.gpu
def test_headline_line_fields():
    print("bench.py --steps 3 --warmup 1 --no-cpu-baseline", flush=True)
    r = subprocess.run([sys.executable, BENCH, "--steps", "3", "--warmup", "1", "--no-cpu-baseline"],
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["dtype"] == "f64" and d["value"] > 0
    assert "workload" in d["config"]
    roof = d["roofline"]
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-3
    assert d["f32c"]["normwise_max_rel"] <= 1e-6
    # the other measured workloads, summarised in the same line
    others = d["other_workloads"]
    assert set(others) == {"box2d_r1_temporal", "star3d_r1_temporal", "box2d_r8", "batched", "star3d_r4_large",
                           "star3d_r4_511_remainder"}
    for name, o in others.items():
        assert "error" not in o, (name, o)
        assert o["value"] > 0 and o["unit"] == "GLup/s", name
    assert others["star3d_r1_temporal"]["vs_one_step"] > 1.0
