"""bench.py's output contract: one JSON line with the keys the benchmark
driver reads, the roofline and CPU-baseline objects, and a refusal of more
than one GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")

TOP_KEYS = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
            "vs_baseline", "dtype", "data", "config", "roofline"}
ROOF_KEYS = {"bound", "achieved", "peak", "unit", "frac", "traffic"}
CPU_KEYS = {"value", "unit", "cores", "kind", "sample"}


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout.strip() == ""


def _run(args):
    p = subprocess.run([sys.executable, BENCH] + args, capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert p.returncode == 0, p.stderr[-3000:]
    lines = [ln for ln in p.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, lines
    return json.loads(lines[0])


@pytest.mark.gpu
@pytest.mark.parametrize("args", [["--steps", "10", "--warmup", "2", "--cpu-seconds", "1", "--no-vendor"],
                                  ["--workload", "config0"]])
def test_json_line(args):
    d = _run(args)
    assert TOP_KEYS <= set(d), TOP_KEYS - set(d)
    assert ROOF_KEYS <= set(d["roofline"])
    assert CPU_KEYS <= set(d["cpu_baseline"])
    assert d["n_gpus"] == 1 and d["higher_is_better"] is True and d["value"] > 0
    assert "workload" in d["config"] and d["dtype"] == "f32" and d["data"] == "synthetic"
    r = d["roofline"]
    assert r["unit"] == "GB/s" and r["peak"] == 8000.0 and abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-3
    assert d["cpu_baseline"]["cores"] >= 1 and d["cpu_baseline"]["value"] > 0
    if "--workload" not in args:
        assert d["steps"] == 10 and d["warmup"] == 2
        assert d["accuracy"]["rel_l2_vs_cpu_max"] <= 1e-5
        # the reference's analytic count for PTPR, N = 1024, S = 256
        assert d["analytic_flops"]["per_system"] == 118963
        assert set(d["batch_stream"]) == {"buffer_sets", "x_stores", "2_streams", "3_streams"}
