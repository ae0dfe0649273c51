"""The tpr_pm driver (src/pm.cpp:118-268 surface): argument handling on the
CPU, GPU runs checked against the reference's known answer and the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TPR_PM = os.path.join(ROOT, "tpr_amd", "tpr_pm")


def run(args, env=None, timeout=120):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([TPR_PM] + [str(a) for a in args], capture_output=True, text=True, env=e,
                          timeout=timeout)


def test_usage_error():
    p = run([])
    assert p.returncode != 0 and "Usage:" in p.stderr


def test_unknown_solver():
    p = run([1024, 256, 1, "nosuch"])
    assert p.returncode != 0 and "Solver Not Found" in p.stderr


@pytest.mark.parametrize("name", ["ptpr", "TPR", "pcr", "lapack"])
def test_cpu_solver_names_are_refused(name):
    p = run([1024, 256, 1, name])
    assert p.returncode != 0 and "CPU solver" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name,s", [("cuptpr", 64), ("CUPTPR", 1024), ("cutpr", 256), ("hipptpr", 128),
                                    ("hiptpr", 512), ("hippcr", 0), ("hipcr", 0), ("cusparse", 0),
                                    ("rocsparse", 0)])
def test_known_answer_through_the_driver(golden, tmp_path, name, s):
    """TPR_PM_INPUT=fixed: the system of test/ans1024.txt; x written with %f.
    cusparse / rocsparse: the vendor slot (src/pm.cpp:240-246) on rocSPARSE."""
    out = tmp_path / "x.txt"
    p = run([1024, s, 3, name, out], env={"TPR_PM_INPUT": "fixed"})
    assert p.returncode == 0, p.stderr
    lines = p.stdout.strip().splitlines()
    assert len(lines) == 3
    for ln in lines:
        m = re.fullmatch(r"1024,%d,%s,([0-9.eE+-]+)" % (s, re.escape(name)), ln)
        assert m, ln
        assert float(m.group(1)) > 0
    x = np.array([float(v) for v in out.read_text().strip().split(",")])
    ans = golden["ans1024"]
    assert x.size == 1024
    assert np.all(np.abs(x - ans) <= 1e-2) and np.max((x - ans) ** 2) <= 1e-3


@pytest.mark.gpu
def test_readme_invocation_config0():
    """README.md:28-30 runs `tpr_pm 2048 512 1000 PTPR`; the GPU solver slot of the same call."""
    p = run([2048, 512, 1000, "cuptpr"], timeout=300)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.strip().splitlines()
    assert len(lines) == 1000
    ms = [float(ln.split(",")[3]) for ln in lines]
    assert all(v > 0 for v in ms)


@pytest.mark.gpu
def test_invalid_shape_reports_error():
    p = run([1000, 256, 1, "cuptpr"])
    assert p.returncode != 0 and "[Error]" in p.stderr


@pytest.mark.gpu
def test_serve_bench_client_closed_loop():
    """tpr_serve_bench, the C++ load generator of the serving loop: a short
    closed-loop run completes every request and reports consistent counts."""
    import json
    exe = os.path.join(ROOT, "tpr_amd", "tpr_serve_bench")
    p = subprocess.run([exe, "--rate", "0", "--requests", "300", "--outstanding", "32", "--budget-gb", "4",
                        "--seed", "7"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    d = json.loads(p.stdout.strip().splitlines()[-1])
    assert d["mode"] == "closed" and d["requests"] == 300
    assert d["systems"] >= 300 and d["equations"] >= 300 * 512
    assert d["systems_per_s"] > 0 and d["latency_ms"]["p50"] <= d["latency_ms"]["p99"] <= d["latency_ms"]["max"]
    assert d["peak_bytes"] <= 4 << 30


@pytest.mark.gpu
def test_serve_bench_client_threads_and_window():
    """The load generator with four client threads and a batching window:
    every request completes (the collector checks each one's state)."""
    import json
    exe = os.path.join(ROOT, "tpr_amd", "tpr_serve_bench")
    p = subprocess.run([exe, "--rate", "0", "--requests", "400", "--outstanding", "64", "--clients", "4",
                        "--budget-gb", "4", "--min-batch", "256", "--max-wait-us", "200", "--seed", "8"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    d = json.loads(p.stdout.strip().splitlines()[-1])
    assert d["requests"] == 400 and d["clients"] == 4 and d["min_batch"] == 256
    assert 0 < d["batches"] <= 400 and d["queued_ms"]["p50"] <= d["latency_ms"]["p99"]


def _build_known_answer(tmp_path):
    exe = tmp_path / "known_answer"
    lib = os.path.join(ROOT, "tpr_amd")
    p = subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "known_answer.cpp"), "-o", str(exe),
                        "-L" + lib, "-ltpr_hip", "-Wl,-rpath," + lib], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return exe


def test_known_answer_program_builds(tmp_path):
    """The C++ restatement of the reference's example tests compiles against
    the headers and the library (its run is the gpu test below)."""
    exe = _build_known_answer(tmp_path)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stderr


@pytest.mark.gpu
def test_known_answer_program(tmp_path):
    """tests/cpp/known_answer.cpp: test/cuexamples.cpp and test/examples.cpp
    in C++ on the engine, with the reference's own acceptance."""
    exe = _build_known_answer(tmp_path)
    p = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden", "ans1024.npy")], capture_output=True,
                       text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failure(s)" in p.stdout
