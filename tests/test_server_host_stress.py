"""The serving loop's host code (tpr_amd/csrc/server.hip) on the CPU: built with
a fake HIP runtime (tests/cpp/fakehip: streams are threads, events complete
in stream order, kernel launchers are stubs whose answers the clients
recompute exactly) and run by tests/cpp/server_stress.cpp with several client
threads, a polling thread and a stats thread against a tight memory budget.

* under ThreadSanitizer: no data race between client calls, the worker, and
  the work the worker queues on its streams (the answers a client reads, the
  pinned table staging buffer, matrices generated on one slot and used on
  another), with and without the batching window;
* every answer is the exact value the stub rule gives, every request
  completes, the arena is empty once everything is released, and the
  server's count of solved systems equals the clients';
* failure handling: with every k-th launch made to fail on the host, exactly
  the requests of those batches fail (TPRH_ERROR_HIP, state "failed", no
  answers), their memory comes back and the server goes on answering the
  rest correctly -- a failure path no GPU test may force.

LLVM's ThreadSanitizer (ROCm's clang) is used: GCC 11's runtime does not
intercept pthread_cond_clockwait, which libstdc++ uses for timed waits on the
steady clock, and then reports every access under the server's mutex.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"


def build(out, sanitize):
    cmd = [CLANG if sanitize else (shutil.which("g++") or "g++"), "-std=c++17", "-O1", "-g", "-pthread",
           "-I", os.path.join(ROOT, "tests", "cpp", "fakehip"),
           "-x", "c++", os.path.join(ROOT, "tpr_amd", "csrc", "server.hip"),
           "-x", "none", os.path.join(ROOT, "tests", "cpp", "server_stress.cpp"), "-o", str(out)]
    if sanitize:
        cmd.insert(1, "-fsanitize=thread")
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-4000:]


def run(exe, clients, iters, seed, sanitize, fail_every=0):
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 exitcode=66 second_deadlock_stack=1"
    p = subprocess.run([str(exe), str(clients), str(iters), str(seed), str(fail_every)], capture_output=True,
                       text=True, env=env, timeout=300)
    if sanitize and "unexpected memory mapping" in p.stderr:
        pytest.skip("ThreadSanitizer cannot map its shadow memory on this kernel")
    assert p.returncode == 0, (p.stdout + p.stderr)[-6000:]
    assert "ThreadSanitizer" not in p.stderr, p.stderr[-6000:]
    assert p.stdout.strip().endswith("ok"), p.stdout
    return p.stdout


@pytest.fixture(scope="module")
def tsan_exe(tmp_path_factory):
    if not os.path.exists(CLANG):
        pytest.skip("no ROCm clang for the ThreadSanitizer build")
    exe = tmp_path_factory.mktemp("server_stress") / "server_stress_tsan"
    build(exe, sanitize=True)
    return exe


# odd seeds batch with a fixed window, seeds 2 and 3 (and so on, (seed / 2) % 2)
# with the adaptive window, which holds batches on the measured device time
@pytest.mark.parametrize("seed", [1, 2, 4])
def test_server_threads_under_tsan(tsan_exe, seed):
    out = run(tsan_exe, 4, 40, seed, sanitize=True)
    assert "max in flight" in out


@pytest.mark.parametrize("seed,fail_every", [(2, 7), (1, 3)])
def test_server_launch_failures_under_tsan(tsan_exe, seed, fail_every):
    out = run(tsan_exe, 4, 40, seed, sanitize=True, fail_every=fail_every)
    failed = int(re.search(r"\((\d+) failed\)", out).group(1))
    assert failed > 0, out


def test_server_stress_plain_build(tmp_path):
    """The same host code built by the system compiler without a sanitizer,
    more clients and requests."""
    exe = tmp_path / "server_stress"
    build(exe, sanitize=False)
    run(exe, 8, 60, 3, sanitize=False)
