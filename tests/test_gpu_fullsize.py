"""Parity at the benchmark workloads' full sizes, not only at reduced
ones (the reference's own tests assert only the last case of their loops,
test/examples.cpp:76-102; these check every class and every system):

* the uneven batch -- 20000 systems, N = 2^U{8..14}, the committed autotuned
  schedule, one persistent launch: exact arithmetic bit-identical to the
  oracle on a stratified sample of 2100 systems (300 of every size class);
  fast arithmetic within 1e-5 of the generator's truth on all 20000;
* the large systems -- 256 weakly dominant systems x N=65536 through the
  multi-workgroup path: exact bit-identical to the oracle on all 256; each
  precision variant's error per system against the host FP64 Thomas solve
  held at the level it reaches (fp32 <= 1e-6, fp32 + 1 refinement step
  <= 5e-8, fp16 coefficients + 1 step <= 1e-7, + 2 steps <= 5e-8);
* serving -- the serving client (tpr_amd/tpr_serve_bench) for 2 s of
  Poisson arrivals at 2000 req/s with the 64 GB budget: every request
  completes, and the answers of every 40th request (all its right-hand
  sides, written out by the client) equal the oracle's within 1e-5.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

import oracle as O
from tpr_amd import autotune
from tpr_amd import batched as B
from tpr_amd.uneven import UnevenPlan, log_uniform_sizes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def bits_equal(x, y):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    same = x.view(np.uint32) == y.view(np.uint32)
    return bool(np.all(same | (np.isnan(x) & np.isnan(y))))


def per_system_rel(x, ref, n):
    x = np.asarray(x, np.float64).reshape(-1, n)
    ref = np.asarray(ref, np.float64).reshape(-1, n)
    return np.linalg.norm(x - ref, axis=1) / np.linalg.norm(ref, axis=1)


@pytest.fixture(scope="module")
def uneven20000():
    sizes = log_uniform_sizes(20000, 8, 14, seed=2020)
    sched = autotune.schedule_of(autotune.load())
    probe = UnevenPlan(sizes, "ptpr", "fast", sched)
    tot = probe.equations
    offs = probe.offsets.copy()
    probe.close()
    a, b, c, d, xt = (torch.empty(tot, device="cuda") for _ in range(5))
    UnevenPlan(sizes, "ptpr", "fast", sched).generate("dominant", 4242, 0, a, b, c, d, xt)
    torch.cuda.synchronize()
    host = {k: v.cpu().numpy() for k, v in (("a", a), ("b", b), ("c", c), ("d", d), ("xt", xt))}
    return {"sizes": sizes, "sched": sched, "offs": offs, "dev": (a, b, c, d), "host": host}


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_uneven_20000_full_batch(uneven20000, method):
    sizes, sched, offs, h = uneven20000["sizes"], uneven20000["sched"], uneven20000["offs"], uneven20000["host"]
    a, b, c, d = uneven20000["dev"]
    x = torch.empty_like(a)
    # fast: every system against the truth
    say("uneven 20000 fast", method)
    with UnevenPlan(sizes, method, "fast", sched) as p:
        p.solve(a, b, c, d, x)
        torch.cuda.synchronize()
    xf = x.cpu().numpy()
    worst = 0.0
    for n in sorted(set(sizes.tolist())):
        ks = np.nonzero(sizes == n)[0]
        idx = np.concatenate([np.arange(offs[k], offs[k + 1]) for k in ks])
        worst = max(worst, per_system_rel(xf[idx], h["xt"][idx], n).max())
    say("uneven 20000 fast worst rel-L2 vs truth", worst)
    assert worst <= 1e-5
    # exact: 300 systems of every class, bit for bit against the oracle at the
    # class's slice size
    say("uneven 20000 exact", method)
    with UnevenPlan(sizes, method, "exact", sched) as p:
        p.solve(a, b, c, d, x)
        torch.cuda.synchronize()
    xe = x.cpu().numpy()
    rng = np.random.default_rng(11)
    checked = 0
    for n in sorted(set(sizes.tolist())):
        ks = np.nonzero(sizes == n)[0]
        pick = np.sort(rng.choice(ks, min(300, ks.size), replace=False))
        idx = np.concatenate([np.arange(offs[k], offs[k + 1]) for k in pick])
        s = sched.get(n, (min(n, 256),))[0]
        ref = O.solve_batch(method, h["a"][idx], h["b"][idx], h["c"][idx], h["d"][idx], n, s)
        assert bits_equal(xe[idx], ref), (method, n)
        checked += pick.size
    assert checked >= 2000


def test_split_256x65536_full_batch():
    nsys, n, s = 256, 65536, 256
    say("split 256x65536 generate")
    t = B.alloc_batch(nsys, n)
    B.generate("weak", 30303, 0, nsys, n, t["a"], t["b"], t["c"], t["d"], t["xt"])
    torch.cuda.synchronize()
    h = {k: t[k].cpu().numpy() for k in "abcd"}
    xf = O.thomas_f64_batch(h["a"], h["b"], h["c"], h["d"], n)
    errs = {}
    x = t["x"]
    # exact: all 256 systems bit for bit
    say("split 256x65536 exact")
    B.solve_ex(t["a"], t["b"], t["c"], t["d"], x, n, s, arith="exact")
    torch.cuda.synchronize()
    assert bits_equal(x.cpu().numpy(), O.solve_batch("ptpr", h["a"], h["b"], h["c"], h["d"], n, s))
    say("split 256x65536 precision variants")
    B.solve_ex(t["a"], t["b"], t["c"], t["d"], x, n, s)
    torch.cuda.synchronize()
    errs["fp32"] = per_system_rel(x.cpu().numpy(), xf, n).max()
    B.solve_ex(t["a"], t["b"], t["c"], t["d"], x, n, s, refine=True)
    torch.cuda.synchronize()
    errs["fp32+1"] = per_system_rel(x.cpu().numpy(), xf, n).max()
    h16 = [t[k].half() for k in "abc"]
    B.solve_ex(h16[0], h16[1], h16[2], t["d"], x, n, s)
    torch.cuda.synchronize()
    errs["fp16"] = per_system_rel(x.cpu().numpy(), xf, n).max()
    for step in (1, 2):
        B.solve_ex(t["a"], t["b"], t["c"], t["d"], x, n, s, refine=True)
        torch.cuda.synchronize()
        errs["fp16+%d" % step] = per_system_rel(x.cpu().numpy(), xf, n).max()
    say(errs)
    assert errs["fp32"] <= 1e-6
    assert errs["fp32+1"] <= 5e-8
    assert errs["fp16"] <= 1e-2
    assert errs["fp16+1"] <= 1e-7
    assert errs["fp16+2"] <= 5e-8


def test_serving_poisson_run_at_full_budget(tmp_path):
    exe = os.path.join(ROOT, "tpr_amd", "tpr_serve_bench")
    dump = str(tmp_path / "answers.bin")
    say("serving serving run")
    p = subprocess.run([exe, "--rate", "2000", "--seconds", "2", "--budget-gb", "64", "--seed", "4044",
                        "--min-batch", "1024", "--max-wait-us", "100", "--dump", dump, "--dump-every", "40"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-3000:]
    res = json.loads(p.stdout.strip().splitlines()[-1])
    say({k: res[k] for k in ("requests", "systems", "systems_per_s", "latency_ms", "max_inflight_seen")})
    assert res["requests"] >= 3000 and res["systems"] >= res["requests"]
    assert res["peak_bytes"] <= 64 << 30
    # the dumped answers: records (k, n, nrhs, j0) + nrhs * n floats
    raw = np.fromfile(dump, dtype=np.uint8)
    pos = checked = 0
    worst = 0.0
    while pos < raw.size:
        k, n, nrhs, j0 = (int(v) for v in raw[pos:pos + 32].view(np.int64))
        pos += 32
        xs = raw[pos:pos + 4 * n * nrhs].view(np.float32).reshape(nrhs, n)
        pos += 4 * n * nrhs
        a, b, c, _, _ = O.gen_system(4044, k, n)
        for j in range(nrhs):
            ref = O.solve_batch("ptpr", a, b, c, O.gen_rhs(4044, k, j0 + j, n), n, min(n, 256))
            worst = max(worst, O.rel_l2(xs[j], ref))
        checked += 1
    say("serving checked requests", checked, "worst", worst)
    assert checked >= res["requests"] // 40 - 1
    assert worst <= 1e-5
