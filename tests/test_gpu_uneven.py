"""GPU parity of the uneven-batch path (tprh_uneven_*): systems of different
sizes in one persistent launch.

* the plan's generator writes, system by system, the oracle's generator bits;
* exact arithmetic: every system bit-identical to the oracle's solve at its
  size class's slice size s;
* fast arithmetic: every system within rel-L2 1e-5 of the oracle;
* a system's bits do not depend on the batch around it (alone vs mixed);
* graph capture: every replay equals the eager launch.
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
from tpr_amd import autotune
from tpr_amd.uneven import UnevenPlan, log_uniform_sizes

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def bits_equal(x, y):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    same = x.view(np.uint32) == y.view(np.uint32)
    return bool(np.all(same | (np.isnan(x) & np.isnan(y))))


def class_s(n, schedule):
    if schedule and n in schedule:
        return schedule[n][0]
    return min(n, 256)


def alloc(plan):
    return [torch.empty(max(plan.equations, 1), device="cuda") for _ in range(6)]


def host_system(plan, k, arrays):
    lo, hi = plan.offset(k), plan.offset(k + 1)
    return [v[lo:hi] for v in arrays]


MIXED = [16, 16384, 256, 32, 1024, 64, 8192, 128, 512, 2048, 4096, 256, 16, 1024, 16384, 128]

SCHEDULES = {
    "default": None,
    "shapes": {256: (256, 8, 32), 512: (512, 8, 64), 1024: (256, 4, 64), 2048: (128, 8, 16),
               4096: (256, 16, 16), 8192: (512, 8, 64), 64: (16, 1, 16), 128: (32, 2, 16)},
    # s alone: exact arithmetic takes a shape of the kernel's menu for s = 512
    "s512": {1024: (512,), 2048: (512,), 4096: (512,), 8192: (512,)},
}


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_uneven_generator_matches_oracle(method):
    sizes = MIXED
    plan = UnevenPlan(sizes, method=method)
    a, b, c, d, x, xt = alloc(plan)
    plan.generate("dominant", 77, 5, a, b, c, d, xt)
    torch.cuda.synchronize()
    h = [v.cpu().numpy() for v in (a, b, c, d, xt)]
    for k, n in enumerate(sizes):
        ga = host_system(plan, k, h)
        ref = O.gen_system(77, 5 + k, n)
        for u, v in zip(ga, ref):
            assert bits_equal(u, v), (k, n)


@pytest.mark.parametrize("sched", list(SCHEDULES))
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_uneven_exact_bit_identical(method, sched):
    schedule = SCHEDULES[sched]
    sizes = MIXED + [1 << int(e) for e in np.random.default_rng(3).integers(4, 15, 40)]
    say("uneven exact", method, sched, len(sizes))
    plan = UnevenPlan(sizes, method=method, arith="exact", schedule=schedule)
    a, b, c, d, x, xt = alloc(plan)
    plan.generate("weak", 78, 0, a, b, c, d, xt)
    plan.solve(a, b, c, d, x)
    torch.cuda.synchronize()
    h = [v.cpu().numpy() for v in (a, b, c, d, x)]
    for k, n in enumerate(sizes):
        sa, sb, sc, sd, sx = host_system(plan, k, h)
        ref = O.solve_batch(method, sa, sb, sc, sd, n, class_s(n, schedule))
        assert bits_equal(sx, ref), (k, n)


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_uneven_fast_within_tolerance(method):
    sizes = log_uniform_sizes(300, 8, 14, seed=9)
    say("uneven fast", method, len(sizes))
    plan = UnevenPlan(sizes, method=method, arith="fast")
    a, b, c, d, x, xt = alloc(plan)
    plan.generate("dominant", 79, 0, a, b, c, d, xt)
    plan.solve(a, b, c, d, x)
    torch.cuda.synchronize()
    h = [v.cpu().numpy() for v in (a, b, c, d, x, xt)]
    worst = 0.0
    for k, n in enumerate(sizes):
        sa, sb, sc, sd, sx, sxt = host_system(plan, k, h)
        ref = O.solve_batch(method, sa, sb, sc, sd, n, class_s(n, None))
        worst = max(worst, O.rel_l2(sx, ref))
        assert O.rel_l2(sx, sxt) <= 1e-5
    say("worst rel-L2 vs oracle", worst)
    assert worst <= FAST_TOL


@pytest.mark.parametrize("arith", ["fast", "exact"])
def test_uneven_result_independent_of_batch(arith):
    """A system solved alone and inside a mixed batch gives the same bits."""
    sizes = MIXED
    mixed = UnevenPlan(sizes, arith=arith)
    a, b, c, d, x, xt = alloc(mixed)
    mixed.generate("dominant", 80, 0, a, b, c, d)
    mixed.solve(a, b, c, d, x)
    torch.cuda.synchronize()
    for k in (0, 1, 4, 6, 13):
        lo, hi = mixed.offset(k), mixed.offset(k + 1)
        alone = UnevenPlan([sizes[k]], arith=arith)
        xa = torch.empty(hi - lo, device="cuda")
        alone.solve(a[lo:hi], b[lo:hi], c[lo:hi], d[lo:hi], xa)
        torch.cuda.synchronize()
        assert torch.equal(xa.view(torch.int32), x[lo:hi].view(torch.int32)), k


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_uneven_unit_diagonal(method):
    sizes = [1024, 64, 4096, 16]
    plan = UnevenPlan(sizes, method=method, arith="exact")
    a, b, c, d, x, xt = alloc(plan)
    plan.generate("dominant", 81, 0, a, b, c, d)
    torch.cuda.synchronize()
    h = [v.cpu().numpy() for v in (a, b, c, d)]
    na, nc, nd = [], [], []
    for k in range(len(sizes)):
        sa, sb, sc, sd = host_system(plan, k, h)
        u, v, w = O.normalize(sa, sb, sc, sd)
        na.append(u)
        nc.append(v)
        nd.append(w)
    da, dc, dd = (torch.from_numpy(np.concatenate(v)).cuda() for v in (na, nc, nd))
    plan.solve(da, None, dc, dd, x)
    torch.cuda.synchronize()
    hx = x.cpu().numpy()
    for k, n in enumerate(sizes):
        ref = (O.ptpr if method == "ptpr" else O.tpr)(na[k], nc[k], nd[k], min(n, 256))
        assert bits_equal(hx[plan.offset(k):plan.offset(k + 1)], ref), k


def test_uneven_graph_replays_match_eager():
    sizes = log_uniform_sizes(500, 8, 14, seed=11)
    plan = UnevenPlan(sizes)
    a, b, c, d, x, xt = alloc(plan)
    plan.generate("dominant", 82, 0, a, b, c, d)
    plan.solve(a, b, c, d, x)
    torch.cuda.synchronize()
    eager = x.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.solve(a, b, c, d, x)  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.solve(a, b, c, d, x)
    for rep in range(4):
        x.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(x.view(torch.int32), eager.view(torch.int32)), rep
        d.add_(0.0)  # eager work between replays


def test_uneven_empty_and_single_class():
    plan = UnevenPlan([])
    assert plan.equations == 0
    x = torch.empty(1, device="cuda")
    plan.solve(x, x, x, x, x)
    plan = UnevenPlan([512] * 1000)
    a, b, c, d, x, xt = alloc(plan)
    plan.generate("dominant", 83, 0, a, b, c, d, xt)
    plan.solve(a, b, c, d, x)
    torch.cuda.synchronize()
    hx = x.cpu().numpy().reshape(1000, 512)
    hxt = xt.cpu().numpy().reshape(1000, 512)
    assert max(O.rel_l2(hx[k], hxt[k]) for k in range(0, 1000, 37)) <= 1e-5


def test_autotune_table_runs_and_is_valid():
    table = autotune.tune([256, 2048], eq_target=1 << 20, reps=3, warmup=1)
    sched = autotune.schedule_of(table)
    for n, cand in sched.items():
        assert tuple(cand) in autotune.candidates(n)
    plan = UnevenPlan([256, 2048, 256], schedule=sched)
    assert plan.equations == 256 + 2048 + 256


def _random_plan(seed):
    rng = np.random.default_rng(seed)
    sizes = [1 << int(e) for e in rng.integers(4, 15, int(rng.integers(5, 60)))]
    schedule = {}
    for n in set(sizes):
        cands = autotune.candidates(n)
        schedule[n] = cands[int(rng.integers(0, len(cands)))]
    return sizes, schedule, ["ptpr", "tpr"][int(rng.integers(0, 2))], bool(rng.integers(0, 2))


# TPR_UNEVEN_SWEEP widens the sweep for a one-off run
@pytest.mark.parametrize("seed", range(int(__import__("os").environ.get("TPR_UNEVEN_SWEEP", "6"))))
def test_uneven_random_schedules_exact(seed):
    """Random batches with a random menu shape per size class: every system
    bit-identical to the oracle at its class's slice size."""
    sizes, schedule, method, with_b = _random_plan(1000 + seed)
    say("uneven sweep", seed, method, with_b, len(sizes), schedule)
    plan = UnevenPlan(sizes, method=method, arith="exact", schedule=schedule)
    a, b, c, d, x, xt = alloc(plan)
    plan.generate("weak", 90 + seed, 0, a, b, c, d)
    torch.cuda.synchronize()
    h = [v.cpu().numpy() for v in (a, b, c, d)]
    if not with_b:
        na, nc, nd = [], [], []
        for k in range(len(sizes)):
            sa, sb, sc, sd = host_system(plan, k, h)
            u, v, w = O.normalize(sa, sb, sc, sd)
            na.append(u)
            nc.append(v)
            nd.append(w)
        h = [np.concatenate(na), None, np.concatenate(nc), np.concatenate(nd)]
        a, c, d = (torch.from_numpy(v).cuda() for v in (h[0], h[2], h[3]))
        b = None
    plan.solve(a, b, c, d, x)
    torch.cuda.synchronize()
    hx = x.cpu().numpy()
    for k, n in enumerate(sizes):
        lo, hi = plan.offset(k), plan.offset(k + 1)
        s = schedule[n][0]
        if with_b:
            ref = O.solve_batch(method, h[0][lo:hi], h[1][lo:hi], h[2][lo:hi], h[3][lo:hi], n, s)
        else:
            ref = (O.ptpr if method == "ptpr" else O.tpr)(h[0][lo:hi], h[2][lo:hi], h[3][lo:hi].copy(), s)
        assert bits_equal(hx[lo:hi], ref), (k, n, schedule[n])
