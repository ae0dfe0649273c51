"""Per-size-class slice schedule for uneven batches, chosen by measurement.

For every size class n (a power of two) the tuner enumerates the slice
shapes the persistent uneven-batch kernel carries, (s, E, L) with s = E * L,
s | n and (n / s) * L <= 1024 lanes, and times each one on a batch of that
class alone (about 16 Mi equations, larger than the 256 MB last-level cache,
so the measurement is HBM-bound like the real batch), in the persistent
kernel itself.  The fastest shape wins.  The result maps n -> (s, E, L) and
is what ``UnevenPlan(schedule=...)`` takes; ``save``/``load`` keep it as
JSON with the timings that chose it.

This generalises the reference's fixed launch shape for one (n, s)
(src/ptpr.cu:648-666, one block of s threads per slice) to a measured
choice per size class.

    python -m tpr_amd.autotune --out tpr_amd/schedules/uneven_gfx950.json

With --batched it measures the batched entry point instead: for each
(method, arithmetic, n, s, batch size) every compiled slice shape (E, L)
with E * L = s that fits a workgroup, whether it gives the same bits as the
default shape, and its time; the fastest same-bits shape is the choice, and
the library's built-in rule (tprh_slice_shape) is recorded beside it.

    python -m tpr_amd.autotune --batched --out tpr_amd/schedules/batched_gfx950.json
"""
import argparse
import json
import os
import time

# (E, L) shapes of the persistent kernel, in the order of its shape table
UNEVEN_SHAPES = [(1, 16), (2, 16), (4, 16), (8, 16), (16, 16), (8, 32), (4, 64), (8, 64)]
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedules", "uneven_gfx950.json")


def candidates(n):
    """(s, E, L) triples the uneven kernel can run for systems of n equations."""
    out = []
    for E, L in UNEVEN_SHAPES:
        s = E * L
        if s > n or n % s:
            continue
        m = n // s
        if m & (m - 1) or m * L > 1024:
            continue
        out.append((s, E, L))
    return out


def default_choice(n):
    """The untuned schedule: s = min(n, 256) with its default shape."""
    s = min(n, 256)
    return (s, s // 16, 16) if s >= 16 else None


def _time_class(n, cand, method, eq_target, reps, warmup):
    import torch
    from .uneven import UnevenPlan
    nsys = max(1, eq_target // n)
    plan = UnevenPlan([n] * nsys, method=method, arith="fast", schedule={n: cand})
    t = plan.equations
    a, b, c, d, x = (torch.empty(t, device="cuda") for _ in range(5))
    plan.generate("dominant", 1234, 0, a, b, c, d)
    for _ in range(warmup):
        plan.solve(a, b, c, d, x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        plan.solve(a, b, c, d, x)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1000.0 / reps
    plan.close()
    return us, nsys


UNEVEN_MARGIN = 0.03


def choose_class(n, us, margin=UNEVEN_MARGIN):
    """The schedule entry for size class n from its measured times us
    ({"s/ExL": microseconds}): the default (default_choice) unless another
    candidate is faster by more than `margin` (a fraction).  Per-class times
    within a few percent of each other do not carry over to a mixed batch in
    one launch (on MI355X a table that took s = 128 for three classes on
    1-2 % per-class leads ran 2-4 % slower on the 20000-system batch than
    s = 256 throughout), so noise-level leads do not move the choice."""
    best = min(us, key=lambda k: us[k])
    dflt = "%d/%dx%d" % default_choice(n)
    if dflt in us and us[dflt] <= us[best] * (1.0 + margin):
        best = dflt
    s, el = best.split("/")
    E, L = el.split("x")
    return [int(s), int(E), int(L)]


def tune(sizes=None, method="ptpr", eq_target=16 << 20, reps=20, warmup=3, log=None, margin=UNEVEN_MARGIN):
    """Measure every candidate for each size class; returns the table dict
    {"method", "classes": {n: {"choice": [s, E, L], "us": {...}, "gbps": ...}}}
    with each choice made by choose_class."""
    sizes = sizes or [1 << k for k in range(8, 15)]
    res = {"method": method, "eq_per_class": eq_target, "margin": margin, "classes": {}}
    for n in sizes:
        row = {"us": {}}
        nsys = None
        for cand in candidates(n):
            us, nsys = _time_class(n, cand, method, eq_target, reps, warmup)
            key = "%d/%dx%d" % cand
            row["us"][key] = round(us, 2)
            if log:
                log("n=%d %s %.2f us" % (n, key, us))
        row["choice"] = choose_class(n, row["us"], margin)
        us = row["us"]["%d/%dx%d" % tuple(row["choice"])]
        row["gbps"] = round(20.0 * n * nsys / (us * 1e-6) / 1e9, 1)
        res["classes"][str(n)] = row
    return res


# (E, L) slice shapes compiled into the batched kernel
BATCHED_SHAPES = [(1, 2), (1, 4), (1, 8), (1, 16), (2, 16), (4, 16), (8, 16), (16, 16), (4, 64), (8, 32), (8, 64),
                  (16, 32), (16, 64), (32, 64)]
BATCHED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedules", "batched_gfx950.json")


def default_batched_shape(s):
    """The library's default (E, L) for slice size s: one DPP row of 16 lanes
    per slice up to s = 256, then 32 or 64 lanes."""
    if s <= 16:
        return (1, s)
    if s <= 256:
        return (s // 16, 16)
    return {512: (16, 32), 1024: (16, 64), 2048: (32, 64)}[s]


def fast_form(method, E, L):
    """Which fast-arithmetic stage 1 a slice shape runs: PTPR on slices of
    one 16-lane DPP row with E >= 2 takes the two-level form (a Thomas sweep
    in each lane, PCR across the lanes' last rows), every other shape the
    register PCR; shapes of one form give the same bits."""
    return "two-level" if method == "ptpr" and L == 16 and E >= 2 else "pcr"


def batched_candidates(n, s):
    """(E, L) shapes the batched entry point can run for (n, s) in one workgroup."""
    m = n // s
    return [(E, L) for E, L in BATCHED_SHAPES if E * L == s and m * L <= 1024]


def batched_key(method, arith, n, s, nsys):
    return "%s/%s/n%d/s%d/nsys%d" % (method, arith, n, s, nsys)


def _time_batched(sets, n, s, method, arith, shape, reps, warmup):
    """Device time per launch: `reps` launches captured in one HIP graph and
    replayed, so that small batches are timed on the GPU, not by the rate at
    which the host issues launches."""
    import torch
    from .batched import solve_batched

    def launch(i):
        t = sets[i % len(sets)]
        solve_batched(t["a"], t["b"], t["c"], t["d"], t["x"], n, s, method, arith, shape=shape)

    for i in range(warmup):
        launch(i)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for i in range(reps):
            launch(i)
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    g.replay()
    e1.record()
    torch.cuda.synchronize()
    del g
    return e0.elapsed_time(e1) * 1000.0 / reps


def pick(row, default, margin):
    """The fastest shape whose answers equal the default shape's; the default
    itself unless another is faster by more than `margin` (a fraction), so
    that timing noise does not move the choice."""
    dkey = "%dx%d" % default
    same = [k for k, v in row["same_bits"].items() if v]
    best = min(same, key=lambda k: row["us"][k])
    if dkey in same and row["us"][dkey] <= row["us"][best] * (1.0 + margin):
        best = dkey
    E, L = best.split("x")
    return int(E), int(L)


def tune_batched(slice_sizes=(256, 512), slices=(1, 2, 4, 8, 16), batches=(1, 256, 4096), methods=("ptpr", "tpr"),
                 ariths=("fast", "exact"), max_equations=16 << 20, reps=50, warmup=3, passes=2, margin=0.03,
                 log=None):
    """Measure every candidate shape of the batched entry point; returns
    {"cases": {key: {"us": {"ExL": t}, "same_bits": {"ExL": bool}, "choice": [E, L], "rule": [E, L]}}}.
    Launches are timed from a replayed HIP graph; buffer sets are cycled so
    that a batch larger than a few MB is read from HBM, not from the
    last-level cache; "same_bits" compares each shape's
    answers with the default shape's on the same inputs; "rule" is the shape
    the library took for these arguments when the table was made."""
    import torch
    from .batched import alloc_batch, generate, slice_shape, solve_batched
    res = {"eq_cap": max_equations, "reps": reps, "passes": passes, "margin": margin, "cases": {}}
    for s in slice_sizes:
        for m in slices:
            n = s * m
            for nsys in batches:
                nsys = max(1, min(nsys, max_equations // n))
                set_bytes = 20 * n * nsys
                nsets = max(1, min(8, -(-(600 << 20) // set_bytes)))
                sets = [alloc_batch(nsys, n, with_xt=False) for _ in range(nsets)]
                for k, t in enumerate(sets):
                    generate("dominant", 77, k * nsys, nsys, n, t["a"], t["b"], t["c"], t["d"])
                t0 = sets[0]
                for method in methods:
                    for arith in ariths:
                        cands = batched_candidates(n, s)
                        rule = slice_shape(n, s, nsys, method, arith)
                        solve_batched(t0["a"], t0["b"], t0["c"], t0["d"], t0["x"], n, s, method, arith,
                                      shape=default_batched_shape(s))
                        ref = t0["x"].clone()
                        row = {"us": {}, "same_bits": {}}
                        for shape in cands:
                            key = "%dx%d" % shape
                            solve_batched(t0["a"], t0["b"], t0["c"], t0["d"], t0["x"], n, s, method, arith,
                                          shape=shape)
                            torch.cuda.synchronize()
                            row["same_bits"][key] = bool(torch.equal(t0["x"].view(torch.int32),
                                                                     ref.view(torch.int32)))
                            row["us"][key] = float("inf")
                        for _ in range(passes):  # interleaved passes, the best time of each shape
                            for shape in cands:
                                key = "%dx%d" % shape
                                us = _time_batched(sets, n, s, method, arith, shape, reps, warmup)
                                row["us"][key] = round(min(row["us"][key], us), 2)
                        row["choice"] = list(pick(row, default_batched_shape(s), margin))
                        row["rule"] = list(rule)
                        key = batched_key(method, arith, n, s, nsys)
                        res["cases"][key] = row
                        if log:
                            log("%s rule %dx%d choice %dx%d %s" % (key, rule[0], rule[1], row["choice"][0],
                                                                 row["choice"][1], row["us"]))
                del sets
                torch.cuda.empty_cache()
    return res


HEADER_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "batched_shapes.inc")


def parse_batched_key(key):
    method, arith, n, s, nsys = key.split("/")
    return method, arith, int(n[1:]), int(s[1:]), int(nsys[4:])


def header_text(table):
    """The rows of the batched table as the C initialiser capi.hip compiles
    in: {method, arith, s, m, nsys, E, L} (method 0 PTPR / 1 TPR, arith 0
    fast / 1 exact), in a fixed order."""
    rows = []
    for key, r in table["cases"].items():
        method, arith, n, s, nsys = parse_batched_key(key)
        rows.append((0 if method == "ptpr" else 1, 0 if arith == "fast" else 1, s, n // s, nsys, r["choice"][0],
                     r["choice"][1]))
    lines = ["// Generated by `python -m tpr_amd.autotune --batched` from schedules/batched_gfx950.json;",
             "// {method (0 PTPR, 1 TPR), arith (0 fast, 1 exact), s, m, nsys, E, L}: the fastest slice shape",
             "// measured on MI355X whose answers equal the default shape's."]
    lines += ["{%d, %d, %d, %d, %d, %d, %d}," % row for row in sorted(rows)]
    return "\n".join(lines) + "\n"


def schedule_of(table):
    """n -> (s, E, L) mapping of a tuned table."""
    return {int(n): tuple(r["choice"]) for n, r in table["classes"].items()}


def save(table, path=DEFAULT_PATH):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")


def load(path=DEFAULT_PATH):
    with open(path) as f:
        return json.load(f)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--method", default="ptpr", choices=["ptpr", "tpr"])
    ap.add_argument("--out", default=None)
    ap.add_argument("--min-log2", type=int, default=8)
    ap.add_argument("--max-log2", type=int, default=14)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--batched", action="store_true", help="tune the batched entry point's slice shapes")
    ap.add_argument("--header", default=None, help="with --batched: also write the table as capi.hip's C rows")
    args = ap.parse_args(argv)
    t0 = time.time()
    if args.batched:
        table = tune_batched(log=lambda m: print(m, flush=True))
        table["tuned_seconds"] = round(time.time() - t0, 1)
        if args.out:
            save(table, args.out)
        if args.header:
            with open(args.header, "w") as f:
                f.write(header_text(table))
        return
    table = tune([1 << k for k in range(args.min_log2, args.max_log2 + 1)], args.method, reps=args.reps,
                 log=lambda m: print(m, flush=True))
    table["tuned_seconds"] = round(time.time() - t0, 1)
    if args.out:
        save(table, args.out)
    print(json.dumps(schedule_of(table)))


if __name__ == "__main__":
    main()
