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


def tune(sizes=None, method="ptpr", eq_target=16 << 20, reps=20, warmup=3, log=None):
    """Measure every candidate for each size class; returns the table dict
    {"method", "classes": {n: {"choice": [s, E, L], "us": {...}, "gbps": ...}}}."""
    sizes = sizes or [1 << k for k in range(8, 15)]
    res = {"method": method, "eq_per_class": eq_target, "classes": {}}
    for n in sizes:
        row = {"us": {}}
        best = None
        for cand in candidates(n):
            us, nsys = _time_class(n, cand, method, eq_target, reps, warmup)
            key = "%d/%dx%d" % cand
            row["us"][key] = round(us, 2)
            if best is None or us < best[0]:
                best = (us, cand, nsys)
            if log:
                log("n=%d %s %.2f us" % (n, key, us))
        us, cand, nsys = best
        row["choice"] = list(cand)
        row["gbps"] = round(20.0 * n * nsys / (us * 1e-6) / 1e9, 1)
        res["classes"][str(n)] = row
    return res


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
    args = ap.parse_args(argv)
    t0 = time.time()
    table = tune([1 << k for k in range(args.min_log2, args.max_log2 + 1)], args.method, reps=args.reps,
                 log=lambda m: print(m, flush=True))
    table["tuned_seconds"] = round(time.time() - t0, 1)
    if args.out:
        save(table, args.out)
    print(json.dumps(schedule_of(table)))


if __name__ == "__main__":
    main()
