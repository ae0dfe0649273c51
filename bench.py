"""Benchmark: batched PTPR on one MI355X.

Workload (one step): one batched PTPR solve of 4096 independent fp32 systems,
N = 1024 each, slice size S = 256, diagonally dominant generator, inputs
resident in HBM.  Throughput is systems/s over the whole timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--method ptpr|tpr]
                    [--arith fast|exact] [--eager] [--no-cpu]

The batch is kept in ROT buffer sets (each 83.9 MB with x) that the steps
cycle through, more than the 256 MB Infinity Cache holds, so every step reads
its inputs from HBM rather than from the cache.

The K timed steps are captured once into a HIP graph and replayed, so the
GPU runs them back to back (--eager launches them one by one from Python).

Roofline: algorithmic bytes per launch = 20 B per equation (read a, b, c, d;
write x) x 1024 x 4096 = 83,886,080 B, divided by the kernel's average
duration (HIP events on the launch stream around the K launches, / K),
against the 8 TB/s HBM peak.

CPU baseline: the reference's own PTPR (oracle/_ref, compiled from its
sources; the oracle's restatement where that build is absent), one solver
per thread over the systems, on this host's cores, on a bounded sample of the
same batch; the GPU answers of that sample are compared with it (rel-L2).

Side figures in the same line: the other method (TPR beside PTPR), the
rocSPARSE comparator, a device copy (the practical HBM ceiling) and
"batch_stream": the same batches issued over 2 and 3 streams so that
consecutive launches overlap (one batch's compute tail under the next one's
loads), over eight buffer sets.

The rest of the metric, from short runs of the other configurations
(--no-side leaves them out):
  "serve"      the serving loop, Poisson arrivals at 2000 req/s for 2 s
               (fixed batching window) and 20000 requests closed-loop
               (adaptive window): p50/p99/max latency, systems/s, against
               one launch per right-hand side; "latency_ms" at the top level
               is its Poisson latency;
  "uneven"     20000 systems of N = 2^U{8..14} in one persistent launch
               (ms, GB/s, fraction of the HBM peak; accuracy of 200 systems);
  "precision"  256 weakly dominant systems x N=65536: fp32, fp16 storage + one
               FP64-residual refinement step, fp32 + one step (ms, errors per
               system against the host FP64 Thomas solve).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "systems/s & % HBM BW at rel L2≤1e-5 vs plain TPR; p50/p99 serve latency"
HBM_PEAK_GBS = 8000.0
NSYS, N, S = 4096, 1024, 256
ROT = 4  # buffer sets cycled through (4 x 83.9 MB > 256 MB cache)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--method", default="ptpr", choices=["ptpr", "tpr"])
    ap.add_argument("--arith", default="fast", choices=["fast", "exact"])
    ap.add_argument("--eager", action="store_true",
                    help="launch the timed steps one by one from Python instead of replaying a captured HIP graph")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocSPARSE comparator leg")
    ap.add_argument("--no-streams", action="store_true",
                    help="skip the batch_stream leg (overlapping launches), e.g. under a kernel trace whose "
                         "per-launch average should be the single-stream figure")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--shape", default=None, help="slice shape E,L (default: library default for S)")
    ap.add_argument("--chunk", type=int, default=0, help="precision workload: systems per round of the split path")
    ap.add_argument("--split-form", type=int, default=0, choices=[0, 1, 2, 3],
                    help="precision workload: form of the multi-workgroup path (0 the library's choice, 1 two "
                         "passes, 2 one pass, 3 pipelined)")
    ap.add_argument("--fp32-only", action="store_true",
                    help="precision workload: the fp32 solve alone (no fp16, refinement or per-system variants), "
                         "e.g. for counter runs")
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "uneven", "precision", "serve", "copy", "config0"],
                    help="batch: 4096 x N=1024, S=256 (default, the headline); uneven: 20000 systems of N = 2^U{8..14} "
                         "in one persistent launch; precision: 256 weakly dominant x N=65536, fp16 storage and "
                         "refinement variants; serve: the serving loop under Poisson load; copy: a 256 MiB "
                         "device copy, the practical HBM ceiling and the counters' calibration; config0: the "
                         "reference's documented run, one system N=2048, S=512, 1000 solves")
    ap.add_argument("--rate", type=float, default=2000.0, help="serve workload: mean request arrivals per second")
    ap.add_argument("--seconds", type=float, default=5.0, help="serve workload: length of the Poisson run")
    ap.add_argument("--retune", action="store_true",
                    help="uneven workload: run the autotuner now instead of loading the committed schedule")
    ap.add_argument("--no-side", action="store_true",
                    help="batch workload: leave out the short config-2/3/4 legs (uneven, precision, serve) that "
                         "the default line carries beside the headline")
    return ap.parse_args()


def gpu_leg(args):
    import torch

    import tpr_amd
    from tpr_amd import batched as B

    torch.cuda.set_device(0)
    shape = tuple(int(v) for v in args.shape.split(",")) if args.shape else None
    sets = []
    for r in range(ROT):
        t = B.alloc_batch(NSYS, N, with_xt=(r == 0))
        B.generate("dominant", 20240 + r, 0, NSYS, N, t["a"], t["b"], t["c"], t["d"], t.get("xt"))
        sets.append(t)
    torch.cuda.synchronize()

    def step(i):
        # launched on the current stream (the capture stream while a graph is captured)
        t = sets[i % ROT]
        tpr_amd.solve_batched(t["a"], t["b"], t["c"], t["d"], t["x"], N, S, method=args.method, arith=args.arith,
                              shape=shape)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()

    graph = None
    if not args.eager:
        # the K timed steps are one captured HIP graph: the GPU runs the
        # launches back to back without host launch latency between them
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for i in range(args.steps):
                step(i)
        graph.replay()
        torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    # timed region: synchronize, K steps, synchronize
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    if graph is not None:
        graph.replay()
    else:
        for i in range(args.steps):
            step(i)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1)

    # answers of set 0 for the accuracy check
    step(0)
    torch.cuda.synchronize()
    x0 = sets[0]["x"].cpu().numpy()
    host = {k: sets[0][k].cpu().numpy() for k in ("a", "b", "c", "d", "xt")}
    vendor = None if args.no_vendor else vendor_leg(sets, host)
    other = None if args.eager else other_method_leg(args, sets, shape)
    streams = None if (args.eager or args.no_streams) else batch_stream_leg(args, sets, shape)
    return {"wall_s": wall, "event_ms": ev_ms, "x0": x0, "host": host, "vendor": vendor, "other": other,
            "streams": streams}


def batch_stream_leg(args, sets, shape):
    """The same batches as a stream of independent launches spread over 2
    and 3 streams (default priority), so that one batch's compute tail runs
    under the next one's loads: the throughput of a caller that keeps
    several batches in flight.  A side figure, not the headline (each launch
    still solves one 4096-system batch, but launches overlap, so a kernel
    trace's per-launch durations are longer than the time per batch here).
    Eight buffer sets (671 MB, well past the 256 MB last-level cache), K
    launches captured in one graph with a fork and a join, HIP events
    around the replay, with the library's default (non-temporal) stores of
    x."""
    import torch

    from tpr_amd import batched as B

    nsets = 2 * ROT
    pool = list(sets)
    for r in range(len(pool), nsets):
        t = B.alloc_batch(NSYS, N, with_xt=False)
        B.generate("dominant", 20240 + r, 0, NSYS, N, t["a"], t["b"], t["c"], t["d"], None)
        pool.append(t)
    torch.cuda.synchronize()

    def step(i):
        t = pool[i % nsets]
        B.solve_ex(t["a"], t["b"], t["c"], t["d"], t["x"], N, S, method=args.method, arith=args.arith, shape=shape)

    out = {"buffer_sets": nsets, "x_stores": "non-temporal"}
    for ns in (2, 3):
        side = [torch.cuda.Stream() for _ in range(ns - 1)]
        for i in range(nsets):
            step(i)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cap = torch.cuda.current_stream()
            fork = torch.cuda.Event()
            fork.record(cap)
            for s in side:
                s.wait_event(fork)
            for i in range(args.steps):
                if i % ns == 0:
                    step(i)
                else:
                    with torch.cuda.stream(side[i % ns - 1]):
                        step(i)
            for s in side:
                join = torch.cuda.Event()
                join.record(s)
                cap.wait_event(join)
        g.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        gbps = 20.0 * N * NSYS / (ms / 1e3) / 1e9
        out["%d_streams" % ns] = {"ms_per_batch": round(ms, 5), "systems_per_s": round(NSYS / (ms / 1e3), 1),
                                  "GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBS, 4)}
    return out


def other_method_leg(args, sets, shape):
    """The other method (TPR beside PTPR, or PTPR beside TPR) on the same
    buffer sets with the same protocol (graph of K steps, HIP events): a
    side figure, not the headline."""
    import torch

    import tpr_amd
    method = "tpr" if args.method == "ptpr" else "ptpr"

    def step(i):
        t = sets[i % ROT]
        tpr_amd.solve_batched(t["a"], t["b"], t["c"], t["d"], t["x"], N, S, method=method, arith=args.arith,
                              shape=shape)

    for i in range(ROT):
        step(i)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for i in range(args.steps):
            step(i)
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    g.replay()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbps = 20.0 * N * NSYS / (ms / 1e3) / 1e9
    return {"method": method, "ms_per_step": round(ms, 5), "systems_per_s": round(NSYS / (ms / 1e3), 1),
            "GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBS, 4)}


def vendor_leg(sets, host):
    """rocSPARSE's batched tridiagonal solver (no pivoting) on the same
    buffer sets: the vendor comparator, 20 steps cycling through the sets,
    launched eagerly (a library call is kept out of graph capture), timed
    with HIP events, with its accuracy against the truth on set 0."""
    import numpy as np
    import torch

    from tpr_amd import nstab
    f = nstab._vendor()
    if f is None:
        return {"error": "libtpr_rocsparse.so not built"}

    def step(i):
        t = sets[i % ROT]
        st = f(t["a"].data_ptr(), t["b"].data_ptr(), t["c"].data_ptr(), t["d"].data_ptr(), t["x"].data_ptr(), N,
               NSYS, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        if st != 0:
            raise RuntimeError("rocSPARSE comparator failed (%d)" % st)

    try:
        reps = 20
        for i in range(ROT):
            step(i)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for i in range(reps):
            step(i)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        step(0)
        torch.cuda.synchronize()
        x = sets[0]["x"].cpu().numpy().reshape(NSYS, N).astype(np.float64)
        xt = host["xt"].reshape(NSYS, N).astype(np.float64)
        err = np.linalg.norm(x - xt, axis=1) / np.linalg.norm(xt, axis=1)
        return {"solver": "rocsparse_sgtsv_no_pivot_strided_batch", "ms_per_step": round(ms, 5),
                "systems_per_s": round(NSYS / (ms / 1e3), 1),
                "GBps_algorithmic": round(20 * N * NSYS / (ms / 1e3) / 1e9, 1),
                "max_rel_l2_vs_truth": float(err.max())}
    except Exception as e:  # the comparator is optional: report, do not fail the bench
        return {"error": str(e)}


def host_info():
    """The host the CPU baseline ran on: logical CPUs, the ones this process
    may use, the CPU model and the OpenMP settings."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")}


def cpu_leg(args, host, x_gpu):
    import numpy as np

    import oracle as O

    kind = "reference" if O.ref_available() else "port"
    solve = O.ref_solve_batch if kind == "reference" else O.solve_batch
    threads = O.max_threads()
    n_sample = 512  # systems per timed call (a sample of the batch)
    sl = slice(0, n_sample * N)
    a, b, c, d = (host[k][sl] for k in "abcd")
    solve(args.method, a, b, c, d, N, S, 1)  # warm
    reps = 0
    t0 = time.perf_counter()
    while True:
        x_cpu = solve(args.method, a, b, c, d, N, S, 1)
        reps += 1
        el = time.perf_counter() - t0
        if el >= args.cpu_seconds:
            break
    cpu_sys_s = reps * n_sample / el
    # parity of the GPU answers on the whole set 0 against the CPU PTPR / TPR
    x_ref = solve(args.method, host["a"], host["b"], host["c"], host["d"], N, S, 1)
    xg = x_gpu.reshape(NSYS, N).astype(np.float64)
    xr = x_ref.reshape(NSYS, N).astype(np.float64)
    rel = np.linalg.norm(xg - xr, axis=1) / np.linalg.norm(xr, axis=1)
    bitexact = bool(np.array_equal(x_gpu.view(np.uint32), x_ref.view(np.uint32)))
    xt = host["xt"].reshape(NSYS, N).astype(np.float64)
    err_t = np.linalg.norm(xg - xt, axis=1) / np.linalg.norm(xt, axis=1)
    base = {"value": round(cpu_sys_s, 1), "unit": "systems/s", "cores": threads, "kind": kind,
            "sample": "%d of the %d systems (N=%d, S=%d), solved %d times in %.1f s, one %s solver per thread "
                      "(OpenMP over systems)" % (n_sample, NSYS, N, S, reps, el, args.method.upper()),
            "host": host_info()}
    acc = {"rel_l2_vs_cpu_max": float(rel.max()), "rel_l2_vs_cpu_mean": float(rel.mean()),
           "bit_identical_to_cpu": bitexact, "rel_l2_vs_truth_max": float(err_t.max()), "cpu_check": kind}
    return base, acc


COPY_BYTES = 256 << 20


def copy_rate(steps=20):
    """A plain 256 MiB device-to-device copy (torch's copy_), timed with HIP
    events over `steps` copies: read + write bytes per second.  It is the
    practical HBM ceiling the solver kernels are compared with, and the
    kernel whose known bytes calibrate the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    src = torch.ones(COPY_BYTES // 4, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return 2.0 * COPY_BYTES / (ms / 1e3) / 1e9, ms


def copy_workload(args):
    torch_gbps, ms = copy_rate(args.steps)
    print(json.dumps({"workload": "256 MiB device-to-device copy (torch copy_)", "bytes_read": COPY_BYTES,
                      "bytes_written": COPY_BYTES, "ms_per_copy": round(ms, 5), "GBps": round(torch_gbps, 1),
                      "frac_of_peak": round(torch_gbps / HBM_PEAK_GBS, 4)}))


PREC_NSYS, PREC_N, PREC_S = 256, 65536, 256


def graph_ms(fn, reps, samples=3):
    """Device time per call of fn: `reps` calls captured in one HIP graph,
    replayed once untimed, then `samples` times between HIP events on the
    current stream; the median of the samples."""
    import torch
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fn()
    g.replay()
    torch.cuda.synchronize()
    times = []
    for _ in range(samples):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / reps)
    return sorted(times)[len(times) // 2]


def precision_setup(chunk=0, split_form=0):
    """Config 3's batch: 256 weakly dominant fp32 systems x N=65536 (and its
    fp16-stored coefficients), generated on the device, with a workspace for
    the multi-workgroup path."""
    import torch

    from tpr_amd import batched as B
    nsys, n, s = PREC_NSYS, PREC_N, PREC_S
    t = B.alloc_batch(nsys, n)
    B.generate("weak", 30303, 0, nsys, n, t["a"], t["b"], t["c"], t["d"], t["xt"])
    h16 = {k: t[k].half() for k in "abc"}
    ws = torch.empty(B.workspace_bytes(B.make_desc(t["a"], t["b"], t["c"], t["d"], t["x"], n, s, chunk=chunk)),
                     dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    return {"t": t, "h16": h16, "ws": ws, "chunk": chunk, "form": split_form}


def precision_variants(P, combos, reps, method="ptpr"):
    """Times each (fp16 coefficients?, refinement steps) variant and keeps
    its answers; algorithmic bytes: solve 20 B/eq (14 with fp16 a, b, c),
    each refinement step 24 B/eq (a, b, c, d, x read; x written)."""
    from tpr_amd import batched as B
    nsys, n, s = PREC_NSYS, PREC_N, PREC_S
    t, h16, ws, ch, form = P["t"], P["h16"], P["ws"], P["chunk"], P["form"]

    def solve(coef16, refine_steps):
        src = h16 if coef16 else t
        B.solve_ex(src["a"], src["b"], src["c"], t["d"], t["x"], n, s, method=method, workspace=ws, chunk=ch,
                   split_form=form)
        for _ in range(refine_steps):
            B.solve_ex(t["a"], t["b"], t["c"], t["d"], t["x"], n, s, method=method, refine=True, workspace=ws,
                       chunk=ch, split_form=form)

    import torch
    out = []
    for coef16, steps in combos:
        ms = graph_ms(lambda: solve(coef16, steps), reps)
        solve(coef16, steps)
        torch.cuda.synchronize()
        alg = nsys * n * ((14 if coef16 else 20) + 24 * steps)
        out.append({"name": "%s%s" % ("fp16" if coef16 else "fp32", "+%d refine" % steps if steps else ""),
                    "ms": round(ms, 4), "systems_per_s": round(nsys / (ms / 1e3), 1),
                    "GBps": round(alg / (ms / 1e3) / 1e9, 1), "x": t["x"].cpu().numpy()})
    return out


def precision_errors(P, variants):
    """Each variant's error per system against the host FP64 Thomas solve
    of the fp32 system (the oracle's, on this host)."""
    import numpy as np

    import oracle as O
    t, n = P["t"], PREC_N
    host = {k: t[k].cpu().numpy() for k in "abcd"}
    xf = O.thomas_f64_batch(host["a"], host["b"], host["c"], host["d"], n).reshape(PREC_NSYS, n)
    for v in variants:
        if v.get("x") is None:
            continue
        x = v["x"].reshape(PREC_NSYS, n).astype(np.float64)
        err = np.linalg.norm(x - xf, axis=1) / np.linalg.norm(xf, axis=1)
        v["rel_l2_vs_fp64_thomas_max"] = float(err.max())
        v["rel_l2_vs_fp64_thomas_median"] = float(np.median(err))
    return host, xf


def precision_side(args):
    """Config 3 for the default line: fp32 solve, fp16-stored coefficients
    followed by one FP64-residual refinement step, and fp32 + one step;
    10 graph-replayed calls each; errors per system vs FP64 Thomas."""
    P = precision_setup()
    vs = precision_variants(P, [(False, 0), (True, 1), (False, 1)], 10)
    if not args.no_cpu:
        precision_errors(P, vs)
    for v in vs:
        v.pop("x", None)
    head = vs[0]
    return {"workload": "%d weakly dominant fp32 systems x N=%d, S=%d, multi-workgroup path" %
                        (PREC_NSYS, PREC_N, PREC_S),
            "ms": head["ms"], "systems_per_s": head["systems_per_s"], "GBps": head["GBps"],
            "frac": round(head["GBps"] / HBM_PEAK_GBS, 4), "variants": vs}


def precision_workload(args):
    """256 weakly dominant systems x N=65536, S=256 (m = 256 slices: the
    multi-workgroup path).  Variants: fp32 PTPR, fp16-stored coefficients,
    each followed by 0..2 FP64-residual refinement steps against the fp32
    system; and the same fp32 solve issued one system per launch (the
    reference's CUDA model).  Errors are against the host FP64 Thomas solve,
    per system, in the CPU leg."""
    import numpy as np
    import torch

    from tpr_amd import batched as B

    nsys, n, s = PREC_NSYS, PREC_N, PREC_S
    torch.cuda.set_device(0)
    ch = args.chunk
    P = precision_setup(ch, args.split_form)
    t, ws = P["t"], P["ws"]

    def per_system():
        for k in range(nsys):
            sl = slice(k * n, (k + 1) * n)
            B.solve_ex(t["a"][sl], t["b"][sl], t["c"][sl], t["d"][sl], t["x"][sl], n, s, workspace=ws)

    combos = [(False, 0)] if args.fp32_only else [(c16, st) for c16 in (False, True) for st in (0, 1, 2)]
    variants = precision_variants(P, combos, 10)
    if not args.fp32_only:
        ms = graph_ms(per_system, 2)
        variants.append({"name": "fp32, one system per launch", "ms": round(ms, 4),
                         "systems_per_s": round(nsys / (ms / 1e3), 1),
                         "GBps": round(20 * nsys * n / (ms / 1e3) / 1e9, 1), "x": None})
    head = variants[0]  # the fp32 solve is the line's value; the rest are variants of it
    out = {
        "metric": METRIC, "value": head["systems_per_s"], "unit": "systems/s", "n_gpus": 1, "steps": 10,
        "warmup": 1, "ms_per_step": head["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "%d weakly dominant fp32 systems x N=%d, S=%d (b = |a|+|c|+eps, eps ~ U[1e-4,1e-2]), "
                               "multi-workgroup path, one step = one batch solve" % (nsys, n, s),
                   "method": "ptpr", "arith": "fast", "chunk": ch, "split_form": args.split_form},
        # algorithmic bytes 20 B/eq over the kernels of a solve, per batch
        "roofline": {"bound": "hbm", "achieved": head["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["GBps"] / HBM_PEAK_GBS, 4), "traffic": None},
        "variants": variants,
    }
    # bytes past L2 per fp32 solve from FETCH_SIZE / WRITE_SIZE runs, scaled
    # by the copy calibration
    tf = os.path.join(ROOT, "profiles", "traffic_ptpr_split256x65536.json")
    if ch == 0 and os.path.exists(tf):
        with open(tf) as f:
            out["roofline"]["traffic"] = json.load(f).get("hbm_bytes_per_step")
    if not args.no_cpu:
        import oracle as O
        host, xf = precision_errors(P, variants)
        kind = "reference" if O.ref_available() else "port"
        solvef = O.ref_solve_batch if kind == "reference" else O.solve_batch
        k = 32
        t0 = time.perf_counter()
        xc = solvef("ptpr", host["a"][:k * n], host["b"][:k * n], host["c"][:k * n], host["d"][:k * n], n, s, 1)
        el = time.perf_counter() - t0
        err = np.linalg.norm(xc.reshape(k, n) - xf[:k], axis=1) / np.linalg.norm(xf[:k], axis=1)
        out["cpu_baseline"] = {"value": round(k / el, 2), "unit": "systems/s", "cores": O.max_threads(),
                               "kind": kind, "sample": "%d of the %d systems, one PTPR per thread" % (k, nsys),
                               "rel_l2_vs_fp64_thomas_max": float(err.max()), "host": host_info()}
    for v in variants:
        v.pop("x", None)
    print(json.dumps(out))


UNEVEN_NSYS = 20000


def uneven_setup(args):
    """Config 2's batch: 20000 systems of N = 2^U{8..14} (log-uniform over
    the powers of two), diagonally dominant generator, in one uneven-batch
    layout; plans for the committed autotuned schedule and for the uniform
    s = min(N, 256) schedule."""
    import torch

    from tpr_amd import autotune
    from tpr_amd.uneven import UnevenPlan, log_uniform_sizes

    sizes = log_uniform_sizes(UNEVEN_NSYS, 8, 14, seed=2020)
    if args.retune or not os.path.exists(autotune.DEFAULT_PATH):
        table = autotune.tune(method=args.method)
        tuned_from = "tuned in this run"
    else:
        table = autotune.load()
        tuned_from = os.path.relpath(autotune.DEFAULT_PATH, ROOT)
    sched = autotune.schedule_of(table)
    plans = {"autotuned": UnevenPlan(sizes, args.method, args.arith, sched),
             "uniform s=256 schedule": UnevenPlan(sizes, args.method, args.arith, None)}
    tot = plans["autotuned"].equations
    a, b, c, d, x, xt = (torch.empty(tot, device="cuda") for _ in range(6))
    plans["autotuned"].generate("dominant", 4242, 0, a, b, c, d, xt)
    torch.cuda.synchronize()
    return {"sizes": sizes, "sched": sched, "tuned_from": tuned_from, "plans": plans, "tot": tot,
            "arr": {"a": a, "b": b, "c": c, "d": d, "x": x, "xt": xt}}


def uneven_time(U, runs, steps):
    """Per-step device time of each run (graph of `steps` calls), as GB/s at
    20 B per equation and as a fraction of the HBM peak."""
    res = {}
    for name, fn in runs.items():
        ms = graph_ms(fn, steps)
        gbps = 20.0 * U["tot"] / (ms / 1e3) / 1e9
        res[name] = {"ms_per_step": round(ms, 4), "systems_per_s": round(UNEVEN_NSYS / (ms / 1e3), 1),
                     "GBps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBS, 4)}
    return res


def uneven_accuracy(args, U, count):
    """The autotuned answers of `count` systems drawn at random, against the
    oracle at the same s per class and against the generator's truth."""
    import numpy as np
    import torch

    import oracle as O
    p = U["plans"]["autotuned"]
    A = U["arr"]
    p.solve(A["a"], A["b"], A["c"], A["d"], A["x"])
    torch.cuda.synchronize()
    h = {k: v.cpu().numpy() for k, v in A.items()}
    pick = np.random.default_rng(7).choice(UNEVEN_NSYS, count, replace=False)
    worst_o = worst_t = 0.0
    for k in pick:
        lo, hi = p.offset(int(k)), p.offset(int(k) + 1)
        n = int(U["sizes"][k])
        ref = O.solve_batch(args.method, h["a"][lo:hi], h["b"][lo:hi], h["c"][lo:hi], h["d"][lo:hi], n,
                            U["sched"].get(n, (min(n, 256),))[0])
        worst_o = max(worst_o, O.rel_l2(h["x"][lo:hi], ref))
        worst_t = max(worst_t, O.rel_l2(h["x"][lo:hi], h["xt"][lo:hi]))
    return {"systems_checked": int(count), "max_rel_l2_vs_oracle_same_s": worst_o, "max_rel_l2_vs_truth": worst_t,
            "tolerance": 1e-5}


def uneven_side(args):
    """Config 2 for the default line: the persistent launch with the
    autotuned and the uniform schedules, 10 graph-replayed steps each;
    accuracy of 200 systems against the oracle."""
    U = uneven_setup(args)
    A = U["arr"]
    runs = {name: (lambda p=p: p.solve(A["a"], A["b"], A["c"], A["d"], A["x"])) for name, p in U["plans"].items()}
    res = uneven_time(U, runs, 10)
    best = res["autotuned"]
    out = {"workload": "%d fp32 systems, N = 2^U{8..14} (%d equations), one persistent launch" %
                       (UNEVEN_NSYS, U["tot"]),
           "ms": best["ms_per_step"], "systems_per_s": best["systems_per_s"], "GBps": best["GBps"],
           "frac": best["hbm_frac"], "schedules": res, "schedule_source": U["tuned_from"]}
    if not args.no_cpu:
        out["accuracy"] = uneven_accuracy(args, U, 200)
    for p in U["plans"].values():
        p.close()
    return out


def uneven_workload(args):
    """20000 systems, N = 2^U{8..14} (log-uniform over the powers of two),
    diagonally dominant generator, one persistent launch per step, replayed
    from a captured graph.  Schedules compared on the same batch:
    * autotuned: s and slice shape per size class from tpr_amd.autotune;
    * uniform schedule: s = min(N, 256) for every class (the 4096 x N=1024
      batch's slice size), same kernel;
    * grouped: the systems sorted by size and one batched launch per size
      class at s = min(N, 256) (the plain batched kernel, 7 launches).
    Algorithmic bytes: 20 B per equation, ~92.9 M equations per step."""
    import numpy as np
    import torch

    from tpr_amd import batched as B
    from tpr_amd.uneven import UnevenPlan

    torch.cuda.set_device(0)
    U = uneven_setup(args)
    sizes, sched, A, tot = U["sizes"], U["sched"], U["arr"], U["tot"]
    nsys = UNEVEN_NSYS
    # class-sorted copy for the grouped launches (stable order inside a class)
    order = np.argsort(sizes, kind="stable")
    srt = UnevenPlan(sizes[order], args.method, args.arith, None)
    pa = U["plans"]["autotuned"]
    idx = np.concatenate([np.arange(pa.offset(k), pa.offset(k + 1)) for k in order])
    idx_t = torch.from_numpy(idx).cuda()
    sa, sb, sc, sd = (A[v][idx_t].contiguous() for v in "abcd")
    sx = torch.empty_like(sa)
    groups = []
    for n in sorted(set(sizes.tolist())):
        ks = np.nonzero(sizes[order] == n)[0]
        lo, hi = srt.offset(int(ks[0])), srt.offset(int(ks[-1]) + 1)
        groups.append((n, lo, hi))
    torch.cuda.synchronize()

    def grouped():
        for n, lo, hi in groups:
            B.solve_batched(sa[lo:hi], sb[lo:hi], sc[lo:hi], sd[lo:hi], sx[lo:hi], n, min(n, 256),
                            method=args.method, arith=args.arith)

    runs = {name: (lambda p=p: p.solve(A["a"], A["b"], A["c"], A["d"], A["x"])) for name, p in U["plans"].items()}
    runs["grouped launches"] = grouped
    steps = args.steps
    res = uneven_time(U, runs, steps)
    best = res["autotuned"]
    out = {
        "metric": METRIC, "value": best["systems_per_s"], "unit": "systems/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "uneven batch: %d fp32 systems, N = 2^U{8..14} (%d equations), diagonally dominant "
                               "generator, one persistent launch per step replayed from a HIP graph" % (nsys, tot),
                   "method": args.method, "arith": args.arith, "schedule": {str(k): v for k, v in sched.items()},
                   "schedule_source": U["tuned_from"],
                   "class_counts": {str(n): int((sizes == n).sum()) for n in sorted(set(sizes.tolist()))}},
        "roofline": {"bound": "hbm", "achieved": best["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": best["hbm_frac"], "traffic": None},
        "schedules": res,
    }
    # HBM bytes per persistent launch from the FETCH_SIZE / WRITE_SIZE runs of
    # this workload, scaled by the copy calibration (fast arithmetic)
    tf = os.path.join(ROOT, "profiles", "traffic_%s_uneven20000.json" % args.method)
    if args.arith == "fast" and os.path.exists(tf):
        with open(tf) as f:
            out["roofline"]["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    if not args.no_cpu:
        import oracle as O
        out["accuracy"] = uneven_accuracy(args, U, 400)
        kind = "reference" if O.ref_available() else "port"
        solvef = O.ref_solve_batch if kind == "reference" else O.solve_batch
        sample = []
        for n, lo, hi in groups:
            cnt = min(16, (hi - lo) // n)
            hs = slice(lo, lo + cnt * n)
            sample.append((n, cnt, [v[hs].cpu().numpy() for v in (sa, sb, sc, sd)]))
        t0 = time.perf_counter()
        done = eqs = 0
        while time.perf_counter() - t0 < args.cpu_seconds:
            for n, cnt, (ha, hb, hc, hd) in sample:
                solvef(args.method, ha, hb, hc, hd, n, min(n, 256), 1)
                done += cnt
                eqs += cnt * n
        el = time.perf_counter() - t0
        # the sample has 16 systems of every class; weight to the batch's mix
        mean_n = tot / nsys
        out["cpu_baseline"] = {"value": round(eqs / el / mean_n, 1), "unit": "systems/s", "cores": O.max_threads(),
                               "kind": kind,
                               "sample": "%d systems (16 per size class per pass) in %.1f s, one %s per thread; "
                                         "systems/s = equations/s over the batch's mean N" %
                                         (done, el, args.method.upper()),
                               "host": host_info()}
    print(json.dumps(out))


SERVE_WINDOW = {"min_batch_systems": 1024, "max_wait_us": 100.0}


def serve_clients(args, seconds, closed_requests):
    """The serving loop driven by the C++ client tpr_amd/tpr_serve_bench in
    child processes (one client thread; the Python client's own call
    overhead would otherwise bound the saturated rate): a Poisson stream of
    args.rate requests per second for `seconds`, then the same request mix
    closed-loop with 256 outstanding.  Each request brings a matrix of
    N = 2^U{9..15} equations (diagonally dominant generator, generated on
    the device and resident in HBM) and 1..64 right-hand sides d ~ U(-1,1);
    64 GB budget.  The Poisson stream uses the fixed batching window (1024
    right-hand sides / 100 us while other batches run); the closed loop the
    adaptive one (capped at 100 us), whose arrival-rate rule keeps batches
    large when requests arrive faster than the server launches them, which
    makes the saturated rate reproducible (the fixed window's depends on the
    batch sizes a run settles into).  Run before this process touches the
    GPU."""
    import subprocess
    exe = os.path.join(ROOT, "tpr_amd", "tpr_serve_bench")
    if not os.path.exists(exe):
        raise RuntimeError("tpr_amd/tpr_serve_bench is not built: run `python -m tpr_amd.build`")
    common = ["--method", args.method, "--arith", args.arith, "--seed", "4044", "--budget-gb", "64",
              "--min-batch", str(SERVE_WINDOW["min_batch_systems"]), "--max-wait-us",
              str(SERVE_WINDOW["max_wait_us"])]

    def client(extra):
        p = subprocess.run([exe] + common + extra, capture_output=True, text=True, timeout=600)
        if p.returncode != 0:
            raise RuntimeError("tpr_serve_bench failed: %s" % p.stderr[-2000:])
        return json.loads(p.stdout.strip().splitlines()[-1])

    poisson = client(["--rate", str(args.rate), "--seconds", str(seconds)])
    closed = client(["--rate", "0", "--requests", str(closed_requests), "--outstanding", "256", "--adaptive", "1"])
    return poisson, closed


def serve_request_mix(count):
    import numpy as np
    rng = np.random.default_rng(4044)
    sizes = (1 << rng.integers(9, 16, count)).astype(np.int64)
    nrhs = rng.integers(1, 65, count).astype(np.int64)
    return sizes, nrhs


def serve_per_system(args, count):
    """The plain per-system GPU solve of the same kind of requests: one
    tprh_solve_ex launch per right-hand side (the reference's CUDA model),
    descriptors, workspaces and inputs prepared beforehand, only the
    launches timed (wall clock, one final synchronisation)."""
    import ctypes

    import torch

    from tpr_amd import _lib
    from tpr_amd import batched as B
    sizes, nrhs = serve_request_mix(count)
    prepared = []
    for k in range(len(sizes)):
        n = int(sizes[k])
        t = B.alloc_batch(1, n, with_xt=True)
        B.generate("dominant", 4044, k, 1, n, t["a"], t["b"], t["c"], t["d"], t["xt"])
        desc = B.make_desc(t["a"], t["b"], t["c"], t["d"], t["x"], n, min(n, 256), args.method, args.arith)
        need = B.workspace_bytes(desc)
        ws = torch.empty(max(need, 1), dtype=torch.uint8, device="cuda")
        desc.workspace, desc.workspace_bytes = ws.data_ptr(), need
        prepared.append((desc, int(nrhs[k]), t, ws))
    torch.cuda.synchronize()
    solve_ex = _lib.lib().tprh_solve_ex
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    t0 = time.perf_counter()
    tot_sys = 0
    for desc, r, _, _ in prepared:
        ref = ctypes.byref(desc)
        for j in range(r):
            if solve_ex(ref, stream) != 0:
                raise RuntimeError("tprh_solve_ex failed")
        tot_sys += r
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"requests": int(len(sizes)), "systems": int(tot_sys), "systems_per_s": round(tot_sys / wall, 1)}


def serve_accuracy(args, count):
    """Every sixth request of the mix (up to two right-hand sides each)
    served through the Python binding, against the oracle at s = min(N, 256)."""
    import oracle as O
    from tpr_amd.serve import Server
    sizes, nrhs = serve_request_mix(count)
    worst = 0.0
    with Server(8 << 30, method=args.method, arith=args.arith) as srv:
        reqs = []
        for k in range(0, len(sizes), 6):
            m = srv.matrix_generate(int(sizes[k]), seed=4044, sys=k)
            reqs.append((k, m, srv.submit(m, int(min(nrhs[k], 2)), rhs_seed=4044)))
        for k, m, r in reqs:
            srv.wait(r)
            x = srv.result(r).cpu().numpy()
            n = int(sizes[k])
            a, b, c, _, _ = O.gen_system(4044, k, n)
            for j in range(x.shape[0]):
                ref = O.solve_batch(args.method, a, b, c, O.gen_rhs(4044, k, j, n), n, min(n, 256))
                worst = max(worst, O.rel_l2(x[j], ref))
            srv.release(r)
            srv.matrix_release(m)
    return {"requests_checked": len(reqs), "max_rel_l2_vs_oracle": worst, "tolerance": 1e-5}


def serve_summary(args, poisson, closed, base):
    return {"workload": "Poisson arrivals at %.0f req/s for %.1f s (%d requests), N = 2^U{9..15}, 1..64 right-hand "
                        "sides per matrix, 64 GB budget" % (args.rate, poisson["wall_s"], poisson["requests"]),
            "systems_per_s": poisson["systems_per_s"], "latency_ms": poisson["latency_ms"],
            "latency_from_arrival_ms": poisson.get("latency_from_arrival_ms"),
            "saturated_systems_per_s": closed["systems_per_s"], "saturated_latency_ms": closed["latency_ms"],
            "saturated_window": "adaptive",
            "per_system_launches_systems_per_s": base["systems_per_s"],
            "served_vs_per_system": round(closed["systems_per_s"] / base["systems_per_s"], 2),
            "max_inflight_seen": poisson.get("max_inflight_seen")}


def serve_workload(args):
    """The serving loop under load (serve_clients): Poisson stream and the
    saturated closed loop; beside them the plain per-system GPU solve (one
    launch per right-hand side) on the first 300 requests of the mix; and
    the answers of a sample of requests against the oracle."""
    poisson, closed = serve_clients(args, args.seconds, 20000)

    import torch
    torch.cuda.set_device(0)
    base = serve_per_system(args, 300)
    out = {
        "metric": METRIC, "value": poisson["systems_per_s"], "unit": "systems/s", "n_gpus": 1,
        "steps": poisson["requests"], "warmup": 7,
        "ms_per_step": round(1e3 * poisson["wall_s"] / poisson["requests"], 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "serving loop: Poisson arrivals at %.0f req/s for %.1f s (%d requests), N = 2^U{9..15}, "
                               "1..64 right-hand sides per matrix, matrices generated and resident in HBM, "
                               "64 GB budget, C++ client" % (args.rate, args.seconds, poisson["requests"]),
                   "method": args.method, "arith": args.arith, "max_inflight": 4, "batching_window": SERVE_WINDOW,
                   "saturated_window": "adaptive, max_wait_us %.0f" % SERVE_WINDOW["max_wait_us"]},
        "latency_ms": poisson["latency_ms"],
        "latency_from_arrival_ms": poisson.get("latency_from_arrival_ms"),
        # latency-bound at the offered load; the solve traffic (12 B per matrix
        # equation, 8 B per right-hand-side equation) against the HBM peak
        "roofline": {"bound": "hbm", "achieved": poisson["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(poisson["GBps"] / HBM_PEAK_GBS, 6), "traffic": None},
        "served": poisson,
        "saturated": closed,
        "per_system_launches": base,
    }
    if not args.no_cpu:
        out["accuracy"] = serve_accuracy(args, 300)
    print(json.dumps(out))


def config0_workload(args):
    """The reference's documented invocation `tpr_pm 2048 512 1000 PTPR`
    (README.md:28-30 of the reference): one fp32 system N=2048, S=512 from
    the diagonally dominant generator (fixed seed), 1000 timed solves.  GPU,
    two ways: the device-resident solve (tprh_solve_batched on one system,
    HIP events around 1000 launches replayed from a captured HIP graph, and
    the same launches issued one by one) and the host-pointer entry point
    tprh_ptpr_cu (the contract of the reference's ptpr_cu: host arrays with
    the diagonal divided out, synchronous; wall time per call, PCIe
    included).  CPU baseline: the reference's PTPR class, 1000 solves one
    after another with its OpenMP over the m = 4 slices (src/pm.cpp:184-191;
    the input copy each solve needs is inside the timing, where the
    reference regenerates it untimed).  rel-L2 against the generator's truth
    for each."""
    import numpy as np
    import torch

    import tpr_amd
    from tpr_amd import batched as B

    n, s, iters = 2048, 512, 1000
    torch.cuda.set_device(0)
    t = B.alloc_batch(1, n)
    B.generate("dominant", 2048, 0, 1, n, t["a"], t["b"], t["c"], t["d"], t["xt"])
    torch.cuda.synchronize()
    def solve():
        B.solve_batched(t["a"], t["b"], t["c"], t["d"], t["x"], n, s, method=args.method, arith=args.arith)

    for _ in range(20):
        solve()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(run):
        torch.cuda.synchronize()
        e0.record()
        run()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / iters

    # launched one by one from Python, and the same 1000 launches replayed
    # from a captured HIP graph (no host launch latency between solves)
    eager_ms = timed(lambda: [solve() for _ in range(iters)])
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(iters):
            solve()
    graph.replay()
    dev_ms = timed(graph.replay) if not args.eager else eager_ms
    xt = t["xt"].cpu().numpy().astype(np.float64)
    host = {k: t[k].cpu().numpy() for k in "abcd"}

    def rel(x):
        return float(np.linalg.norm(x.astype(np.float64) - xt) / np.linalg.norm(xt))

    err_dev = rel(t["x"].cpu().numpy())
    # the host-pointer path takes the diagonal divided out, as the
    # reference's 4-argument setter leaves the arrays (fp32 divisions)
    na, nc, nd = (host[k] / host["b"] for k in "acd")
    x = np.zeros(n, np.float32)
    cu = tpr_amd.ptpr_cu if args.method == "ptpr" else tpr_amd.tpr_cu
    for _ in range(20):
        cu(na, nc, nd, x, n, s)
    t0 = time.perf_counter()
    for _ in range(iters):
        cu(na, nc, nd, x, n, s)
    cu_ms = (time.perf_counter() - t0) * 1e3 / iters
    err_cu = rel(x)
    # the same entry point called straight through ctypes with the pointers
    # prepared once (what a C caller pays), and for comparison a
    # device-resident launch followed by a stream synchronisation, the floor
    # of any synchronous call
    from tpr_amd import _lib
    L = _lib.lib()
    entry = L.tprh_ptpr_cu if args.method == "ptpr" else L.tprh_tpr_cu
    P = [v.ctypes.data_as(ctypes.c_void_p) for v in (na, nc, nd, x)]

    def per_call_us(fn, k=2000):
        for _ in range(200):
            fn()
        t1 = time.perf_counter()
        for _ in range(k):
            fn()
        return (time.perf_counter() - t1) / k * 1e6

    c_us = per_call_us(lambda: entry(*P, n, s))
    err_c = rel(x)
    D = [ctypes.c_void_p(t[k].data_ptr()) for k in "abcdx"]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    m_id = 0 if args.method == "ptpr" else 1

    def launch_sync():
        L.tprh_solve_batched(m_id, 1, *D, n, s, 1, stream)
        torch.cuda.synchronize()

    floor_us = per_call_us(launch_sync)
    out = {
        "metric": METRIC, "value": round(1e3 / dev_ms, 1), "unit": "systems/s", "n_gpus": 1,
        "steps": iters, "warmup": 20, "ms_per_step": round(dev_ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "the reference's documented run: one fp32 system N=2048, S=512, diagonally "
                               "dominant generator (seed 2048), %d solves" % iters,
                   "n": n, "s": s, "method": args.method, "arith": args.arith, "graph": not args.eager},
        "eager_ms_per_solve": round(eager_ms, 5),
        # one system moves 40 KB: the solve is latency-bound, far below any
        # bandwidth roof
        "roofline": {"bound": "hbm", "achieved": round(20 * n / (dev_ms / 1e3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(20 * n / (dev_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 6),
                     "traffic": None},
        "host_pointer_call": {"entry": "tprh_%s_cu" % args.method, "ms_per_call": round(cu_ms, 4),
                              "ms_per_call_from_c": round(c_us / 1e3, 4),
                              "device_launch_and_sync_ms": round(floor_us / 1e3, 4),
                              "rel_l2_vs_truth": err_cu, "rel_l2_vs_truth_from_c": err_c,
                              "what": "ms_per_call through the Python mirror (numpy checks included); "
                                      "ms_per_call_from_c straight through the C ABI with prepared pointers; "
                                      "device_launch_and_sync_ms: one device-resident exact launch and a stream "
                                      "synchronisation, the floor of a synchronous call"},
        "accuracy": {"rel_l2_vs_truth": err_dev},
    }
    if not args.no_cpu:
        import oracle as O
        kind = "reference" if O.ref_available() else "port"
        solvef = O.ref_solve_batch if kind == "reference" else O.solve_batch
        reps = {k: np.tile(host[k], iters) for k in "abcd"}
        solvef(args.method, host["a"], host["b"], host["c"], host["d"], n, s, 0)  # warm
        t0 = time.perf_counter()
        xc = solvef(args.method, reps["a"], reps["b"], reps["c"], reps["d"], n, s, 0)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(iters / el, 1), "unit": "systems/s", "cores": O.max_threads(),
                               "kind": kind, "ms_per_solve": round(el * 1e3 / iters, 4),
                               "sample": "the same system solved %d times one after another (OpenMP over the "
                                         "m = %d slices of each solve)" % (iters, n // s),
                               "rel_l2_vs_truth": rel(xc[:n]), "host": host_info()}
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (systems are independent and "
              "would be sharded by count across GPUs); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "precision":
        precision_workload(args)
        return
    if args.workload == "uneven":
        uneven_workload(args)
        return
    if args.workload == "serve":
        serve_workload(args)
        return
    if args.workload == "copy":
        copy_workload(args)
        return
    if args.workload == "config0":
        config0_workload(args)
        return
    side = not args.no_side and args.shape is None and not args.eager
    if side:
        # the serving clients run in child processes before this one touches
        # the GPU
        poisson, closed = serve_clients(args, 2.0, 20000)
    g = gpu_leg(args)
    ms_per_step = g["event_ms"] / args.steps
    value = NSYS / (ms_per_step / 1e3)
    alg_bytes = 20 * N * NSYS
    # every step is one launch of the solver kernel, so the kernel's average
    # duration over the timed region is the event time divided by K
    kernel_ms = ms_per_step
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s_batch4096x1024.json" % args.method)
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "systems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched %s: %d fp32 systems x N=%d, S=%d, diagonally dominant generator, "
                               "inputs resident in HBM" % (args.method.upper(), NSYS, N, S),
                   "nsys": NSYS, "n": N, "s": S, "method": args.method, "arith": args.arith,
                   "graph": not args.eager, "buffer_sets": ROT, "shape": args.shape},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "kernel_ms_avg": round(kernel_ms, 5),
        "wall_s": round(g["wall_s"], 4),
    }
    cg, _ = copy_rate()
    out["copy_ceiling"] = {"GBps": round(cg, 1), "frac": round(achieved / cg, 4),
                           "what": "256 MiB device copy measured in this run (read + write bytes / time)"}
    # the reference's analytic operation count (src/ptpr.cpp:104-106,
    # src/tpr.cpp:96-98), as the solver class's solve() returns it
    import tpr_amd
    solver = (tpr_amd.PTPR if args.method == "ptpr" else tpr_amd.TPR)(N, S)
    h = g["host"]
    solver.set_tridiagonal_system(*(h[k][:N] for k in "abcd"))
    per_sys = int(solver.solve())
    out["analytic_flops"] = {"per_system": per_sys, "GFLOPs": round(per_sys * value / 1e9, 1),
                             "source": "the reference's count returned by solve(), times systems/s"}
    if g.get("other") is not None:
        out["other_method"] = g["other"]
    if g.get("streams") is not None:
        out["batch_stream"] = g["streams"]
    if g.get("vendor") is not None:
        out["vendor_comparator"] = g["vendor"]
    if not args.no_cpu:
        base, acc = cpu_leg(args, g["host"], g["x0"])
        out["cpu_baseline"] = base
        out["accuracy"] = acc
    if side:
        # the rest of the metric, from short runs: the serving loop's
        # latency, the 20000-system uneven batch and the precision variants
        # of 256 x N=65536
        sv = serve_summary(args, poisson, closed, serve_per_system(args, 100))
        out["latency_ms"] = sv["latency_ms"]
        out["serve"] = sv
        out["uneven"] = uneven_side(args)
        out["precision"] = precision_side(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
