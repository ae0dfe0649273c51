"""Per-stage timing of the batched fast-PTPR solve on the GPU, from the
diagnostic build libtpr_diag.so (tpr_amd/diag/stage_clock.hip): the shipped
kernel's device functions with the GPU's 100 MHz clock read between stages,
six stamps per wavefront.

    python -m tpr_amd.stage_clock [--nsys 4096] [--n 1024] [--s 256]
    python -m tpr_amd.stage_clock --split [--nsys 256] [--n 65536] [--lag 0]

Prints one JSON line: the kernel's span, per-stage durations over the
wavefronts (median, 10th and 90th percentile, microseconds), and when the
wavefronts finish loading and finish storing relative to the kernel's start
-- the picture of how loads and arithmetic overlap across the batch.  The
answers are checked bit for bit against tprh_solve_batched's.
"""
import argparse
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DIAG = os.path.join(HERE, "libtpr_diag.so")
STAGES = ["load", "stage1", "reduced", "stage3", "store"]
TICK_US = 0.01  # s_memrealtime runs at 100 MHz


def run(nsys=4096, n=1024, s=256, seed=20240, reps=3):
    import torch

    from . import _lib
    from . import batched as B
    _lib.lib()  # loads torch's HIP runtime first
    if not os.path.exists(DIAG):
        raise RuntimeError("tpr_amd/libtpr_diag.so is not built: run `python -m tpr_amd.build`")
    L = ctypes.CDLL(DIAG)
    L.tprh_diag_ptpr_stage_clock.restype = ctypes.c_int
    t = B.alloc_batch(nsys, n, with_xt=False)
    B.generate("dominant", seed, 0, nsys, n, t["a"], t["b"], t["c"], t["d"])
    x_ref = torch.empty_like(t["x"])
    B.solve_batched(t["a"], t["b"], t["c"], t["d"], x_ref, n, s, method="ptpr", arith="fast", shape=(s // 16, 16))
    m = n // s
    waves = -(-nsys // (256 // (m * 16))) * 4
    stamps = torch.zeros(waves * 6, dtype=torch.int64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = [ctypes.c_void_p(t[k].data_ptr()) for k in "abcdx"]
    runs = []
    for _ in range(reps):
        torch.cuda.synchronize()
        e = L.tprh_diag_ptpr_stage_clock(*ptr, n, s, nsys, ctypes.c_void_p(stamps.data_ptr()), stream)
        if e != 0:
            raise RuntimeError("diagnostic launch failed (%d)" % e)
        torch.cuda.synchronize()
        runs.append(stamps.cpu().numpy().reshape(waves, 6).astype(np.float64))
    same = bool(torch.equal(t["x"].view(torch.int32), x_ref.view(torch.int32)))
    st = runs[-1]
    st = st[st[:, 0] > 0]
    t0 = st[:, 0].min()
    out = {"workload": "%d fast PTPR systems x N=%d, S=%d (16 x %d slices)" % (nsys, n, s, s // 16),
           "waves": int(st.shape[0]), "bit_identical_to_shipped_kernel": same,
           "kernel_span_us": round((st[:, 5].max() - t0) * TICK_US, 2),
           "span_us_per_run": [round((r[:, 5].max() - r[:, 0].min()) * TICK_US, 2) for r in runs]}
    for k, name in enumerate(STAGES):
        d = (st[:, k + 1] - st[:, k]) * TICK_US
        out[name + "_us"] = {"median": round(float(np.median(d)), 3), "p10": round(float(np.percentile(d, 10)), 3),
                             "p90": round(float(np.percentile(d, 90)), 3)}
    for k, name in ((0, "start"), (1, "loads_complete"), (2, "stage1_done"), (5, "stores_complete")):
        d = (st[:, k] - t0) * TICK_US
        out[name + "_at_us"] = {"first": round(float(d.min()), 2), "median": round(float(np.median(d)), 2),
                                "last": round(float(d.max()), 2)}
    return out


SPLIT_KINDS = {0: ("rows_task_us", ["loads_and_arrival", "stage1", "rows_issued", "-", "next_ticket"]),
               1: ("reduced_task_us", ["arrival_and_arrivals_seen", "solve", "flag_raised", "-", "next_ticket"]),
               2: ("stage3_task_us", ["loads_and_arrival", "stage1", "flag_seen", "xed_read",
                                      "stage3_store_next_ticket"])}


def _dist(v):
    v = np.asarray(v, np.float64) * TICK_US
    if v.size == 0:
        return None
    return {"median": round(float(np.median(v)), 3), "p10": round(float(np.percentile(v, 10)), 3),
            "p90": round(float(np.percentile(v, 90)), 3), "max": round(float(v.max()), 3)}


def run_split(nsys=256, n=65536, lag=0, seed=20241, reps=3):
    """The multi-workgroup path's pipelined form (split_form 3), persistent,
    with six stamps per task: where each kind of task spends its time."""
    import torch

    from . import _lib
    from . import batched as B
    _lib.lib()
    if not os.path.exists(DIAG):
        raise RuntimeError("tpr_amd/libtpr_diag.so is not built: run `python -m tpr_amd.build`")
    L = ctypes.CDLL(DIAG)
    L.tprh_diag_split_clock.restype = ctypes.c_int
    L.tprh_diag_split_tickets.restype = ctypes.c_longlong
    ntk = L.tprh_diag_split_tickets(n, nsys)
    if ntk <= 0:
        raise RuntimeError("tprh_diag_split_tickets failed")
    t = B.alloc_batch(nsys, n, with_xt=False)
    B.generate("weak", seed, 0, nsys, n, t["a"], t["b"], t["c"], t["d"])
    x_ref = torch.empty_like(t["x"])
    B.solve_ex(t["a"], t["b"], t["c"], t["d"], x_ref, n, 256, method="ptpr", arith="fast", chunk=nsys,
               split_form=3)
    ws = torch.empty(B.workspace_bytes(B.make_desc(t["a"], t["b"], t["c"], t["d"], x_ref, n, 256, chunk=nsys)),
                     dtype=torch.uint8, device="cuda")
    stamps = torch.zeros(ntk * 8, dtype=torch.int64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = [ctypes.c_void_p(t[k].data_ptr()) for k in "abcdx"]
    runs = []
    for _ in range(reps):
        torch.cuda.synchronize()
        e = L.tprh_diag_split_clock(*ptr, n, nsys, ctypes.c_void_p(ws.data_ptr()), int(lag),
                                    ctypes.c_void_p(stamps.data_ptr()), stream)
        if e != 0:
            raise RuntimeError("diagnostic launch failed (%d)" % e)
        torch.cuda.synchronize()
        runs.append(stamps.cpu().numpy().reshape(ntk, 8))
    same = bool(torch.equal(t["x"].view(torch.int32), x_ref.view(torch.int32)))
    st = runs[-1]
    kind = (st[:, 6] & 0xFF).astype(int)
    wg = (st[:, 6] >> 8).astype(int)
    ts = st[:, :6].astype(np.float64)
    t0 = ts[:, 0].min()
    out = {"workload": "%d weak fast-PTPR systems x N=%d, S=256 (16 x 16 slices), pipelined split form" % (nsys, n),
           "tickets": int(ntk), "workgroups": int(wg.max() + 1), "bit_identical_to_shipped_kernel": same,
           "kernel_span_us": round((ts[:, 5].max() - t0) * TICK_US, 2),
           "span_us_per_run": [round((r[:, 5].max() - r[:, 0].min()) * TICK_US, 2) for r in runs],
           "tasks": {SPLIT_KINDS[k][0]: int((kind == k).sum()) for k in SPLIT_KINDS}}
    for k, (name, steps) in SPLIT_KINDS.items():
        sel = ts[kind == k]
        if sel.shape[0] == 0:
            continue
        d = {"whole": _dist(sel[:, 5] - sel[:, 0])}
        for c, step in enumerate(steps):
            if step != "-":
                d[step] = _dist(sel[:, c + 1] - sel[:, c])
        out[name] = d
    # busy time of all workgroups against the span: the share spent per kind
    span = ts[:, 5].max() - t0
    out["share_of_workgroup_time"] = {SPLIT_KINDS[k][0]: round(float((ts[kind == k, 5] - ts[kind == k, 0]).sum() /
                                                                     (span * (wg.max() + 1))), 3)
                                      for k in SPLIT_KINDS}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nsys", type=int, default=None)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--s", type=int, default=256)
    ap.add_argument("--split", action="store_true",
                    help="the multi-workgroup path's pipelined form (default 256 systems x N=65536)")
    ap.add_argument("--lag", type=int, default=0, help="--split: lag in workgroups (0: the library's)")
    args = ap.parse_args(argv)
    if args.split:
        print(json.dumps(run_split(args.nsys or 256, args.n or 65536, args.lag)))
    else:
        print(json.dumps(run(args.nsys or 4096, args.n or 1024, args.s)))


if __name__ == "__main__":
    main()
