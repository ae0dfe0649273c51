"""Per-stage timing of the batched fast-PTPR solve on the GPU, from the
diagnostic build libtpr_diag.so (tpr_amd/diag/stage_clock.hip): the shipped
kernel's device functions with the GPU's 100 MHz clock read between stages,
six stamps per wavefront.

    python -m tpr_amd.stage_clock [--nsys 4096] [--n 1024] [--s 256]

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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nsys", type=int, default=4096)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--s", type=int, default=256)
    args = ap.parse_args(argv)
    print(json.dumps(run(args.nsys, args.n, args.s)))


if __name__ == "__main__":
    main()
