"""Accuracy report in the reference nstab driver's protocol and format
(src/numerical_stability.cpp:143-195), on the MI355X solvers.

    python -m tpr_amd.nstab INPUTS [--fast] [--out report.csv]

INPUTS is either a directory holding the reference's dense-CSV inputs
(eq*.txt with n = 512 rows of n + 1 comma-separated values, xt.txt) or an
.npz with the same systems already reduced to bands (arrays eqK of shape
(4, n): a, diag, c, rhs in fp32; and xt).  For every system, in glob order
(eq1, eq10, eq11, ..., eq2, eq20, eq3, ...), one CSV line ``file,solver,N,S,L2``
for PTPR and TPR at s = 16 .. n, PCR (S = 0) and, when libtpr_rocsparse.so
is built, the rocSPARSE comparator REFERENCE_ROCSPARSE (S = 0).

L2 is ||x - xt|| / ||xt|| with the reference's mixed precision (differences
and squares in fp32, sums in fp64; src/numerical_stability.cpp:100-115).
"""
import argparse
import ctypes
import glob
import os
import sys

import numpy as np

N = 512


def read_dense_csv(path, n=N):
    """Bands (a, diag, c, rhs) of one input file, read in double and cast to
    fp32 as src/system.cpp:54-101 does (a[0] = c[n-1] = 0)."""
    rows = np.loadtxt(path, delimiter=",", dtype=np.float64, ndmin=2)
    if rows.shape != (n, n + 1):
        raise ValueError("%s: expected %d x %d values, got %s" % (path, n, n + 1, rows.shape))
    i = np.arange(n)
    a = np.zeros(n)
    c = np.zeros(n)
    a[1:] = rows[i[1:], i[1:] - 1]
    c[:-1] = rows[i[:-1], i[:-1] + 1]
    d = rows[i, i]
    return tuple(v.astype(np.float32) for v in (a, d, c, rows[:, n]))


def _glob_key(name):
    return name  # glob order of the reference: plain lexicographic


def load_inputs(src, n=N):
    """[(file name, (a, diag, c, rhs))], xt"""
    if os.path.isdir(src):
        xt = np.loadtxt(os.path.join(src, "xt.txt"), dtype=np.float64)[:n].astype(np.float32)
        files = sorted(glob.glob(os.path.join(src, "eq*.txt")), key=lambda p: _glob_key(os.path.basename(p)))
        return [(os.path.basename(p), read_dense_csv(p, n)) for p in files], xt
    z = np.load(src)
    names = sorted((k + ".txt" for k in z.files if k.startswith("eq")), key=_glob_key)
    return [(nm, tuple(np.ascontiguousarray(v, dtype=np.float32) for v in z[nm[:-4]])) for nm in names], \
        np.asarray(z["xt"], np.float32)


def rel_l2(x, xt):
    with np.errstate(over="ignore", invalid="ignore"):     # inf / nan propagate as in the reference
        diff = (np.asarray(x, np.float32) - np.asarray(xt, np.float32))
        sq = (diff * diff).astype(np.float64)              # fp32 square, fp64 sum
        t = np.asarray(xt, np.float32)
        tsq = (t * t).astype(np.float64)
    s1 = 0.0
    s2 = 0.0
    for v in sq:                                           # sequential sums, the reference's order
        s1 += v
    for v in tsq:
        s2 += v
    return float(np.sqrt(s1) / np.sqrt(s2))


def _vendor():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtpr_rocsparse.so")
    if not os.path.exists(path):
        return None
    from ._lib import _load_torch_first
    _load_torch_first()
    L = ctypes.CDLL(path)
    f = L.tprh_vendor_gtsv_batch
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return f


def _solve_vendor(f, bands):
    import torch
    a, d, c, r = (torch.from_numpy(v.copy()).cuda() for v in bands)
    x = torch.empty_like(r)
    st = f(a.data_ptr(), d.data_ptr(), c.data_ptr(), r.data_ptr(), x.data_ptr(), a.numel(), 1, None)
    if st != 0:
        raise RuntimeError("rocSPARSE comparator failed (%d)" % st)
    torch.cuda.synchronize()
    return x.cpu().numpy()


def report(inputs, xt, arith="exact", vendor=True):
    """Rows (file, solver, N, S, L2) in the reference's order."""
    import tpr_amd
    rows = []
    vf = _vendor() if vendor else None
    n = xt.size

    def run(solver, bands):
        solver.set_arith(arith)
        a, d, c, r = (v.copy() for v in bands)    # a fresh copy per solver
        solver.set_tridiagonal_system(a, d, c, r)
        solver.solve()
        x = np.zeros(n, np.float32)
        solver.get_ans(x)
        return x

    for fname, bands in inputs:
        s = 16
        while s <= n:
            rows.append((fname, "PTPR", n, s, rel_l2(run(tpr_amd.PTPR(n, s), bands), xt)))
            rows.append((fname, "TPR", n, s, rel_l2(run(tpr_amd.TPR(n, s), bands), xt)))
            s *= 2
        rows.append((fname, "PCR", n, 0, rel_l2(run(tpr_amd.PCR(n), bands), xt)))
        if vf is not None:
            rows.append((fname, "REFERENCE_ROCSPARSE", n, 0, rel_l2(_solve_vendor(vf, bands), xt)))
    return rows


def format_rows(rows):
    return "".join("%s,%s,%d,%d,%e\n" % r for r in rows)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("inputs")
    ap.add_argument("--fast", action="store_true")
    ap.add_argument("--no-vendor", action="store_true")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    inputs, xt = load_inputs(args.inputs)
    text = format_rows(report(inputs, xt, "fast" if args.fast else "exact", not args.no_vendor))
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    sys.stdout.write(text)


if __name__ == "__main__":
    main()
