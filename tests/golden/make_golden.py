"""Writes the golden fixtures under tests/golden/ (run in a tree where the
reference is present and oracle/_ref was built by `make -C oracle ref`).

  ans1024.npy      the reference's known answer test/ans1024.txt (data file of
                   its own tests): solution of tridiag(-1/6, 1, -1/6) x = 1..1024
  nstab.npz        the tridiagonal bands of src/inputs/eq1..eq20.txt (dense CSV
                   rows, rhs in the last column, read in double and cast to
                   float as src/system.cpp:54-116 does), xt.txt (n = 512) and
                   cond.dat
  ref_vectors.npz  inputs and outputs of the reference's own solver classes,
                   compiled from its sources (oracle/_ref/libtprref.so), on
                   seeded generator systems: PTPR / TPR for a set of (n, s),
                   PCR / CR, the 4-argument normalisation, and systems that
                   force the s1 = -1 and key = 0 branches (src/ptpr.cpp:196-200,
                   :276-279)
  ref_flops.json   what the reference classes' solve() returns (the analytic
                   flop count, src/ptpr.cpp:103-106, src/tpr.cpp:95-98) for
                   every (n, s) of FLOP_CASES, taken from the return value of
                   a real solve

    python tests/golden/make_golden.py [--ref /root/reference] [--only flops]
"""
import argparse
import glob
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

PTPR_CASES = [(16, 2), (16, 4), (64, 8), (64, 64), (256, 16), (512, 32), (1024, 256), (1024, 4),
              (2048, 512), (2048, 64), (4096, 1024), (128, 128)]
PCR_CR_N = [2, 16, 256, 1024]
# every solve shape of the parity cases, the benchmark configurations
# (2048/512, 1024/256, 65536/256) and the uneven size classes at s = 256
FLOP_CASES = sorted(set(PTPR_CASES + [(2048, 512), (1024, 256), (65536, 256), (256, 256), (512, 256),
                                      (2048, 256), (4096, 256), (8192, 256), (16384, 256), (32768, 256),
                                      (4096, 2), (8192, 8), (1024, 1024), (32768, 1024)]))


def write_flops():
    out = {"ptpr": {}, "tpr": {}}
    for (n, s) in FLOP_CASES:
        a, b, c, d, _ = O.gen_system(1, 3, n)
        na, nc, nd = O.ref_normalize(a, b, c, d, s)
        out["ptpr"]["%d,%d" % (n, s)] = int(O.ref_ptpr(na, nc, nd, s, with_flops=True)[1])
        out["tpr"]["%d,%d" % (n, s)] = int(O.ref_tpr(na, nc, nd, s, with_flops=True)[1])
    with open(os.path.join(HERE, "ref_flops.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


def read_ans1024(ref):
    txt = open(os.path.join(ref, "test", "ans1024.txt")).read()
    vals = [float(v) for v in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", txt)]
    assert len(vals) == 1024, len(vals)
    return np.array(vals, dtype=np.float64)


def read_dense_band(path, n):
    """src/system.cpp:54-116: row i holds n matrix entries and the rhs."""
    a = np.zeros(n)
    b = np.zeros(n)
    c = np.zeros(n)
    d = np.zeros(n)
    with open(path) as f:
        rows = [ln for ln in f.read().splitlines() if ln.strip()]
    assert len(rows) == n, (path, len(rows))
    for i, ln in enumerate(rows):
        v = [float(t) for t in ln.split(",")]
        assert len(v) == n + 1
        if i > 0:
            a[i] = v[i - 1]
        b[i] = v[i]
        if i < n - 1:
            c[i] = v[i + 1]
        d[i] = v[n]
    return a.astype(np.float32), b.astype(np.float32), c.astype(np.float32), d.astype(np.float32)


def branch_system(n, s, seed):
    """Dominant system with c = 0 at every slice end: decoupled slices, so the
    s1 = -1 branch (|c_ed| < eps) and the key = 0 branch (c_ed == 0) run in
    every slice."""
    a, b, c, d, _ = O.gen_system(seed, 0, n)
    c[s - 1::s] = 0.0
    a[::s] = 0.0
    return a, b, c, d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", choices=["flops"], help="write only this fixture")
    args = ap.parse_args()
    if not O.ref_available():
        sys.exit("oracle/_ref/libtprref.so is not built: run `make -C oracle ref`")
    write_flops()
    if args.only == "flops":
        print("wrote", os.path.join(HERE, "ref_flops.json"))
        return

    np.save(os.path.join(HERE, "ans1024.npy"), read_ans1024(args.ref))

    inp = os.path.join(args.ref, "src", "inputs")
    files = sorted(glob.glob(os.path.join(inp, "eq*.txt")), key=lambda p: int(re.findall(r"\d+", os.path.basename(p))[0]))
    n = 512
    bands = {}
    for p in files:
        k = int(re.findall(r"\d+", os.path.basename(p))[0])
        a, b, c, d = read_dense_band(p, n)
        bands["eq%d" % k] = np.stack([a, b, c, d])
    xt = np.array([float(v) for v in open(os.path.join(inp, "xt.txt")).read().split()][:n]).astype(np.float32)
    cond = np.array([float(v) for v in open(os.path.join(inp, "cond.dat")).read().split()])
    np.savez_compressed(os.path.join(HERE, "nstab.npz"), xt=xt, cond=cond, **bands)

    vec = {}
    for (n, s) in PTPR_CASES:
        for seed in (1, 2):
            a, b, c, d, _ = O.gen_system(seed, 3, n)
            na, nc, nd = O.ref_normalize(a, b, c, d, s)
            key = "n%d_s%d_seed%d" % (n, s, seed)
            if seed == 1:
                vec[key + "_raw"] = np.stack([a, b, c, d])
            vec[key + "_norm"] = np.stack([na, nc, nd])
            vec[key + "_ptpr"] = O.ref_ptpr(na, nc, nd, s)
            vec[key + "_tpr"] = O.ref_tpr(na, nc, nd, s)
        a, b, c, d = branch_system(n, s, 9)
        na, nc, nd = O.ref_normalize(a, b, c, d, s)
        key = "n%d_s%d_branch" % (n, s)
        vec[key + "_norm"] = np.stack([na, nc, nd])
        vec[key + "_ptpr"] = O.ref_ptpr(na, nc, nd, s)
        vec[key + "_tpr"] = O.ref_tpr(na, nc, nd, s)
    for n in PCR_CR_N:
        a, b, c, d, _ = O.gen_system(4, 5, n)
        na, nc, nd = O.ref_normalize(a, b, c, d, 2 if n >= 2 else 1)
        key = "n%d" % n
        vec[key + "_pcrcr_norm"] = np.stack([na, nc, nd])
        vec[key + "_pcr"] = O.ref_pcr(na, nc, nd)
        vec[key + "_cr"] = O.ref_cr(na, nc, nd)
    # the fixed system of the reference's tests (ExampleFixedInput(1024))
    a, b, c, d, _ = O.gen_system(0, 0, 1024, O.GEN_FIXED)
    vec["fixed1024_norm"] = np.stack([a, c, d])
    for s in (4, 16, 64, 256, 1024):
        vec["fixed1024_s%d_ptpr" % s] = O.ref_ptpr(a, c, d, s)
        vec["fixed1024_s%d_tpr" % s] = O.ref_tpr(a, c, d, s)
    vec["fixed1024_pcr"] = O.ref_pcr(a, c, d)
    vec["fixed1024_cr"] = O.ref_cr(a, c, d)
    # the reference classes on the nstab matrices (4-argument setter, as nstab
    # does, src/numerical_stability.cpp:117-126)
    for k in range(1, 21):
        a, b, c, d = bands["eq%d" % k]
        for s in (16, 64, 512):
            na, nc, nd = O.ref_normalize(a, b, c, d, s)
            vec["eq%d_s%d_ptpr" % (k, s)] = O.ref_ptpr(na, nc, nd, s)
            vec["eq%d_s%d_tpr" % (k, s)] = O.ref_tpr(na, nc, nd, s)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **vec)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
