"""Seeded random sweep of shapes through the general entry point
(tprh_solve_ex): n = 2^k, s = 2^j with every supported combination of the
one-workgroup and the multi-workgroup path, both methods, with and without a
diagonal.  Exact arithmetic must equal the oracle bit for bit on every
system (strictly or weakly dominant inputs); fast arithmetic must stay
within rel-L2 1e-5 of it on strictly dominant inputs."""
import os
import sys

import numpy as np
import pytest
import torch

import oracle as O
from tpr_amd import batched as B

pytestmark = pytest.mark.gpu


def bits_equal(x, y):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    same = x.view(np.uint32) == y.view(np.uint32)
    return bool(np.all(same | (np.isnan(x) & np.isnan(y))))


def dev(v):
    return torch.from_numpy(np.ascontiguousarray(v)).cuda()


def sweep_cases(count=48, seed=20251202):
    rng = np.random.default_rng(seed)
    cases = []
    while len(cases) < count:
        k = int(rng.integers(1, 18))          # n = 2 .. 131072
        n = 1 << k
        j = int(rng.integers(1, min(k, 11) + 1))  # s = 2 .. min(n, 2048)
        s = 1 << j
        m = n // s
        if m > 2048:
            continue
        nsys = int(rng.integers(1, 40))
        nsys = max(1, min(nsys, (1 << 21) // n))  # at most 2 Mi equations per case
        arith = ["exact", "fast"][int(rng.integers(0, 2))]
        kind = int(rng.integers(0, 2))
        if arith == "fast":
            # the fast tolerance is for well-conditioned systems: a weakly
            # dominant 2 x 2 block has a determinant near eps * (|a| + |c|),
            # where two correct fp32 orders of operations differ by ~1e-5
            kind = 0
        cases.append((n, s, nsys, ["ptpr", "tpr"][int(rng.integers(0, 2))], bool(rng.integers(0, 2)), arith, kind))
    return cases


# TPR_SWEEP_COUNT / TPR_SWEEP_SEED widen the sweep for a one-off run
@pytest.mark.parametrize("n,s,nsys,method,with_b,arith,kind",
                         sweep_cases(int(os.environ.get("TPR_SWEEP_COUNT", "48")),
                                     int(os.environ.get("TPR_SWEEP_SEED", "20251202"))))
def test_sweep(n, s, nsys, method, with_b, arith, kind):
    print("sweep", n, s, nsys, method, with_b, arith, kind, flush=True)
    sys.stdout.flush()
    a, b, c, d, xt = O.gen_batch(41 + n + s, 0, nsys, n, kind)
    if with_b:
        ref = O.solve_batch(method, a, b, c, d, n, s)
        bd = dev(b)
    else:
        na, nc, nd = [], [], []
        for q in range(nsys):
            sl = slice(q * n, (q + 1) * n)
            u, v, w = O.normalize(a[sl], b[sl], c[sl], d[sl])
            na.append(u)
            nc.append(v)
            nd.append(w)
        a, c, d = np.concatenate(na), np.concatenate(nc), np.concatenate(nd)
        solve = O.ptpr if method == "ptpr" else O.tpr
        ref = np.concatenate([solve(a[q * n:(q + 1) * n], c[q * n:(q + 1) * n], d[q * n:(q + 1) * n].copy(), s)
                              for q in range(nsys)])
        bd = None
    x = torch.empty(nsys * n, device="cuda")
    B.solve_ex(dev(a), bd, dev(c), dev(d), x, n, s, method=method, arith=arith)
    torch.cuda.synchronize()
    got = x.cpu().numpy()
    if arith == "exact":
        assert bits_equal(got, ref)
    else:
        g = got.reshape(nsys, n).astype(np.float64)
        r = ref.reshape(nsys, n).astype(np.float64)
        rel = np.linalg.norm(g - r, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-30)
        assert rel.max() <= 1e-5, rel.max()
