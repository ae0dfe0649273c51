"""GPU side of the nstab harness: the report over the reference's nstab
inputs (committed as bands) equals, line for line, the report computed from
the oracle's solutions (exact arithmetic: identical x, identical L2); the
C++ driver prints the same CSV from the dense-CSV text files; the rocSPARSE
comparator solves dominant batches to fp32 accuracy."""
import os
import subprocess

import numpy as np
import pytest
import torch

import oracle as O
from tpr_amd import nstab

from test_nstab_host import write_dense

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def oracle_rows(inputs, xt):
    rows = []
    n = xt.size
    for fname, (a, d, c, r) in inputs:
        na, nc, nr = O.normalize(a, d, c, r)
        s = 16
        while s <= n:
            rows.append((fname, "PTPR", n, s, O.norm(O.ptpr(na, nc, nr, s), xt)))
            rows.append((fname, "TPR", n, s, O.norm(O.tpr(na, nc, nr, s), xt)))
            s *= 2
        rows.append((fname, "PCR", n, 0, O.norm(O.pcr(na, nc, nr), xt)))
    return rows


def test_nstab_report_matches_oracle(golden):
    path = os.path.join(ROOT, "tests", "golden", "nstab.npz")
    inputs, xt = nstab.load_inputs(path)
    got = nstab.report(inputs, xt, "exact", vendor=False)
    want = oracle_rows(inputs, xt)
    assert len(got) == len(want) == 20 * 13
    for g, w in zip(got, want):
        assert g[:4] == w[:4]
        assert g[4] == w[4] or (np.isnan(g[4]) and np.isnan(w[4])), (g, w)


def test_cli_prints_the_same_csv(tmp_path, golden):
    z = golden["nstab"]
    np.savetxt(tmp_path / "xt.txt", z["xt"].astype(np.float64), fmt="%.9g")
    for k in ("eq1", "eq12", "eq2", "eq5"):
        write_dense(tmp_path / (k + ".txt"), z[k])
    exe = os.path.join(ROOT, "tpr_amd", "tpr_nstab")
    print("tpr_nstab", tmp_path, flush=True)
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    inputs, xt = nstab.load_inputs(str(tmp_path))
    rows = nstab.report(inputs, xt, "exact")
    cli = out.stdout.splitlines()
    py = nstab.format_rows(rows).splitlines()
    assert len(cli) == len(py)
    for u, v in zip(cli, py):
        if "REFERENCE_ROCSPARSE" in v:
            # the vendor's answers are its own: both drivers must run it, on
            # the same inputs; its bits on ill-conditioned inputs are not ours
            assert u.split(",")[:4] == v.split(",")[:4]
            if u != v:
                print("vendor lines differ:", u, v, flush=True)
        else:
            assert u == v
    if os.path.exists(os.path.join(ROOT, "tpr_amd", "libtpr_rocsparse.so")):
        assert "REFERENCE_ROCSPARSE" in out.stdout


def test_rocsparse_comparator_accuracy():
    f = nstab._vendor()
    if f is None:
        pytest.skip("rocSPARSE comparator not built")
    nsys, n = 64, 1024
    a, b, c, d, xt = O.gen_batch(12, 0, nsys, n)
    ta, tb, tc, td = (torch.from_numpy(v).cuda() for v in (a, b, c, d))
    x = torch.empty_like(td)
    assert f(ta.data_ptr(), tb.data_ptr(), tc.data_ptr(), td.data_ptr(), x.data_ptr(), n, nsys, None) == 0
    torch.cuda.synchronize()
    xf = O.thomas_f64_batch(a, b, c, d, n).reshape(nsys, n)
    got = x.cpu().numpy().reshape(nsys, n).astype(np.float64)
    err = np.linalg.norm(got - xf, axis=1) / np.linalg.norm(xf, axis=1)
    assert err.max() <= 1e-5
    assert np.array_equal(td.cpu().numpy(), d)   # d untouched
