"""The nstab harness's host side: the dense-CSV reader against the committed
band fixture, the L2 metric against the oracle's restatement of
src/numerical_stability.cpp:100-115, and the comparator library's exports."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from tpr_amd import nstab

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_dense(path, bands):
    a, d, c, r = (np.asarray(v, np.float64) for v in bands)
    n = a.size
    m = np.zeros((n, n + 1))
    i = np.arange(n)
    m[i, i] = d
    m[i[1:], i[1:] - 1] = a[1:]
    m[i[:-1], i[:-1] + 1] = c[:-1]
    m[:, n] = r
    np.savetxt(path, m, delimiter=",", fmt="%.9g")


def test_reader_round_trips_fixture(golden, tmp_path):
    z = golden["nstab"]
    for k in ("eq1", "eq7", "eq20"):
        p = tmp_path / (k + ".txt")
        write_dense(p, z[k])
        got = nstab.read_dense_csv(str(p))
        for u, v in zip(got, z[k]):
            assert np.array_equal(u.view(np.uint32), np.asarray(v, np.float32).view(np.uint32))


def test_glob_order_matches_reference(tmp_path, golden):
    z = golden["nstab"]
    np.savetxt(tmp_path / "xt.txt", z["xt"].astype(np.float64), fmt="%.9g")
    for k in ("eq1", "eq10", "eq2", "eq3"):
        write_dense(tmp_path / (k + ".txt"), z[k])
    inputs, xt = nstab.load_inputs(str(tmp_path))
    assert [f for f, _ in inputs] == ["eq1.txt", "eq10.txt", "eq2.txt", "eq3.txt"]
    assert np.array_equal(xt, z["xt"])


def test_rel_l2_matches_oracle():
    rng = np.random.default_rng(1)
    for n in (1, 7, 512):
        x = rng.standard_normal(n).astype(np.float32)
        xt = rng.standard_normal(n).astype(np.float32)
        assert nstab.rel_l2(x, xt) == O.norm(x, xt)
    xt = np.arange(1, 513, dtype=np.float32)
    assert nstab.rel_l2(xt, xt) == 0.0


def test_comparator_library_exports():
    path = os.path.join(ROOT, "tpr_amd", "libtpr_rocsparse.so")
    if not os.path.exists(path):
        pytest.skip("rocSPARSE comparator not built (no rocSPARSE in the image)")
    txt = open(os.path.join(ROOT, "include", "tpr_rocsparse.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(tprh_[a-z0-9_]+)\s*\(", txt))
    L = ctypes.CDLL(path)
    assert names == {"tprh_vendor_gtsv_batch", "tprh_vendor_last_error", "tprh_vendor_version"}
    for nm in names:
        assert hasattr(L, nm)
    f = L.tprh_vendor_gtsv_batch
    f.restype = ctypes.c_int
    assert f(None, None, None, None, None, 1, 1, None) != 0   # n < 2 refused before any device call
    L.tprh_vendor_last_error.restype = ctypes.c_char_p
    assert b"n must be" in L.tprh_vendor_last_error()
