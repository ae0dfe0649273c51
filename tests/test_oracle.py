"""Pins the CPU oracle (oracle/tpr_oracle.c) before anything is checked
against it.

* bit-for-bit against golden vectors written by the reference's own solver
  classes compiled from its sources (tests/golden/ref_vectors.npz), and
  against that build live where it is present (oracle/_ref);
* against the reference's known answer test/ans1024.txt with the reference's
  tolerances (test/examples.cpp:36-102), at every s (the reference's own
  TPRTest/PTPRTest only assert the last s of their loop);
* against closed forms and independent solvers: diagonal systems (x = d/b
  exactly), FP64 Thomas (itself checked against scipy.linalg.solve_banded),
  dense numpy.linalg.solve for small n.
"""
import numpy as np
import pytest

import oracle as O

PTPR_CASES = [(16, 2), (16, 4), (64, 8), (64, 64), (256, 16), (512, 32), (1024, 256), (1024, 4),
              (2048, 512), (2048, 64), (4096, 1024), (128, 128)]


def bits_equal(x, y):
    """Identical float32 bits; two NaNs count as equal whatever their payload
    (x86 and other producers differ in NaN sign and payload only)."""
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    if x.shape != y.shape:
        return False
    same = x.view(np.uint32) == y.view(np.uint32)
    return bool(np.all(same | (np.isnan(x) & np.isnan(y))))


# ---------------------------------------------------------------------------
# golden vectors of the reference build

@pytest.mark.parametrize("n,s", PTPR_CASES)
def test_ptpr_tpr_match_reference_vectors(golden, n, s):
    ref = golden["ref"]
    for tag in ("seed1", "seed2", "branch"):
        key = "n%d_s%d_%s" % (n, s, tag)
        na, nc, nd = ref[key + "_norm"]
        assert bits_equal(O.ptpr(na, nc, nd, s), ref[key + "_ptpr"]), key
        assert bits_equal(O.tpr(na, nc, nd, s), ref[key + "_tpr"]), key


@pytest.mark.parametrize("n,s", PTPR_CASES)
def test_normalisation_matches_reference(golden, n, s):
    ref = golden["ref"]
    key = "n%d_s%d_seed1" % (n, s)
    a, b, c, d = ref[key + "_raw"]
    na, nc, nd = O.normalize(a, b, c, d)
    exp = ref[key + "_norm"]
    assert bits_equal(na, exp[0]) and bits_equal(nc, exp[1]) and bits_equal(nd, exp[2])


def test_generator_matches_fixture_inputs(golden):
    ref = golden["ref"]
    for (n, s) in PTPR_CASES:
        a, b, c, d, _ = O.gen_system(1, 3, n)
        raw = ref["n%d_s%d_seed1_raw" % (n, s)]
        for got, exp in zip((a, b, c, d), raw):
            assert bits_equal(got, exp), (n, s)


@pytest.mark.parametrize("n", [2, 16, 256, 1024])
def test_pcr_cr_match_reference_vectors(golden, n):
    ref = golden["ref"]
    na, nc, nd = ref["n%d_pcrcr_norm" % n]
    assert bits_equal(O.pcr(na, nc, nd), ref["n%d_pcr" % n])
    assert bits_equal(O.cr(na, nc, nd), ref["n%d_cr" % n])


def test_fixed_input_matches_reference_vectors(golden):
    ref = golden["ref"]
    a, c, d = ref["fixed1024_norm"]
    for s in (4, 16, 64, 256, 1024):
        assert bits_equal(O.ptpr(a, c, d, s), ref["fixed1024_s%d_ptpr" % s]), s
        assert bits_equal(O.tpr(a, c, d, s), ref["fixed1024_s%d_tpr" % s]), s
    assert bits_equal(O.pcr(a, c, d), ref["fixed1024_pcr"])
    assert bits_equal(O.cr(a, c, d), ref["fixed1024_cr"])


@pytest.mark.parametrize("k", range(1, 21))
def test_nstab_matrices_match_reference_vectors(golden, k):
    """All 20 stability classes, NaN/Inf patterns included."""
    a, b, c, d = golden["nstab"]["eq%d" % k]
    for s in (16, 64, 512):
        na, nc, nd = O.normalize(a, b, c, d)
        assert bits_equal(O.ptpr(na, nc, nd, s), golden["ref"]["eq%d_s%d_ptpr" % (k, s)]), (k, s)
        assert bits_equal(O.tpr(na, nc, nd, s), golden["ref"]["eq%d_s%d_tpr" % (k, s)]), (k, s)


# ---------------------------------------------------------------------------
# live reference build, where present

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built in this tree")


@needs_ref
@pytest.mark.parametrize("n,s", [(32, 2), (512, 512), (2048, 8), (8192, 256), (16384, 128), (65536, 256)])
def test_oracle_matches_live_reference(n, s):
    for seed in (11, 12):
        for kind in (O.GEN_DOMINANT, O.GEN_WEAK):
            a, b, c, d, _ = O.gen_system(seed, 1, n, kind)
            na, nc, nd = O.normalize(a, b, c, d)
            ra, rc, rd = O.ref_normalize(a, b, c, d, s)
            assert bits_equal(na, ra) and bits_equal(nc, rc) and bits_equal(nd, rd)
            assert bits_equal(O.ptpr(na, nc, nd, s), O.ref_ptpr(na, nc, nd, s))
            assert bits_equal(O.tpr(na, nc, nd, s), O.ref_tpr(na, nc, nd, s))


@needs_ref
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_batch_shapes_match_live_reference(method):
    n, s, nsys = 1024, 256, 24
    a, b, c, d, _ = O.gen_batch(5, 0, nsys, n)
    for mode in (0, 1):
        x = O.solve_batch(method, a, b, c, d, n, s, mode)
        xr = O.ref_solve_batch(method, a, b, c, d, n, s, mode)
        assert bits_equal(x, xr)


# ---------------------------------------------------------------------------
# known answer of the reference's tests

def _ans_check(ans, x):
    # test/examples.cpp:36-56: |x - ans| <= 1e-2 per element, max squared diff <= 1e-3
    x = np.asarray(x, np.float64)
    assert np.all(np.abs(x - ans) <= 1e-2)
    assert np.max((x - ans) ** 2) <= 1e-3


def test_known_answer_every_s(golden):
    ans = golden["ans1024"]
    a, b, c, d, _ = O.gen_system(0, 0, 1024, O.GEN_FIXED)
    _ans_check(ans, O.cr(a, c, d))
    _ans_check(ans, O.pcr(a, c, d))
    s = 4
    while s <= 1024:
        _ans_check(ans, O.tpr(a, c, d, s))
        _ans_check(ans, O.ptpr(a, c, d, s))
        s *= 2


def test_fixed_input_matches_reference_generator():
    """ExampleFixedInput (src/system.cpp:118-130)."""
    a, b, c, d, _ = O.gen_system(0, 0, 16, O.GEN_FIXED)
    assert a[0] == 0 and c[-1] == 0
    assert np.all(a[1:] == np.float32(-1.0 / 6.0)) and np.all(c[:-1] == np.float32(-1.0 / 6.0))
    assert np.all(b == 1) and np.array_equal(d, np.arange(1, 17, dtype=np.float32))


# ---------------------------------------------------------------------------
# closed forms and independent solvers

@pytest.mark.parametrize("n,s", [(64, 8), (1024, 256), (2048, 2)])
def test_diagonal_system_is_exact(n, s):
    rng = np.random.default_rng(3)
    b = rng.uniform(1, 2, n).astype(np.float32)
    d = rng.uniform(-1, 1, n).astype(np.float32)
    z = np.zeros(n, np.float32)
    na, nc, nd = O.normalize(z, b, z, d)
    exp = d / b
    assert bits_equal(O.ptpr(na, nc, nd, s), exp)
    assert bits_equal(O.tpr(na, nc, nd, s), exp)
    assert bits_equal(O.pcr(na, nc, nd), exp)
    assert bits_equal(O.cr(na, nc, nd), exp)


def test_thomas_matches_scipy_banded():
    from scipy.linalg import solve_banded
    for n in (5, 64, 1000):
        a, b, c, d, _ = O.gen_system(7, n, n, O.GEN_WEAK)
        ab = np.zeros((3, n))
        ab[0, 1:] = c[:-1]
        ab[1] = b
        ab[2, :-1] = a[1:]
        xs = solve_banded((1, 1), ab, d.astype(np.float64))
        xt = O.thomas_f64(a, b, c, d)
        assert np.max(np.abs(xs - xt)) <= 1e-10 * max(1.0, np.max(np.abs(xs)))


@pytest.mark.parametrize("n,s", [(4, 2), (16, 4), (64, 16), (64, 64)])
def test_dense_solve_small(n, s):
    for seed in range(5):
        a, b, c, d, xt = O.gen_system(seed, 2, n)
        A = np.diag(b.astype(np.float64)) + np.diag(a[1:].astype(np.float64), -1) + np.diag(c[:-1].astype(np.float64), 1)
        xd = np.linalg.solve(A, d.astype(np.float64))
        na, nc, nd = O.normalize(a, b, c, d)
        for f in (lambda: O.ptpr(na, nc, nd, s), lambda: O.tpr(na, nc, nd, s),
                  lambda: O.pcr(na, nc, nd), lambda: O.cr(na, nc, nd)):
            assert O.rel_l2(f(), xd) < 1e-6


@pytest.mark.parametrize("n,s", [(1024, 256), (2048, 512), (16384, 128), (65536, 256)])
def test_ptpr_tpr_close_to_fp64_thomas(n, s):
    a, b, c, d, xt = O.gen_system(21, 0, n)
    xf = O.thomas_f64(a, b, c, d)
    assert O.rel_l2(xf, xt) < 1e-6
    na, nc, nd = O.normalize(a, b, c, d)
    assert O.rel_l2(O.ptpr(na, nc, nd, s), xf) < 1e-6
    assert O.rel_l2(O.tpr(na, nc, nd, s), xf) < 1e-6


def test_s_equals_n_is_pcr_up_to_rounding():
    """m = 1: PTPR is one slice of PCR (src/ptpr.cpp:238-240 + src/pcr.cpp:16)."""
    a, b, c, d, _ = O.gen_system(9, 9, 512)
    na, nc, nd = O.normalize(a, b, c, d)
    assert O.rel_l2(O.ptpr(na, nc, nd, 512), O.pcr(na, nc, nd)) < 1e-6


def test_residual_f64_of_truth_is_rounding_only():
    a, b, c, d, xt = O.gen_system(3, 4, 4096)
    r = O.residual_f64(a, b, c, d, xt.astype(np.float64))
    # d is A*xt rounded once to float: |r| <= ulp(d) / 2
    assert np.all(np.abs(r) <= np.spacing(np.abs(d)).astype(np.float64))


# ---------------------------------------------------------------------------
# generator and metric

def test_generator_independent_of_batch_composition():
    n = 256
    a, b, c, d, xt = O.gen_batch(42, 10, 5, n)
    for k in range(5):
        a1, b1, c1, d1, x1 = O.gen_system(42, 10 + k, n)
        sl = slice(k * n, (k + 1) * n)
        for got, exp in ((a[sl], a1), (b[sl], b1), (c[sl], c1), (d[sl], d1), (xt[sl], x1)):
            assert bits_equal(got, exp)


@pytest.mark.parametrize("kind", [O.GEN_DOMINANT, O.GEN_WEAK])
def test_generator_properties(kind):
    n = 4096
    a, b, c, d, xt = O.gen_system(1, 2, n, kind)
    assert a[0] == 0 and c[-1] == 0
    assert np.all(np.abs(a) <= 1) and np.all(np.abs(c) <= 1) and np.all(np.abs(xt) <= 1)
    margin = b - (np.abs(a) + np.abs(c))
    if kind == O.GEN_DOMINANT:
        assert np.all(margin >= 0.99) and np.all(margin <= 2.01)
    else:
        assert np.all(margin >= 0.5e-4) and np.all(margin <= 1.01e-2)
    # U(-1,1) values are multiples of 2^-23
    assert np.all((a * 2 ** 23) == np.round(a * 2 ** 23))
    dd = b.astype(np.float64) * xt
    dd[1:] += a[1:].astype(np.float64) * xt[:-1]
    dd[:-1] += c[:-1].astype(np.float64) * xt[1:]
    assert np.max(np.abs(dd.astype(np.float32) - d)) <= 1e-6


def test_norm_formula():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(512).astype(np.float32)
    t = rng.standard_normal(512).astype(np.float32)
    diff = (x - t).astype(np.float32)
    exp = np.sqrt(np.sum((diff * diff).astype(np.float64))) / np.sqrt(np.sum((t * t).astype(np.float64)))
    assert abs(O.norm(x, t) - exp) <= 1e-12 * exp
