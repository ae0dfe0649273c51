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
  dense numpy.linalg.solve for small n;
* the analytic flop counts against what the reference classes' solve()
  returned (tests/golden/ref_flops.json, and live);
* the seeded generator against SplitMix64's published output (its mixing
  function and Weyl sequence) and against properties of its right-hand-side
  stream: range, grid, index layout, independence.
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


# ---------------------------------------------------------------------------
# analytic flop counts: what the reference classes' solve() returns

def test_flops_match_reference_fixture(golden):
    """orc_ptpr_flops / orc_tpr_flops against the return value of the
    reference's own solve() (src/ptpr.cpp:103-106, src/tpr.cpp:95-98), over
    slice sizes 2 .. 1024 and 1 .. 256 slices, so that a swapped log2(s) /
    log2(m) or a dropped stage term changes some case."""
    fl = golden["flops"]
    assert len(fl["ptpr"]) >= 20 and set(fl["ptpr"]) == set(fl["tpr"])
    for key, want in fl["ptpr"].items():
        n, s = map(int, key.split(","))
        assert O.ptpr_flops(n, s) == want, ("ptpr", n, s)
    for key, want in fl["tpr"].items():
        n, s = map(int, key.split(","))
        assert O.tpr_flops(n, s) == want, ("tpr", n, s)


def test_flops_fixture_holds_the_documented_count(golden):
    # the count the reference's documented run (tpr_pm 2048 512 ... PTPR)
    # reports per solve, and that of N=1024, S=256
    assert golden["flops"]["ptpr"]["2048,512"] == 266419
    assert golden["flops"]["ptpr"]["1024,256"] == 118963


@needs_ref
@pytest.mark.parametrize("n,s", [(64, 2), (1024, 256), (2048, 512), (8192, 32), (65536, 256)])
def test_flops_match_live_reference(n, s):
    a, b, c, d, _ = O.gen_system(2, 1, n)
    na, nc, nd = O.normalize(a, b, c, d)
    x, f = O.ref_ptpr(na, nc, nd, s, with_flops=True)
    assert f == O.ptpr_flops(n, s)
    assert bits_equal(x, O.ptpr(na, nc, nd, s))
    x, f = O.ref_tpr(na, nc, nd, s, with_flops=True)
    assert f == O.tpr_flops(n, s)
    assert bits_equal(x, O.tpr(na, nc, nd, s))


# ---------------------------------------------------------------------------
# the generator: SplitMix64 and the right-hand-side stream

_M64 = (1 << 64) - 1
_GAMMA = 0x9E3779B97F4A7C15


def _splitmix64(state, count):
    """SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c): the
    state advances by the golden gamma and each output is the state
    through the variant-13 finaliser."""
    out = []
    for _ in range(count):
        state = (state + _GAMMA) & _M64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_published_vectors():
    # the first outputs of SplitMix64 seeded with 0, as published with the
    # reference implementation
    assert _splitmix64(0, 5) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
                                 0xF88BB8A8724C81EC, 0x1B39896A51A8749B]


@pytest.mark.parametrize("seed,sys", [(0, 0), (1, 7), (12345, 3)])
def test_gen_rhs_is_splitmix64(seed, sys):
    """Right-hand side j of system `sys`: element i is the top 24 bits k of
    output 8(j n + i) + 6 of a SplitMix64 stream whose state is output
    sys + 1 of a SplitMix64 stream seeded with `seed`, mapped to
    (k - 2^23) 2^-23."""
    n, j = 8, 3
    key = _splitmix64(seed, sys + 1)[-1]
    outs = _splitmix64(key, 8 * (j * n + n - 1) + 6)
    exp = np.array([((outs[8 * (j * n + i) + 5] >> 40) - (1 << 23)) / float(1 << 23) for i in range(n)],
                   dtype=np.float32)
    assert bits_equal(O.gen_rhs(seed, sys, j, n), exp)


def test_gen_rhs_range_and_grid():
    d = O.gen_rhs(9, 4, 0, 1 << 16)
    assert d.dtype == np.float32
    assert np.all(d >= -1.0) and np.all(d < 1.0)
    k = d.astype(np.float64) * 2 ** 23
    assert np.all(k == np.round(k))  # multiples of 2^-23: exact in fp32
    # uniform: mean 0 and variance 1/3 within 5 standard errors
    assert abs(d.mean()) < 5 * np.sqrt(1 / 3 / d.size)
    assert abs(d.var() - 1 / 3) < 5 * np.sqrt(4 / 45 / d.size)
    # about as many distinct values as a 2^24-point grid gives
    assert np.unique(d).size > 0.99 * d.size


def test_gen_rhs_index_layout():
    """Right-hand side j of length n covers stream indices [j n, (j+1) n):
    splitting it in halves gives right-hand sides 2j and 2j+1 of length n/2."""
    for j in (0, 1, 5):
        full = O.gen_rhs(3, 11, j, 64)
        halves = np.concatenate([O.gen_rhs(3, 11, 2 * j, 32), O.gen_rhs(3, 11, 2 * j + 1, 32)])
        assert bits_equal(full, halves), j


def test_gen_rhs_independent_of_order_and_other_systems():
    n = 1024
    fwd = [O.gen_rhs(5, 2, j, n) for j in range(6)]
    rev = [O.gen_rhs(5, 2, j, n) for j in reversed(range(6))][::-1]
    for x, y in zip(fwd, rev):
        assert bits_equal(x, y)
    # other systems, seeds and right-hand sides draw other values
    others = [O.gen_rhs(5, 3, 0, n), O.gen_rhs(6, 2, 0, n), fwd[1]]
    for o in others:
        assert np.mean(o == fwd[0]) < 0.01
    # and none of them repeats a stream of the system generator
    a, b, c, d, xt = O.gen_system(5, 2, n)
    for v in (a, c, xt, d):
        assert np.mean(v == fwd[0]) < 0.01
    # successive right-hand sides are uncorrelated
    r = np.corrcoef(fwd[0], fwd[1])[0, 1]
    assert abs(r) < 5 / np.sqrt(n)


def test_thomas_batch_equals_per_system():
    n, nsys = 500, 7
    a, b, c, d, _ = O.gen_batch(8, 2, nsys, n, O.GEN_WEAK)
    xb = O.thomas_f64_batch(a, b, c, d, n)
    for k in range(nsys):
        sl = slice(k * n, (k + 1) * n)
        assert np.array_equal(xb[sl], O.thomas_f64(a[sl], b[sl], c[sl], d[sl])), k


@pytest.mark.parametrize("kind", [O.GEN_DOMINANT, O.GEN_WEAK])
def test_gen_system_streams_are_splitmix64(kind):
    """a, c, the diagonal's U[0,1) and x_t are streams 0, 1, 2, 3 of the
    system's SplitMix64 sequence (element i: output 8 i + stream + 1); the
    diagonal is b = (|a| + |c|) + extra in fp32, d = A x_t summed in FP64."""
    seed, sys, n = 77, 5, 12
    key = _splitmix64(seed, sys + 1)[-1]
    outs = _splitmix64(key, 8 * n)
    top = lambda st, i: outs[8 * i + st] >> 40  # noqa: E731
    pm1 = lambda k: np.float32((k - (1 << 23)) / float(1 << 23))  # noqa: E731
    a = np.array([pm1(top(0, i)) for i in range(n)], np.float32)
    c = np.array([pm1(top(1, i)) for i in range(n)], np.float32)
    u = np.array([np.float32(top(2, i) / float(1 << 24)) for i in range(n)], np.float32)
    xt = np.array([pm1(top(3, i)) for i in range(n)], np.float32)
    a[0] = 0
    c[-1] = 0
    extra = (np.float32(1) + u) if kind == O.GEN_DOMINANT else (np.float32(1e-4) + u * np.float32(0.0099))
    b = (np.abs(a) + np.abs(c)) + extra.astype(np.float32)
    d = b.astype(np.float64) * xt
    d[1:] = a[1:].astype(np.float64) * xt[:-1] + d[1:]
    d[:-1] = d[:-1] + c[:-1].astype(np.float64) * xt[1:]
    ga, gb, gc, gd, gx = O.gen_system(seed, sys, n, kind)
    for got, exp in ((ga, a), (gc, c), (gb, b), (gx, xt), (gd, d.astype(np.float32))):
        assert bits_equal(got, exp)
