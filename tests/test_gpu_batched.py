"""GPU parity of the batched PTPR / TPR kernels, the generator and the
reference-surface entry points, through the C ABI.

Bar:
  * arith="exact": bit-identical to the oracle (which is itself bit-identical
    to the reference's CPU classes);
  * arith="fast":  relative L2 distance to the oracle <= 1e-5 per system
    (tolerance of the north-star parity requirement, fp32).
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
import tpr_amd
from tpr_amd import batched as B

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def bits_equal(x, y):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    same = x.view(np.uint32) == y.view(np.uint32)
    return bool(np.all(same | (np.isnan(x) & np.isnan(y))))


def to_dev(*arrs):
    return [None if v is None else torch.from_numpy(np.ascontiguousarray(v, np.float32)).cuda() for v in arrs]


def gpu_solve(a, b, c, d, n, s, method, arith, shape=None):
    da, db, dc, dd = to_dev(a, b, c, d)
    dx = torch.empty_like(da)
    B.solve_batched(da, db, dc, dd, dx, n, s, method=method, arith=arith, shape=shape)
    torch.cuda.synchronize()
    return dx.cpu().numpy()


def per_system_rel(x, ref, n):
    x = x.reshape(-1, n).astype(np.float64)
    ref = ref.reshape(-1, n).astype(np.float64)
    return np.linalg.norm(x - ref, axis=1) / np.linalg.norm(ref, axis=1)


# ---------------------------------------------------------------------------
# generator

@pytest.mark.parametrize("kind", ["dominant", "weak", "fixed", "random_rhs"])
@pytest.mark.parametrize("n,nsys", [(16, 7), (1024, 33), (4096, 3)])
def test_generator_bit_identical(kind, n, nsys):
    say("generator", kind, n, nsys)
    t = B.alloc_batch(nsys, n)
    B.generate(kind, 1234, 17, nsys, n, t["a"], t["b"], t["c"], t["d"], t["xt"])
    torch.cuda.synchronize()
    code = {"dominant": O.GEN_DOMINANT, "weak": O.GEN_WEAK, "fixed": O.GEN_FIXED, "random_rhs": O.GEN_RANDOM_RHS}[kind]
    ra, rb, rc, rd, rxt = O.gen_batch(1234, 17, nsys, n, code)
    for k, ref in (("a", ra), ("b", rb), ("c", rc), ("d", rd)):
        assert bits_equal(t[k].cpu().numpy(), ref), k
    if kind in ("dominant", "weak"):
        assert bits_equal(t["xt"].cpu().numpy(), rxt)


# ---------------------------------------------------------------------------
# batched PTPR / TPR, default slice shapes

CASES = [(16, 2), (16, 4), (64, 8), (256, 16), (512, 32), (1024, 64), (2048, 128), (1024, 256),
         (2048, 512), (4096, 1024), (8192, 2048), (256, 256), (1024, 4), (16384, 256), (1024, 16)]


@pytest.mark.parametrize("n,s", CASES)
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_exact_is_bit_identical(n, s, method):
    nsys = 5 if n >= 8192 else 37
    say("exact", method, n, s, nsys)
    a, b, c, d, _ = O.gen_batch(99, 0, nsys, n)
    # 4-argument form: the kernel divides by b
    x = gpu_solve(a, b, c, d, n, s, method, "exact")
    ref = O.solve_batch(method, a, b, c, d, n, s)
    assert bits_equal(x, ref)
    # 3-argument form: unit diagonal
    na, nc, nd = O.normalize(a, b, c, d)
    x = gpu_solve(na, None, nc, nd, n, s, method, "exact")
    assert bits_equal(x, ref)


@pytest.mark.parametrize("n,s", CASES)
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_fast_within_tolerance(n, s, method):
    nsys = 5 if n >= 8192 else 37
    say("fast", method, n, s, nsys)
    a, b, c, d, xt = O.gen_batch(7, 100, nsys, n)
    x = gpu_solve(a, b, c, d, n, s, method, "fast")
    ref = O.solve_batch(method, a, b, c, d, n, s)
    rel = per_system_rel(x, ref, n)
    assert rel.max() <= FAST_TOL, rel.max()
    # and as accurate as the reference against the truth
    err_gpu = per_system_rel(x, xt, n).max()
    err_ref = per_system_rel(ref, xt, n).max()
    assert err_gpu <= max(2 * err_ref, 1e-6)


ALT_SHAPES = [(256, 8, 32), (256, 4, 64), (512, 8, 64), (512, 16, 32), (1024, 16, 64), (2048, 32, 64)]


@pytest.mark.parametrize("s,E,L", ALT_SHAPES)
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_alternative_shapes_bit_identical(s, E, L, method):
    n = max(s * 4, 1024)
    nsys = 9
    say("shape", method, n, s, E, L)
    a, b, c, d, _ = O.gen_batch(3, 0, nsys, n)
    x = gpu_solve(a, b, c, d, n, s, method, "exact", shape=(E, L))
    assert bits_equal(x, O.solve_batch(method, a, b, c, d, n, s))


# Loads and stores staged through LDS as 1 KB runs per wavefront (E = 8 and
# 16): several systems per wavefront, one system per wavefront, systems of
# 128 to 1024 lanes (the last above 64 KB of LDS), with a last workgroup or
# wavefront only partly filled.
STAGED_CASES = [(256, 256, 16, 16, 37), (1024, 256, 16, 16, 37), (512, 128, 8, 16, 37), (4096, 256, 16, 16, 3),
                (1024, 256, 8, 32, 5), (16384, 256, 16, 16, 2), (8192, 512, 8, 64, 3)]


@pytest.mark.parametrize("n,s,E,L,nsys", STAGED_CASES)
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_staged_io_exact_and_in_bounds(method, n, s, E, L, nsys):
    """Exact answers bit for bit, and nothing written past the batch's x
    (a guard region behind it keeps its values)."""
    say("staged", method, n, s, E, L, nsys)
    a, b, c, d, _ = O.gen_batch(23, 5, nsys, n)
    da, db, dc, dd = to_dev(a, b, c, d)
    guard = 4096
    xfull = torch.full((nsys * n + guard,), 7.25, dtype=torch.float32, device="cuda")
    B.solve_batched(da, db, dc, dd, xfull[:nsys * n], n, s, method=method, arith="exact", shape=(E, L))
    torch.cuda.synchronize()
    xf = xfull.cpu().numpy()
    assert bits_equal(xf[:nsys * n], O.solve_batch(method, a, b, c, d, n, s))
    assert np.all(xf[nsys * n:] == 7.25)


# systems of one wavefront, of up to 512 lanes, and of 1024 lanes
@pytest.mark.parametrize("n,s", [(1024, 256), (4096, 256), (16384, 256)])
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_cached_x_stores_same_answers(method, n, s):
    """x_cached = 1 (x stored through the caches) and the default
    non-temporal stores write the same bits: the oracle's in exact mode, and
    each other's in fast mode."""
    nsys = 21
    say("x_cached", method, n, s)
    a, b, c, d, _ = O.gen_batch(17, 0, nsys, n)
    da, db, dc, dd = to_dev(a, b, c, d)
    ref = O.solve_batch(method, a, b, c, d, n, s)
    for arith in ("exact", "fast"):
        xs = []
        for cached in (False, True):
            dx = torch.full_like(da, float("nan"))
            B.solve_ex(da, db, dc, dd, dx, n, s, method=method, arith=arith, x_cached=cached)
            torch.cuda.synchronize()
            xs.append(dx.cpu().numpy())
        assert bits_equal(xs[0], xs[1])
        if arith == "exact":
            assert bits_equal(xs[1], ref)


@pytest.mark.parametrize("m", [1, 4, 16])
@pytest.mark.parametrize("s", [256, 512])
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_every_tuner_candidate_both_modes(method, s, m):
    """Every slice shape the batched tuner times (E * L = s, one workgroup
    per system) up to 16 slices per system: exact answers bit-identical to
    the oracle; fast answers within 1e-5 of it, and bit-identical to the
    default shape's wherever the shape runs the same fast stage 1 (one
    operation order per form: the two-level form of PTPR on 16-lane slices,
    the register PCR everywhere else), on four input kinds."""
    from tpr_amd.autotune import batched_candidates, default_batched_shape, fast_form
    n, nsys = s * m, 3
    default = default_batched_shape(s)
    for kind in (O.GEN_DOMINANT, O.GEN_WEAK, O.GEN_FIXED, O.GEN_RANDOM_RHS):
        a, b, c, d, _ = O.gen_batch(5 + kind, 0, nsys, n, kind)
        ref = O.solve_batch(method, a, b, c, d, n, s)
        x_def = gpu_solve(a, b, c, d, n, s, method, "fast", shape=default)
        for shape in batched_candidates(n, s):
            say("candidate", method, n, s, shape, kind)
            assert bits_equal(gpu_solve(a, b, c, d, n, s, method, "exact", shape=shape), ref)
            x = gpu_solve(a, b, c, d, n, s, method, "fast", shape=shape)
            if kind != O.GEN_WEAK:  # weakly dominant systems (condition up to ~1e4) exceed 1e-5 in fp32
                assert per_system_rel(x, ref, n).max() <= FAST_TOL
            if fast_form(method, *shape) == fast_form(method, *default):
                assert bits_equal(x, x_def)


def _tuned_rows_off_default():
    from tpr_amd import autotune as A
    out = []
    for key, row in sorted(A.load(A.BATCHED_PATH)["cases"].items()):
        method, arith, n, s, nsys = A.parse_batched_key(key)
        if tuple(row["choice"]) != A.default_batched_shape(s):
            out.append((method, arith, n, s, tuple(row["choice"]), A.default_batched_shape(s)))
    return sorted(set(out))


@pytest.mark.parametrize("method,arith,n,s,shape,default", _tuned_rows_off_default())
def test_tuned_default_shapes_keep_the_answers(method, arith, n, s, shape, default):
    """Every shape the tuned table makes a default (where it is not the
    library's plain default shape) gives the plain default's bits, in fast
    arithmetic too, on four input kinds: the table changes speed only, never
    a result, whatever the batch size."""
    for kind in (O.GEN_DOMINANT, O.GEN_WEAK, O.GEN_FIXED, O.GEN_RANDOM_RHS):
        say("tuned shape", method, arith, n, s, shape, kind)
        a, b, c, d, _ = O.gen_batch(21 + kind, 0, 3, n, kind)
        x = gpu_solve(a, b, c, d, n, s, method, arith, shape=shape)
        assert bits_equal(x, gpu_solve(a, b, c, d, n, s, method, arith, shape=default))


@pytest.mark.parametrize("n", [1024, 2048, 4096, 8192])
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_s512_default_shape_matches_16x32(method, n):
    """s = 512 takes 8 x 64 slices where that measured faster (PTPR up to 8
    slices per system, TPR up to 4) and 16 x 32 otherwise: in fast arithmetic
    too the default's answers equal the 16 x 32 answers bit for bit, so the
    choice never changes a result."""
    s, nsys = 512, 5
    say("s512 default", method, n)
    a, b, c, d, _ = O.gen_batch(11, 0, nsys, n)
    x_def = gpu_solve(a, b, c, d, n, s, method, "fast")
    x_ref = gpu_solve(a, b, c, d, n, s, method, "fast", shape=(16, 32))
    assert bits_equal(x_def, x_ref)


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_full_batch_4096x1024(method):
    """4096 systems x N = 1024, S = 256, generated on the GPU: exact is
    bit-identical to the oracle, fast within 1e-5 per system."""
    n, s, nsys = 1024, 256, 4096
    say("4096 x 1024", method)
    t = B.alloc_batch(nsys, n)
    B.generate("dominant", 2024, 0, nsys, n, t["a"], t["b"], t["c"], t["d"], t["xt"])
    a, b, c, d = (t[k].cpu().numpy() for k in "abcd")
    ref = O.solve_batch(method, a, b, c, d, n, s)
    B.solve_batched(t["a"], t["b"], t["c"], t["d"], t["x"], n, s, method=method, arith="exact")
    torch.cuda.synchronize()
    assert bits_equal(t["x"].cpu().numpy(), ref)
    B.solve_batched(t["a"], t["b"], t["c"], t["d"], t["x"], n, s, method=method, arith="fast")
    torch.cuda.synchronize()
    assert per_system_rel(t["x"].cpu().numpy(), ref, n).max() <= FAST_TOL


# ---------------------------------------------------------------------------
# edge cases the reference exercises or documents

@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s", [(64, 8), (1024, 256), (512, 16)])
def test_branch_systems(method, n, s):
    """c = 0 at slice ends: s1 = -1 (src/ptpr.cpp:196-200) and key = 0
    (src/ptpr.cpp:276-279) in every slice."""
    say("branch", method, n, s)
    a, b, c, d, _ = O.gen_batch(5, 0, 3, n)
    c.reshape(3, n)[:, s - 1::s] = 0.0
    a.reshape(3, n)[:, ::s] = 0.0
    for arith in ("exact", "fast"):
        x = gpu_solve(a, b, c, d, n, s, method, arith)
        ref = O.solve_batch(method, a, b, c, d, n, s)
        if arith == "exact":
            assert bits_equal(x, ref)
        else:
            assert per_system_rel(x, ref, n).max() <= FAST_TOL


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_diagonal_only(method):
    n, s = 512, 64
    rng = np.random.default_rng(1)
    b = rng.uniform(1, 2, n).astype(np.float32)
    d = rng.uniform(-1, 1, n).astype(np.float32)
    z = np.zeros(n, np.float32)
    x = gpu_solve(z, b, z, d, n, s, method, "exact")
    assert bits_equal(x, d / b)


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s", [(256, 256), (2, 2), (256, 16), (32768, 2048), (16384, 256)])
def test_edge_shapes(method, n, s):
    """m = 1 (s = n), the smallest system, N = s^2 and the largest reduced
    system one workgroup holds (m * L = 1024 lanes)."""
    say("edge", method, n, s)
    nsys = 3
    a, b, c, d, _ = O.gen_batch(8, 0, nsys, n)
    x = gpu_solve(a, b, c, d, n, s, method, "exact")
    assert bits_equal(x, O.solve_batch(method, a, b, c, d, n, s))


def test_empty_batch_is_a_noop():
    da = torch.zeros(1024, device="cuda")
    B.solve_batched(da[:0], None, da[:0], da[:0], da[:0], 1024, 256)


# ---------------------------------------------------------------------------
# reference surface: PTPR_CU::ptpr_cu & co., the solver classes

def test_cu_dropins_known_answer(golden):
    """test/cuexamples.cpp:63-83: tpr_cu / ptpr_cu for s = 64..1024 and
    cr_cu / pcr_cu against ans1024 -- and bit-identical to the CPU classes."""
    ans = golden["ans1024"]
    a, b, c, d, _ = O.gen_system(0, 0, 1024, O.GEN_FIXED)
    x = np.zeros(1024, np.float32)
    s = 64
    while s <= 1024:
        say("cu dropin s", s)
        tpr_amd.tpr_cu(a, c, d, x, 1024, s)
        assert np.all(np.abs(x - ans) <= 1e-2) and np.max((x - ans) ** 2) <= 1e-3
        assert bits_equal(x, O.tpr(a, c, d, s))
        tpr_amd.ptpr_cu(a, c, d, x, 1024, s)
        assert np.all(np.abs(x - ans) <= 1e-2) and np.max((x - ans) ** 2) <= 1e-3
        assert bits_equal(x, O.ptpr(a, c, d, s))
        s *= 2
    tpr_amd.cr_cu(a, c, d, x, 1024)
    assert np.all(np.abs(x - ans) <= 1e-2)
    assert bits_equal(x, O.cr(a, c, d))
    tpr_amd.pcr_cu(a, c, d, x, 1024)
    assert np.all(np.abs(x - ans) <= 1e-2)
    assert bits_equal(x, O.pcr(a, c, d))


def test_cu_dropins_from_many_threads():
    """The host-pointer entry points from 6 threads at once, each with its
    own systems and shapes (one-workgroup sizes, solved from mapped pinned
    memory, and a multi-workgroup size, copied): every answer is the CPU
    class's, bit for bit."""
    import threading
    cases = [(1024, 256, "ptpr"), (2048, 512, "tpr"), (4096, 64, "ptpr"), (65536, 256, "ptpr"),
             (512, 16, "tpr"), (32768, 2048, "ptpr")]
    errors = []

    def worker(k, n, s, method):
        try:
            fn = tpr_amd.ptpr_cu if method == "ptpr" else tpr_amd.tpr_cu
            ref_fn = O.ptpr if method == "ptpr" else O.tpr
            for rep in range(8):
                a, b, c, d, _ = O.gen_system(50 + k, rep, n)
                na, nc, nd = O.normalize(a, b, c, d)
                x = np.full(n, np.nan, np.float32)
                fn(na, nc, nd, x, n, s)
                if not bits_equal(x, ref_fn(na, nc, nd, s)):
                    errors.append((k, n, s, method, rep))
        except Exception as e:  # reported below, in the main thread
            errors.append((k, repr(e)))

    say("cu dropins from threads", len(cases))
    th = [threading.Thread(target=worker, args=(k,) + cs) for k, cs in enumerate(cases)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors


@pytest.mark.parametrize("n", [2, 64, 1024, 4096])
def test_pcr_cr_match_oracle(n):
    say("pcr/cr", n)
    a, b, c, d, _ = O.gen_system(6, 1, n)
    na, nc, nd = O.normalize(a, b, c, d)
    x = np.zeros(n, np.float32)
    tpr_amd.pcr_cu(na, nc, nd, x, n)
    assert bits_equal(x, O.pcr(na, nc, nd))
    tpr_amd.cr_cu(na, nc, nd, x, n)
    assert bits_equal(x, O.cr(na, nc, nd))


@pytest.mark.parametrize("cls,method", [(tpr_amd.PTPR, "ptpr"), (tpr_amd.TPR, "tpr")])
def test_solver_class_surface(cls, method):
    """PTPR(n, s) / TPR(n, s): 3- and 4-argument setters, solve() returns the
    reference's flop count, get_ans fills x; caller arrays are not modified."""
    n, s = 2048, 512
    a, b, c, d, _ = O.gen_system(12, 3, n)
    keep = [v.copy() for v in (a, b, c, d)]
    t = cls(n, s)
    t.set_tridiagonal_system(a, b, c, d)
    flops = t.solve()
    exp_flops = O.ptpr_flops(n, s) if method == "ptpr" else O.tpr_flops(n, s)
    assert flops == exp_flops
    x = np.zeros(n, np.float32)
    assert t.get_ans(x) == 0
    for v, k in zip((a, b, c, d), keep):
        assert np.array_equal(v, k)
    na, nc, nd = O.normalize(a, b, c, d)
    ref = (O.ptpr if method == "ptpr" else O.tpr)(na, nc, nd, s)
    assert bits_equal(x, ref)
    t.set_tridiagonal_system(na, nc, nd)
    t.solve()
    t.get_ans(x)
    assert bits_equal(x, ref)
    t.set_arith("fast")
    t.solve()
    t.get_ans(x)
    assert O.rel_l2(x, ref) <= FAST_TOL


def test_pcr_cr_class_surface():
    """PCR's 4-argument setter divides by diag (src/pcr.hpp:42-53); CR's
    ignores it (src/cr.cpp:17-21)."""
    n = 512
    a, b, c, d, _ = O.gen_system(2, 2, n)
    na, nc, nd = O.normalize(a, b, c, d)
    x = np.zeros(n, np.float32)
    p = tpr_amd.PCR(n)
    p.set_tridiagonal_system(a, b, c, d)
    p.solve()
    p.get_ans(x)
    assert bits_equal(x, O.pcr(na, nc, nd))
    q = tpr_amd.CR(n)
    q.set_tridiagonal_system(na, b, nc, nd)
    q.solve()
    q.get_ans(x)
    assert bits_equal(x, O.cr(na, nc, nd))


@pytest.mark.parametrize("k", range(1, 21))
def test_nstab_matrices(golden, k):
    """The 20 stability classes of src/inputs (including NaN/Inf cases) give
    the reference's bits."""
    a, b, c, d = golden["nstab"]["eq%d" % k]
    for s in (16, 64, 512):
        for method in ("ptpr", "tpr"):
            x = gpu_solve(a, b, c, d, 512, s, method, "exact")
            assert bits_equal(x, golden["ref"]["eq%d_s%d_%s" % (k, s, method)]), (k, s, method)


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("arith", ["exact", "fast"])
@pytest.mark.parametrize("n,s,nsys", [(1024, 256, 4097), (256, 16, 1001), (64, 8, 333), (512, 2, 77)])
def test_ragged_batches(method, arith, n, s, nsys):
    """Batch sizes that leave the last workgroup partly empty."""
    say("ragged", method, arith, n, s, nsys)
    t = B.alloc_batch(nsys, n)
    B.generate("dominant", 77, 5, nsys, n, t["a"], t["b"], t["c"], t["d"], t["xt"])
    B.solve_batched(t["a"], t["b"], t["c"], t["d"], t["x"], n, s, method=method, arith=arith)
    torch.cuda.synchronize()
    a, b, c, d = (t[k].cpu().numpy() for k in "abcd")
    ref = O.solve_batch(method, a, b, c, d, n, s)
    if arith == "exact":
        assert bits_equal(t["x"].cpu().numpy(), ref)
    else:
        assert per_system_rel(t["x"].cpu().numpy(), ref, n).max() <= FAST_TOL


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_graph_capture_replays_match_eager(method):
    """The batched solve captured in a HIP graph (as bench.py times it):
    every replay, with eager work and synchronisations in between, equals
    the eager result bit for bit."""
    n, s, nsys = 1024, 256, 512
    say("graph", method)
    t = B.alloc_batch(nsys, n)
    B.generate("dominant", 5, 0, nsys, n, t["a"], t["b"], t["c"], t["d"], t["xt"])
    eager = torch.empty_like(t["x"])
    B.solve_batched(t["a"], t["b"], t["c"], t["d"], eager, n, s, method=method)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        B.solve_batched(t["a"], t["b"], t["c"], t["d"], t["x"], n, s, method=method)
    ref = eager.cpu().numpy()
    for rep in range(4):
        t["x"].fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert bits_equal(t["x"].cpu().numpy(), ref), rep
        # eager work between replays
        B.solve_batched(t["a"], t["b"], t["c"], t["d"], eager, n, s, method=method)
        torch.cuda.synchronize()
