"""GPU parity of the general entry point (tprh_solve_ex): systems larger than
one workgroup (split path), fp16-stored coefficients and the FP64-residual
refinement step.

Every case runs the exact arithmetic against the oracle bit for bit:
* split path: the same operations as the one-workgroup path, distributed;
* fp16 storage: the kernel widens fp16 to fp32 exactly, so it must equal the
  oracle run on the widened coefficients;
* refinement: x1 = x0 + A^-1 fl32(d - A x0) with the residual in FP64 in the
  oracle's operation order (orc_residual_f64) -- identical bits.
The fast arithmetic is checked against the same references at rel-L2 <= 1e-5.
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


def per_system_rel(x, ref, n):
    x = x.reshape(-1, n).astype(np.float64)
    ref = ref.reshape(-1, n).astype(np.float64)
    return np.linalg.norm(x - ref, axis=1) / np.linalg.norm(ref, axis=1)


def dev(v, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(v)).to(dtype).cuda()


SPLIT_CASES = [(65536, 256, 3, 0), (32768, 128, 4, 3), (16384, 16, 5, 2), (131072, 1024, 2, 1),
               (8192, 4, 3, 0), (65536, 2048, 2, 0), (4096, 2, 2, 1)]


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s,nsys,chunk", SPLIT_CASES)
def test_split_exact_bit_identical(method, n, s, nsys, chunk):
    say("split exact", method, n, s, nsys, chunk)
    a, b, c, d, xt = O.gen_batch(31, 0, nsys, n, O.GEN_WEAK)
    x = torch.empty(nsys * n, device="cuda")
    B.solve_ex(dev(a), dev(b), dev(c), dev(d), x, n, s, method=method, arith="exact", chunk=chunk)
    torch.cuda.synchronize()
    assert bits_equal(x.cpu().numpy(), O.solve_batch(method, a, b, c, d, n, s))


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s,nsys,chunk", [(65536, 256, 3, 2), (32768, 512, 2, 0), (16384, 128, 3, 0)])
def test_split_staged_stores_in_bounds(method, n, s, nsys, chunk):
    """The two-pass form's staged loads and stores (E = 8 and 16): exact
    answers bit for bit and nothing written past the batch's x."""
    say("split staged", method, n, s, nsys, chunk)
    a, b, c, d, _ = O.gen_batch(37, 0, nsys, n, O.GEN_WEAK)
    guard = 4096
    xfull = torch.full((nsys * n + guard,), -3.5, dtype=torch.float32, device="cuda")
    B.solve_ex(dev(a), dev(b), dev(c), dev(d), xfull[:nsys * n], n, s, method=method, arith="exact", chunk=chunk,
               split_form=1)
    torch.cuda.synchronize()
    xf = xfull.cpu().numpy()
    assert bits_equal(xf[:nsys * n], O.solve_batch(method, a, b, c, d, n, s))
    assert np.all(xf[nsys * n:] == -3.5)


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s", [(1 << 22, 2048), (1 << 19, 256)])
def test_split_maximum_sizes(method, n, s):
    """The largest systems the engine takes: m = n / s = 2048 slices (the
    multi-workgroup path's limit) at the largest slice size, 4 Mi equations,
    and at s = 256; exact answers bit-identical, fast within tolerance, and
    one slice more is refused."""
    say("split maximum", method, n, s)
    a, b, c, d, xt = O.gen_batch(33, 0, 1, n, O.GEN_WEAK)
    ref = O.solve_batch(method, a, b, c, d, n, s)
    x = torch.empty(n, device="cuda")
    da, db, dc, dd = dev(a), dev(b), dev(c), dev(d)
    B.solve_ex(da, db, dc, dd, x, n, s, method=method, arith="exact")
    torch.cuda.synchronize()
    assert bits_equal(x.cpu().numpy(), ref)
    B.solve_ex(da, db, dc, dd, x, n, s, method=method, arith="fast")
    torch.cuda.synchronize()
    assert per_system_rel(x.cpu().numpy(), ref, n).max() <= FAST_TOL
    with pytest.raises(tpr_amd.TprError):
        B.solve_ex(da, db, dc, dd, x, 2 * n, s, method=method, arith="exact")


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s,nsys,chunk", SPLIT_CASES[:4])
def test_split_fast_within_tolerance(method, n, s, nsys, chunk):
    say("split fast", method, n, s, nsys, chunk)
    a, b, c, d, xt = O.gen_batch(32, 0, nsys, n)
    x = torch.empty(nsys * n, device="cuda")
    B.solve_ex(dev(a), dev(b), dev(c), dev(d), x, n, s, method=method, arith="fast", chunk=chunk)
    torch.cuda.synchronize()
    ref = O.solve_batch(method, a, b, c, d, n, s)
    assert per_system_rel(x.cpu().numpy(), ref, n).max() <= FAST_TOL


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_split_unit_diagonal_and_host_path(method):
    """3-argument form through the split path, and the host-pointer entry
    point (PTPR_CU::ptpr_cu surface) for an n the reference's GPU path cannot
    solve (n / s > s)."""
    n, s = 65536, 128
    say("split unit diag", method)
    a, b, c, d, _ = O.gen_system(40, 0, n)
    na, nc, nd = O.normalize(a, b, c, d)
    ref = (O.ptpr if method == "ptpr" else O.tpr)(na, nc, nd, s)
    x = torch.empty(n, device="cuda")
    B.solve_ex(dev(na), None, dev(nc), dev(nd), x, n, s, method=method, arith="exact")
    torch.cuda.synchronize()
    assert bits_equal(x.cpu().numpy(), ref)
    xh = np.zeros(n, np.float32)
    (tpr_amd.ptpr_cu if method == "ptpr" else tpr_amd.tpr_cu)(na, nc, nd, xh, n, s)
    assert bits_equal(xh, ref)


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s,nsys", [(1024, 256, 65), (65536, 256, 2), (2048, 16, 9)])
def test_fp16_storage_is_exact_on_widened_coefficients(method, n, s, nsys):
    say("fp16", method, n, s, nsys)
    a, b, c, d, xt = O.gen_batch(33, 0, nsys, n, O.GEN_WEAK)
    a16, b16, c16 = (v.astype(np.float16) for v in (a, b, c))
    x = torch.empty(nsys * n, device="cuda")
    B.solve_ex(dev(a16, torch.float16), dev(b16, torch.float16), dev(c16, torch.float16), dev(d), x, n, s,
               method=method, arith="exact")
    torch.cuda.synchronize()
    ref = O.solve_batch(method, a16.astype(np.float32), b16.astype(np.float32), c16.astype(np.float32), d, n, s)
    assert bits_equal(x.cpu().numpy(), ref)


def oracle_refine(method, a, b, c, d, x0, n, s):
    """x0 + A^-1 fl32(d - A x0), residual in FP64 (orc_residual_f64)."""
    nsys = a.size // n
    r = np.empty_like(d)
    for k in range(nsys):
        sl = slice(k * n, (k + 1) * n)
        r[sl] = O.residual_f64(a[sl], b[sl], c[sl], d[sl], x0[sl].astype(np.float64)).astype(np.float32)
    dx = O.solve_batch(method, a, b, c, r, n, s)
    return (x0 + dx).astype(np.float32)


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s,nsys", [(1024, 256, 33), (65536, 256, 2), (512, 16, 7)])
def test_refinement_step_bit_identical(method, n, s, nsys):
    say("refine", method, n, s, nsys)
    a, b, c, d, xt = O.gen_batch(34, 0, nsys, n, O.GEN_WEAK)
    da, db, dc, dd = dev(a), dev(b), dev(c), dev(d)
    x = torch.empty(nsys * n, device="cuda")
    B.solve_ex(da, db, dc, dd, x, n, s, method=method, arith="exact")
    torch.cuda.synchronize()
    x0 = x.cpu().numpy()
    assert bits_equal(x0, O.solve_batch(method, a, b, c, d, n, s))
    B.solve_ex(da, db, dc, dd, x, n, s, method=method, arith="exact", refine=True)
    torch.cuda.synchronize()
    x1 = x.cpu().numpy()
    assert bits_equal(x1, oracle_refine(method, a, b, c, d, x0, n, s))


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_refinement_reaches_fp64_accuracy(method):
    """Weakly dominant systems (the `weak` generator, smaller n): the fp16-stored
    solve is the least accurate, fp32 next; refinement steps against the fp32
    system bring both to the FP64 Thomas solution within fp32 rounding."""
    n, s, nsys = 16384, 256, 8
    say("refine accuracy", method)
    a, b, c, d, xt = O.gen_batch(35, 0, nsys, n, O.GEN_WEAK)
    xf = O.thomas_f64_batch(a, b, c, d, n)
    da, db, dc, dd = dev(a), dev(b), dev(c), dev(d)
    x = torch.empty(nsys * n, device="cuda")
    errs = {}
    B.solve_ex(dev(a.astype(np.float16), torch.float16), dev(b.astype(np.float16), torch.float16),
               dev(c.astype(np.float16), torch.float16), dd, x, n, s, method=method)
    torch.cuda.synchronize()
    errs["f16"] = per_system_rel(x.cpu().numpy(), xf, n).max()
    for step in (1, 2, 3):
        B.solve_ex(da, db, dc, dd, x, n, s, method=method, refine=True)
        torch.cuda.synchronize()
        errs["f16+%d" % step] = per_system_rel(x.cpu().numpy(), xf, n).max()
    B.solve_ex(da, db, dc, dd, x, n, s, method=method)
    torch.cuda.synchronize()
    errs["f32"] = per_system_rel(x.cpu().numpy(), xf, n).max()
    B.solve_ex(da, db, dc, dd, x, n, s, method=method, refine=True)
    torch.cuda.synchronize()
    errs["f32+1"] = per_system_rel(x.cpu().numpy(), xf, n).max()
    say(errs)
    assert errs["f32+1"] <= errs["f32"]
    assert errs["f32+1"] <= 2e-7
    assert errs["f16+3"] <= 2e-7
    assert errs["f16"] > errs["f32"]


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s,nrhs,with_b", [(1024, 256, 37, True), (256, 16, 9, False), (65536, 256, 3, True),
                                             (16384, 128, 4, True)])
def test_shared_matrix_many_rhs_bit_identical(method, n, s, nrhs, with_b):
    """coef_shared: nrhs right-hand sides against one matrix (ADI-style
    reuse), through the one-workgroup and the multi-workgroup path; each
    answer equals the oracle's solve of that right-hand side, bit for bit."""
    say("shared matrix", method, n, s, nrhs, with_b)
    a, b, c, _, _ = O.gen_system(36, 5, n, O.GEN_WEAK)
    rhs = np.concatenate([O.gen_rhs(36, 5, j, n) for j in range(nrhs)])
    if not with_b:
        a, c, _ = O.normalize(a, b, c, np.zeros(n, np.float32))
    x = torch.empty(nrhs * n, device="cuda")
    B.solve_ex(dev(a), dev(b) if with_b else None, dev(c), dev(rhs), x, n, s, method=method, arith="exact",
               shared=True)
    torch.cuda.synchronize()
    got = x.cpu().numpy().reshape(nrhs, n)
    for j in range(nrhs):
        r = rhs[j * n:(j + 1) * n]
        if with_b:
            ref = O.solve_batch(method, a, b, c, r, n, s)
        else:
            ref = (O.ptpr if method == "ptpr" else O.tpr)(a, c, r.copy(), s)
        assert bits_equal(got[j], ref), j


def test_shared_matrix_fp16_and_refinement():
    n, s, nrhs = 2048, 256, 6
    a, b, c, _, _ = O.gen_system(37, 2, n, O.GEN_WEAK)
    rhs = np.concatenate([O.gen_rhs(37, 2, j, n) for j in range(nrhs)])
    x = torch.empty(nrhs * n, device="cuda")
    a16, b16, c16 = (v.astype(np.float16) for v in (a, b, c))
    B.solve_ex(dev(a16, torch.float16), dev(b16, torch.float16), dev(c16, torch.float16), dev(rhs), x, n, s,
               arith="exact", shared=True)
    torch.cuda.synchronize()
    got = x.cpu().numpy().reshape(nrhs, n)
    wa, wb, wc = (v.astype(np.float32) for v in (a16, b16, c16))
    for j in range(nrhs):
        assert bits_equal(got[j], O.solve_batch("ptpr", wa, wb, wc, rhs[j * n:(j + 1) * n], n, s))
    # one refinement step against the fp32 matrix, every right-hand side
    x0 = x.cpu().numpy()
    B.solve_ex(dev(a), dev(b), dev(c), dev(rhs), x, n, s, arith="exact", refine=True, shared=True)
    torch.cuda.synchronize()
    x1 = x.cpu().numpy().reshape(nrhs, n)
    for j in range(nrhs):
        sl = slice(j * n, (j + 1) * n)
        ref = oracle_refine("ptpr", a, b, c, rhs[sl], x0[sl], n, s)
        assert bits_equal(x1[j], ref), j


def test_solve_ex_side_stream_keeps_its_workspace():
    """solve_ex allocates the split path's workspace itself on the current
    stream and returns while its kernels may still run on `stream`: the
    caching allocator must not hand those bytes to an allocation made on
    the current stream at once.  Answers equal a same-stream solve."""
    n, s, nsys, chunk = 1 << 18, 256, 4, 1
    say("side stream workspace", n, s, nsys, chunk)
    a, b, c, d, _ = O.gen_batch(38, 0, nsys, n, O.GEN_WEAK)
    da, db, dc, dd = dev(a), dev(b), dev(c), dev(d)
    x_ref = torch.empty(nsys * n, device="cuda")
    B.solve_ex(da, db, dc, dd, x_ref, n, s, arith="exact", chunk=chunk)
    torch.cuda.synchronize()
    need = B.workspace_bytes(B.make_desc(da, db, dc, dd, x_ref, n, s, arith="exact", chunk=chunk))
    assert need > 0
    side = torch.cuda.Stream()
    x = torch.zeros(nsys * n, device="cuda")
    side.wait_stream(torch.cuda.current_stream())
    for _ in range(3):
        B.solve_ex(da, db, dc, dd, x, n, s, arith="exact", chunk=chunk, stream=side)
        # allocations of the workspace's size, filled at once on the current stream
        junk = [torch.full((need // 4 + 1,), float("nan"), device="cuda") for _ in range(4)]
        del junk
    torch.cuda.synchronize()
    assert bits_equal(x.cpu().numpy(), x_ref.cpu().numpy())


@pytest.mark.parametrize("form", [1, 2, 3])
@pytest.mark.parametrize("method", ["ptpr", "tpr"])
@pytest.mark.parametrize("n,s,nsys,chunk", [SPLIT_CASES[0], SPLIT_CASES[2], SPLIT_CASES[3], SPLIT_CASES[6],
                                            (65536, 256, 70, 0)])
def test_split_forms_bit_identical(form, method, n, s, nsys, chunk):
    """Every form of the multi-workgroup path, forced: two passes (stage 1
    and the reduced systems, then stage 3), one pass (slices in registers
    throughout; 70 systems of 65536 are more blocks than the chip holds at
    once, so its blocks come in several waves) and pipelined (rows tasks and
    stage-3 tasks in one launch, the system's last arriver solving its
    reduced system): exact answers bit-identical to the oracle, fast ones
    within tolerance, with and without a refinement step."""
    say("split form", form, method, n, s, nsys, chunk)
    a, b, c, d, xt = O.gen_batch(39, 0, nsys, n, O.GEN_WEAK)
    da, db, dc, dd = dev(a), dev(b), dev(c), dev(d)
    x = torch.empty(nsys * n, device="cuda")
    B.solve_ex(da, db, dc, dd, x, n, s, method=method, arith="exact", chunk=chunk, split_form=form)
    torch.cuda.synchronize()
    x0 = x.cpu().numpy()
    assert bits_equal(x0, O.solve_batch(method, a, b, c, d, n, s))
    B.solve_ex(da, db, dc, dd, x, n, s, method=method, arith="exact", chunk=chunk, split_form=form, refine=True)
    torch.cuda.synchronize()
    if nsys <= 8:
        assert bits_equal(x.cpu().numpy(), oracle_refine(method, a, b, c, d, x0, n, s))
    xf = torch.empty_like(x)
    ga, gb, gc, gd, _ = O.gen_batch(40, 0, nsys, n)
    B.solve_ex(dev(ga), dev(gb), dev(gc), dev(gd), xf, n, s, method=method, arith="fast", chunk=chunk,
               split_form=form)
    torch.cuda.synchronize()
    ref = O.solve_batch(method, ga, gb, gc, gd, n, s)
    assert per_system_rel(xf.cpu().numpy(), ref, n).max() <= FAST_TOL
