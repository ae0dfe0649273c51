"""tprh_solve_ex_noalloc, the entry point that allocates nothing: its
multi-workgroup workspace is the library's fixed device array.  Checked
through the C ABI against tprh_solve_ex (caller-provided workspace) on the
same inputs -- bit for bit -- on consecutive calls, the second smaller than
the first, so that the array's contents from one call cannot leak into the
next; rounds cut to fit the array; and the refusal of an explicit chunk
whose round does not fit (before any device work)."""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O
from tpr_amd import _lib
from tpr_amd import batched as B

SCRATCH = 16 << 20


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def bits_equal(x, y):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    same = x.view(np.uint32) == y.view(np.uint32)
    return bool(np.all(same | (np.isnan(x) & np.isnan(y))))


def noalloc(desc, stream=None):
    import torch
    st = _lib.lib().tprh_solve_ex_noalloc(ctypes.byref(desc), ctypes.c_void_p(
        (stream or torch.cuda.current_stream()).cuda_stream))
    return st


def test_noalloc_refuses_an_explicit_chunk_that_does_not_fit():
    """No device is touched: the plan is refused first (runs without a GPU)."""
    desc = _lib.SolveDesc()
    desc.method, desc.arith = _lib.PTPR_M, _lib.ARITH_EXACT
    desc.n, desc.s, desc.nsys = 32768, 16, 300
    desc.coef = _lib.STORE_F32
    desc.a = desc.c = desc.d = desc.x = 1 << 20  # never dereferenced
    desc.chunk = 300  # 300 * 2048 slices * 36 B > 16 MiB
    L = _lib.lib()
    assert L.tprh_solve_ex_noalloc(ctypes.byref(desc), None) == _lib.ERROR_NOT_SUPPORTED
    assert b"fixed buffer" in L.tprh_last_error()
    desc.chunk = 0
    desc.workspace, desc.workspace_bytes = 1 << 20, 4096
    assert L.tprh_solve_ex_noalloc(ctypes.byref(desc), None) == _lib.ERROR_INVALID_VALUE


CASES = [
    # (n, s, nsys, kind, arith, method): the first needs two rounds to fit the
    # array (300 systems of 2048 slices), the second is smaller, the third
    # needs no workspace at all
    (32768, 16, 300, O.GEN_WEAK, "exact", "ptpr"),
    (65536, 256, 4, O.GEN_WEAK, "exact", "tpr"),
    (1024, 256, 64, O.GEN_DOMINANT, "fast", "ptpr"),
]


@pytest.mark.gpu
def test_noalloc_equals_solve_ex_on_consecutive_calls():
    import torch
    for (n, s, nsys, kind, arith, method) in CASES:
        say("noalloc", n, s, nsys, arith, method)
        a, b, c, d, _ = O.gen_batch(61, 0, nsys, n, kind)
        da, db, dc, dd = (torch.from_numpy(v).cuda() for v in (a, b, c, d))
        x_ref = torch.full((nsys * n,), float("nan"), device="cuda")
        B.solve_ex(da, db, dc, dd, x_ref, n, s, method=method, arith=arith)
        x = torch.full((nsys * n,), float("nan"), device="cuda")
        desc = B.make_desc(da, db, dc, dd, x, n, s, method=method, arith=arith)
        assert noalloc(desc) == _lib.SUCCESS, _lib.lib().tprh_last_error()
        torch.cuda.synchronize()
        got, ref = x.cpu().numpy(), x_ref.cpu().numpy()
        assert not np.isnan(ref).any()
        assert bits_equal(got, ref), (n, s, nsys)
        if arith == "exact":
            assert bits_equal(got, O.solve_batch(method, a, b, c, d, n, s)), (n, s, nsys)


@pytest.mark.gpu
def test_noalloc_refinement_and_fp16_match_solve_ex():
    import torch
    n, s, nsys = 65536, 256, 3
    say("noalloc refine/fp16", n, s, nsys)
    a, b, c, d, _ = O.gen_batch(62, 0, nsys, n, O.GEN_WEAK)
    da, db, dc, dd = (torch.from_numpy(v).cuda() for v in (a, b, c, d))
    h = [torch.from_numpy(v.astype(np.float16)).cuda() for v in (a, b, c)]
    x_ref = torch.empty(nsys * n, device="cuda")
    x = torch.empty(nsys * n, device="cuda")
    B.solve_ex(h[0], h[1], h[2], dd, x_ref, n, s)
    assert noalloc(B.make_desc(h[0], h[1], h[2], dd, x, n, s)) == _lib.SUCCESS
    torch.cuda.synchronize()
    assert bits_equal(x.cpu().numpy(), x_ref.cpu().numpy())
    B.solve_ex(da, db, dc, dd, x_ref, n, s, refine=True)
    assert noalloc(B.make_desc(da, db, dc, dd, x, n, s, refine=True)) == _lib.SUCCESS
    torch.cuda.synchronize()
    assert bits_equal(x.cpu().numpy(), x_ref.cpu().numpy())
