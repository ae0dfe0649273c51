"""Device-resident batched solves and the synthetic generator.

Arrays are CUDA (HIP) torch tensors of float32, ``nsys * n`` elements, system
k at ``[k*n, (k+1)*n)``.  PyTorch provides device memory and the stream; the
solve itself is one launch of a hand-written gfx950 kernel in libtpr_hip.so.
"""
import ctypes

from . import _lib

METHODS = {"ptpr": _lib.PTPR_M, "tpr": _lib.TPR_M}
ARITH = {"fast": _lib.ARITH_FAST, "exact": _lib.ARITH_EXACT}
GEN = {"dominant": _lib.GEN_DOMINANT, "weak": _lib.GEN_WEAK, "fixed": _lib.GEN_FIXED,
       "random_rhs": _lib.GEN_RANDOM_RHS}


def _dptr(t, name, numel=None):
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % name)
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError("%s must be a contiguous float32 GPU tensor" % name)
    if numel is not None and t.numel() < numel:
        raise ValueError("%s has %d elements, need %d" % (name, t.numel(), numel))
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def solve_batched(a, b, c, d, x, n, s, method="ptpr", arith="fast", shape=None, stream=None):
    """Solve nsys = numel // n systems; b may be None (unit diagonal).

    shape: optional (E, L) slice shape (equations per lane, lanes per slice).
    """
    nsys = a.numel() // n
    tot = nsys * n
    args = [_dptr(a, "a", tot), _dptr(b, "b", tot), _dptr(c, "c", tot), _dptr(d, "d", tot), _dptr(x, "x", tot)]
    if shape is None:
        _lib.call("tprh_solve_batched", METHODS[method], ARITH[arith], *args, int(n), int(s), int(nsys),
                  _stream(stream))
    else:
        E, L = shape
        _lib.call("tprh_solve_batched_shape", METHODS[method], ARITH[arith], *args, int(n), int(s), int(nsys),
                  int(E), int(L), _stream(stream))
    return x


def _dptr_any(t, name, numel, dtypes):
    import torch
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous() or t.dtype not in dtypes:
        raise TypeError("%s must be a contiguous GPU tensor of %s" % (name, dtypes))
    if t.numel() < numel:
        raise ValueError("%s has %d elements, need %d" % (name, t.numel(), numel))
    return t.data_ptr()


def make_desc(a, b, c, d, x, n, s, method="ptpr", arith="fast", refine=False, shape=None, chunk=0,
              shared=False, split_form=0, x_cached=False):
    """tprh_solve_desc for torch tensors; a, b, c float32 or float16 (same
    dtype).  shared: a, b, c hold one matrix (n values) for every right-hand
    side in d.  x_cached: store x through the caches (for a caller whose
    next kernel reads x) instead of non-temporally."""
    import torch
    nsys = d.numel() // n
    tot = nsys * n
    ctot = n if shared else tot
    coef16 = a.dtype == torch.float16
    cd = (torch.float16,) if coef16 else (torch.float32,)
    desc = _lib.SolveDesc()
    desc.method = METHODS[method]
    desc.arith = ARITH[arith]
    desc.n, desc.s, desc.nsys = int(n), int(s), int(nsys)
    desc.coef = _lib.STORE_F16 if coef16 else _lib.STORE_F32
    desc.a = _dptr_any(a, "a", ctot, cd)
    desc.b = _dptr_any(b, "b", ctot, cd)
    desc.c = _dptr_any(c, "c", ctot, cd)
    desc.d = _dptr_any(d, "d", tot, (torch.float32,))
    desc.x = _dptr_any(x, "x", tot, (torch.float32,))
    desc.refine = 1 if refine else 0
    desc.E, desc.L = shape if shape is not None else (0, 0)
    desc.chunk = int(chunk)
    desc.coef_shared = 1 if shared else 0
    desc.split_form = int(split_form)
    desc.x_cached = 1 if x_cached else 0
    return desc


def workspace_bytes(desc):
    v = _lib.lib().tprh_workspace_bytes(ctypes.byref(desc))
    if v == ctypes.c_size_t(-1).value:
        _lib.check(_lib.ERROR_INVALID_VALUE, "tprh_workspace_bytes")
    return v


def solve_ex(a, b, c, d, x, n, s, method="ptpr", arith="fast", refine=False, shape=None, chunk=0,
             workspace=None, stream=None, shared=False, split_form=0, x_cached=False):
    """The general entry point (tprh_solve_ex): large systems (through a
    workspace, allocated here from torch if not given), fp16 coefficient
    storage (a, b, c as float16 tensors), one refinement step (refine=True:
    x += A^-1 (d - A x) with an FP64 residual), many right-hand sides of one
    matrix (shared=True: a, b, c of n values, d and x of nsys * n); split_form
    0 / 1 / 2 / 3: the library's choice / two passes / one pass / pipelined
    for systems above one workgroup; x_cached=True stores x through the
    caches instead of non-temporally (same answers)."""
    import torch
    desc = make_desc(a, b, c, d, x, n, s, method, arith, refine, shape, chunk, shared, split_form, x_cached)
    need = workspace_bytes(desc)
    own = None
    if need:
        if workspace is None:
            workspace = own = torch.empty(need, dtype=torch.uint8, device=d.device)
        if workspace.numel() * workspace.element_size() < need:
            raise ValueError("workspace smaller than %d bytes" % need)
        desc.workspace = workspace.data_ptr()
        desc.workspace_bytes = workspace.numel() * workspace.element_size()
    _lib.call("tprh_solve_ex", ctypes.byref(desc), _stream(stream))
    if own is not None and stream is not None:
        # the workspace was allocated on the current stream and is freed when
        # this function returns, while the kernels on `stream` may still use
        # it: tell the caching allocator to wait for `stream` before reuse
        own.record_stream(stream)
    return x


def shape_supported(E, L):
    return bool(_lib.lib().tprh_shape_supported(int(E), int(L)))


def slice_shape(n, s, nsys=1, method="ptpr", arith="fast"):
    """(E, L): the slice shape solve_batched takes for these arguments (no
    device call)."""
    E, L = ctypes.c_int(), ctypes.c_int()
    _lib.call("tprh_slice_shape", METHODS[method], ARITH[arith], int(n), int(s), int(nsys), ctypes.byref(E),
              ctypes.byref(L))
    return E.value, L.value


def generate(kind, seed, sys0, nsys, n, a, b, c, d, xt=None, stream=None):
    """Fill device tensors with the seeded synthetic systems sys0 .. sys0+nsys-1."""
    tot = nsys * n
    _lib.call("tprh_generate", GEN[kind], int(seed), int(sys0), int(nsys), int(n), _dptr(a, "a", tot),
              _dptr(b, "b", tot), _dptr(c, "c", tot), _dptr(d, "d", tot), _dptr(xt, "xt", tot), _stream(stream))


def alloc_batch(nsys, n, device="cuda", with_xt=True):
    """Device tensors a, b, c, d, x (and xt) for nsys systems of n equations."""
    import torch
    tot = nsys * n
    names = ["a", "b", "c", "d", "x"] + (["xt"] if with_xt else [])
    return {k: torch.empty(tot, dtype=torch.float32, device=device) for k in names}
