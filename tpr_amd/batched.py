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


def shape_supported(E, L):
    return bool(_lib.lib().tprh_shape_supported(int(E), int(L)))


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
