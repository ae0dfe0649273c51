"""Parity oracle for the TPR/PTPR solvers -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker.  The product
(``tpr_amd``) never imports it and has no CPU fallback.

Two CPU implementations are exposed:

* ``liboracle.so`` -- ``tpr_oracle.c``, a C restatement of the reference's CPU
  solvers (src/ptpr.cpp, src/tpr.cpp, src/pcr.cpp, src/cr.cpp) plus the seeded
  generator, an FP64 Thomas solve, an FP64 residual and the rel-L2 metric of
  src/numerical_stability.cpp:100-115.
* ``_ref/libtprref.so`` -- the reference's own solver sources compiled in
  place (see ``Makefile``), reached through a small C harness.  It exists only
  where it was built; ``ref_available()`` says whether it loaded.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libtprref.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i = ctypes.c_int
_u64 = ctypes.c_uint64

GEN_DOMINANT = 0
GEN_WEAK = 1
GEN_FIXED = 2
GEN_RANDOM_RHS = 3

_orc = None
_ref = None


def _load_oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(_ORC_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        lib = ctypes.CDLL(_ORC_PATH)
        for name in ("orc_ptpr_solve", "orc_tpr_solve"):
            getattr(lib, name).argtypes = [_f32p, _f32p, _f32p, _f32p, _i, _i]
            getattr(lib, name).restype = _i
        for name in ("orc_pcr_solve", "orc_cr_solve"):
            getattr(lib, name).argtypes = [_f32p, _f32p, _f32p, _f32p, _i]
            getattr(lib, name).restype = _i
        lib.orc_normalize.argtypes = [_f32p, _f32p, _f32p, _f32p, _i]
        lib.orc_normalize.restype = None
        lib.orc_solve_batch.argtypes = [_i, _f32p, _vp, _f32p, _f32p, _f32p, _i, _i, _i, _i]
        lib.orc_solve_batch.restype = _i
        lib.orc_norm.argtypes = [_f32p, _f32p, _i]
        lib.orc_norm.restype = ctypes.c_double
        lib.orc_thomas_f64.argtypes = [_f32p, _f32p, _f32p, _f32p, _f64p, _i]
        lib.orc_thomas_f64.restype = None
        lib.orc_thomas_f64_batch.argtypes = [_f32p, _f32p, _f32p, _f32p, _f64p, _i, _i]
        lib.orc_thomas_f64_batch.restype = None
        lib.orc_residual_f64.argtypes = [_f32p, _f32p, _f32p, _f32p, _f64p, _f64p, _i]
        lib.orc_residual_f64.restype = None
        lib.orc_gen_system.argtypes = [_u64, _u64, _i, _i, _f32p, _f32p, _f32p, _f32p, _vp]
        lib.orc_gen_system.restype = None
        lib.orc_gen_batch.argtypes = [_u64, _u64, _i, _i, _i, _f32p, _f32p, _f32p, _f32p, _vp]
        lib.orc_gen_batch.restype = None
        lib.orc_gen_rhs.argtypes = [_u64, _u64, _i, _i, _f32p]
        lib.orc_gen_rhs.restype = None
        lib.orc_ptpr_flops.argtypes = [_i, _i]
        lib.orc_ptpr_flops.restype = _i
        lib.orc_tpr_flops.argtypes = [_i, _i]
        lib.orc_tpr_flops.restype = _i
        lib.orc_max_threads.argtypes = []
        lib.orc_max_threads.restype = _i
        _orc = lib
    return _orc


def _load_ref():
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        lib = ctypes.CDLL(_REF_PATH)
        for name in ("ref_ptpr_solve", "ref_tpr_solve"):
            getattr(lib, name).argtypes = [_f32p, _f32p, _f32p, _f32p, _i, _i]
            getattr(lib, name).restype = _i
        for name in ("ref_pcr_solve", "ref_cr_solve"):
            getattr(lib, name).argtypes = [_f32p, _f32p, _f32p, _f32p, _i]
            getattr(lib, name).restype = _i
        for name in ("ref_ptpr_batch", "ref_tpr_batch"):
            getattr(lib, name).argtypes = [_f32p, _vp, _f32p, _f32p, _f32p, _i, _i, _i, _i]
            getattr(lib, name).restype = _i
        lib.ref_ptpr_normalize.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i, _i]
        lib.ref_ptpr_normalize.restype = None
        _ref = lib
    return _ref


def ref_available():
    return _load_ref() is not None


def _c32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def _vptr(x):
    return None if x is None else x.ctypes.data_as(_vp)


# --------------------------------------------------------------------------
# Solvers (unit diagonal, inputs untouched, answer returned)

def ptpr(a, c, rhs, s):
    a, c, rhs = _c32(a), _c32(c), _c32(rhs)
    x = np.empty_like(a)
    if _load_oracle().orc_ptpr_solve(a, c, rhs, x, a.size, s) < 0:
        raise MemoryError("orc_ptpr_solve")
    return x


def tpr(a, c, rhs, s):
    a, c, rhs = _c32(a), _c32(c), _c32(rhs)
    x = np.empty_like(a)
    if _load_oracle().orc_tpr_solve(a, c, rhs, x, a.size, s) < 0:
        raise MemoryError("orc_tpr_solve")
    return x


def pcr(a, c, rhs):
    a, c, rhs = _c32(a), _c32(c), _c32(rhs)
    x = np.empty_like(a)
    _load_oracle().orc_pcr_solve(a, c, rhs, x, a.size)
    return x


def cr(a, c, rhs):
    a, c, rhs = _c32(a), _c32(c), _c32(rhs)
    x = np.empty_like(a)
    _load_oracle().orc_cr_solve(a, c, rhs, x, a.size)
    return x


def normalize(a, b, c, d):
    """The 4-argument setter's in-place division by the diagonal, on copies."""
    a, b, c, d = (np.array(v, dtype=np.float32, copy=True) for v in (a, b, c, d))
    _load_oracle().orc_normalize(a, b, c, d, a.size)
    return a, c, d


def solve_batch(method, a, b, c, d, n, s, mode=1):
    """method 'ptpr' | 'tpr'; arrays hold nsys back-to-back systems; b may be None."""
    a, c, d = _c32(a), _c32(c), _c32(d)
    b = None if b is None else _c32(b)
    x = np.empty_like(a)
    m = {"ptpr": 0, "tpr": 1}[method]
    if _load_oracle().orc_solve_batch(m, a, _vptr(b), c, d, x, n, s, a.size // n, mode) != 0:
        raise MemoryError("orc_solve_batch")
    return x


def max_threads():
    return _load_oracle().orc_max_threads()


def ptpr_flops(n, s):
    return _load_oracle().orc_ptpr_flops(n, s)


def tpr_flops(n, s):
    return _load_oracle().orc_tpr_flops(n, s)


# --------------------------------------------------------------------------
# Reference build (only where oracle/_ref was built)

def _need_ref():
    lib = _load_ref()
    if lib is None:
        raise RuntimeError("oracle/_ref/libtprref.so is not built")
    return lib


def ref_ptpr(a, c, rhs, s, with_flops=False):
    """The reference PTPR class; with_flops: also what its solve() returned
    (the analytic flop count, src/ptpr.cpp:103-106)."""
    a, c, rhs = _c32(a), _c32(c), _c32(rhs)
    x = np.empty_like(a)
    f = _need_ref().ref_ptpr_solve(a, c, rhs, x, a.size, s)
    return (x, f) if with_flops else x


def ref_tpr(a, c, rhs, s, with_flops=False):
    """The reference TPR class; with_flops as ref_ptpr (src/tpr.cpp:95-98)."""
    a, c, rhs = _c32(a), _c32(c), _c32(rhs)
    x = np.empty_like(a)
    f = _need_ref().ref_tpr_solve(a, c, rhs, x, a.size, s)
    return (x, f) if with_flops else x


def ref_pcr(a, c, rhs):
    a, c, rhs = _c32(a), _c32(c), _c32(rhs)
    x = np.empty_like(a)
    _need_ref().ref_pcr_solve(a, c, rhs, x, a.size)
    return x


def ref_cr(a, c, rhs):
    a, c, rhs = _c32(a), _c32(c), _c32(rhs)
    x = np.empty_like(a)
    _need_ref().ref_cr_solve(a, c, rhs, x, a.size)
    return x


def ref_normalize(a, b, c, d, s):
    a, b, c, d = _c32(a), _c32(b), _c32(c), _c32(d)
    na, nc, nd = np.empty_like(a), np.empty_like(a), np.empty_like(a)
    _need_ref().ref_ptpr_normalize(a, b, c, d, na, nc, nd, a.size, s)
    return na, nc, nd


def ref_solve_batch(method, a, b, c, d, n, s, mode=1):
    a, c, d = _c32(a), _c32(c), _c32(d)
    b = None if b is None else _c32(b)
    x = np.empty_like(a)
    fn = {"ptpr": _need_ref().ref_ptpr_batch, "tpr": _need_ref().ref_tpr_batch}[method]
    fn(a, _vptr(b), c, d, x, n, s, a.size // n, mode)
    return x


# --------------------------------------------------------------------------
# Generator, accuracy metrics

def gen_system(seed, sys, n, kind=GEN_DOMINANT):
    a, b, c, d = (np.empty(n, np.float32) for _ in range(4))
    xt = np.zeros(n, np.float32)
    _load_oracle().orc_gen_system(seed, sys, n, kind, a, b, c, d, _vptr(xt))
    return a, b, c, d, xt


def gen_batch(seed, sys0, nsys, n, kind=GEN_DOMINANT):
    tot = nsys * n
    a, b, c, d = (np.empty(tot, np.float32) for _ in range(4))
    xt = np.zeros(tot, np.float32)
    _load_oracle().orc_gen_batch(seed, sys0, nsys, n, kind, a, b, c, d, _vptr(xt))
    return a, b, c, d, xt


def gen_rhs(seed, sys, j, n):
    d = np.empty(n, np.float32)
    _load_oracle().orc_gen_rhs(seed, sys, j, n, d)
    return d


def norm(solution, xtruth):
    """rel-L2 as src/numerical_stability.cpp:100-115."""
    s, t = _c32(solution), _c32(xtruth)
    return _load_oracle().orc_norm(s, t, s.size)


def thomas_f64(a, b, c, d):
    a, b, c, d = _c32(a), _c32(b), _c32(c), _c32(d)
    x = np.empty(a.size, np.float64)
    _load_oracle().orc_thomas_f64(a, b, c, d, x, a.size)
    return x


def thomas_f64_batch(a, b, c, d, n):
    a, b, c, d = _c32(a), _c32(b), _c32(c), _c32(d)
    x = np.empty(a.size, np.float64)
    _load_oracle().orc_thomas_f64_batch(a, b, c, d, x, n, a.size // n)
    return x


def residual_f64(a, b, c, d, x):
    a, b, c, d = _c32(a), _c32(b), _c32(c), _c32(d)
    x = np.ascontiguousarray(x, dtype=np.float64)
    r = np.empty(a.size, np.float64)
    _load_oracle().orc_residual_f64(a, b, c, d, x, r, a.size)
    return r


def rel_l2(x, ref):
    """||x - ref||_2 / ||ref||_2 in double."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))
