"""Python mirror of the reference's solver interface, backed by libtpr_hip.

Reference surface (src/ptpr.hpp:53-97, src/tpr.hpp:53-97, src/pcr.hpp,
src/cr.hpp) and the CUDA free functions (src/ptpr.cuh:50-57,
src/tpr.cuh:50-57)::

    t = PTPR(n, s)                              # PTPR t(n, s);
    t.set_tridiagonal_system(a, c, rhs)         # unit diagonal, 3-argument form
    t.set_tridiagonal_system(a, diag, c, rhs)   # 4-argument form (divides by diag)
    flops = t.solve()                           # int flops = t.solve();
    t.get_ans(x)                                # t.get_ans(x);

    ptpr_cu(a, c, rhs, x, n, s)                 # PTPR_CU::ptpr_cu
    tpr_cu(a, c, rhs, x, n, s)                  # TPR_CU::tpr_cu
    pcr_cu(a, c, rhs, x, n)                     # PTPR_CU::pcr_cu
    cr_cu(a, c, rhs, x, n)                      # TPR_CU::cr_cu

Arrays are host numpy float32 arrays here.  The solve runs on the GPU;
there is no CPU path.  As in the reference, the solver keeps references to
the arrays given to set_tridiagonal_system and reads them in solve().  It
does not modify them: the division by the diagonal happens on the GPU
(the reference divides the caller's arrays in place, src/ptpr.hpp:80-90).

Errors raise ``TprError`` instead of exiting the process.
"""
import ctypes

import numpy as np

from . import _lib


def _host_f32(x, name):
    if not isinstance(x, np.ndarray) or x.dtype != np.float32 or not x.flags.c_contiguous:
        raise TypeError("%s must be a C-contiguous numpy float32 array" % name)
    return x


def _ptr(x):
    return x.ctypes.data_as(ctypes.c_void_p)


class _SolverBase:
    _method = None

    def __init__(self, n, s=0):
        h = ctypes.c_void_p()
        _lib.call("tprh_solver_create", ctypes.byref(h), self._method, int(n), int(s))
        self._h = h
        self.n = int(n)
        self.s = int(s)
        self._keep = ()

    def set_arith(self, arith):
        """'exact' (default: bit-identical to the CPU reference) or 'fast'."""
        code = {"exact": _lib.ARITH_EXACT, "fast": _lib.ARITH_FAST}[arith]
        _lib.call("tprh_solver_set_arith", self._h, code)

    def set_tridiagonal_system(self, *arrays):
        if len(arrays) == 3:
            a, c, rhs = arrays
            diag = None
        elif len(arrays) == 4:
            a, diag, c, rhs = arrays
        else:
            raise TypeError("set_tridiagonal_system(a, c, rhs) or (a, diag, c, rhs)")
        for v, nm in ((a, "a"), (c, "c"), (rhs, "rhs")) + (((diag, "diag"),) if diag is not None else ()):
            _host_f32(v, nm)
            if v.size != self.n:
                raise ValueError("%s has %d entries, expected n = %d" % (nm, v.size, self.n))
        self._keep = (a, diag, c, rhs)
        _lib.call("tprh_solver_set_system", self._h, _ptr(a), None if diag is None else _ptr(diag), _ptr(c),
                  _ptr(rhs))

    def solve(self):
        f = ctypes.c_int(0)
        _lib.call("tprh_solver_solve", self._h, ctypes.byref(f))
        return f.value

    def get_ans(self, x):
        _host_f32(x, "x")
        if x.size < self.n:
            raise ValueError("x has %d entries, expected n = %d" % (x.size, self.n))
        _lib.call("tprh_solver_get_ans", self._h, _ptr(x))
        return 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.lib().tprh_solver_destroy(h)
            self._h = None


class PTPR(_SolverBase):
    """PCR-like tree partitioning reduction, PTPR(n, s) (src/ptpr.hpp:44-110)."""
    _method = _lib.PTPR_M


class TPR(_SolverBase):
    """Tree partitioning reduction, TPR(n, s) (src/tpr.hpp:44-112)."""
    _method = _lib.TPR_M


class PCR(_SolverBase):
    """Parallel cyclic reduction, PCR(n) (src/pcr.hpp:9-64), n <= 4096."""
    _method = _lib.PCR_M

    def __init__(self, n):
        super().__init__(n, 0)


class CR(_SolverBase):
    """Cyclic reduction, CR(n) (src/cr.hpp:24-81), n <= 4096."""
    _method = _lib.CR_M

    def __init__(self, n):
        super().__init__(n, 0)


def _cu(name, a, c, rhs, x, *dims):
    for v, nm in ((a, "a"), (c, "c"), (rhs, "rhs"), (x, "x")):
        _host_f32(v, nm)
        if v.size < dims[0]:
            raise ValueError("%s shorter than n" % nm)
    _lib.call(name, _ptr(a), _ptr(c), _ptr(rhs), _ptr(x), *[int(v) for v in dims])


def ptpr_cu(a, c, rhs, x, n, s):
    """PTPR_CU::ptpr_cu(a, c, rhs, x, n, s): unit diagonal, answer into x."""
    _cu("tprh_ptpr_cu", a, c, rhs, x, n, s)


def tpr_cu(a, c, rhs, x, n, s):
    """TPR_CU::tpr_cu(a, c, rhs, x, n, s)."""
    _cu("tprh_tpr_cu", a, c, rhs, x, n, s)


def pcr_cu(a, c, rhs, x, n):
    """PTPR_CU::pcr_cu(a, c, rhs, x, n)."""
    _cu("tprh_pcr_cu", a, c, rhs, x, n)


def cr_cu(a, c, rhs, x, n):
    """TPR_CU::cr_cu(a, c, rhs, x, n)."""
    _cu("tprh_cr_cu", a, c, rhs, x, n)
