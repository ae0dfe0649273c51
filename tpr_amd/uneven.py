"""Batches of systems of different sizes: one persistent launch per solve.

A plan fixes the batch's layout (system k of n[k] equations at
``offset(k)``, back to back) and the per-size-class slice schedule; solving
is then two kernels on the caller's stream (a counter reset and the
persistent solver), with no allocation, so it can be captured in a graph.

schedule: None (s = min(n, 256) for every class) or a mapping
``n -> (s, E, L)``; ``tpr_amd.autotune`` writes such tables.
"""
import ctypes

import numpy as np

from . import _lib
from .batched import ARITH, GEN, METHODS, _dptr, _stream

N_LOG2 = 32  # schedule rows, indexed by log2(n)


def schedule_array(schedule):
    """The C ABI's 32 x (s, E, L) int table for a mapping n -> (s, E, L)."""
    tab = np.zeros((N_LOG2, 3), np.int32)
    for n, sel in (schedule or {}).items():
        n = int(n)
        if n <= 0 or n & (n - 1):
            raise ValueError("schedule keys must be powers of two, got %r" % n)
        s, E, L = (tuple(sel) + (0, 0))[:3]
        tab[n.bit_length() - 1] = (s, E, L)
    return tab


class UnevenPlan:
    """tprh_uneven_plan: layout and schedule of one uneven batch."""

    def __init__(self, sizes, method="ptpr", arith="fast", schedule=None):
        self.sizes = np.ascontiguousarray(np.asarray(sizes, dtype=np.int32))
        self.method, self.arith = method, arith
        self._tab = schedule_array(schedule)
        h = ctypes.c_void_p()
        _lib.call("tprh_uneven_plan_create", ctypes.byref(h), METHODS[method], ARITH[arith],
                  self.sizes.ctypes.data_as(ctypes.c_void_p), int(self.sizes.size),
                  self._tab.ctypes.data_as(ctypes.c_void_p) if schedule else None)
        self._h = h
        self.nsys = int(self.sizes.size)
        self.equations = int(_lib.lib().tprh_uneven_equations(h))
        L = _lib.lib()
        self.offsets = np.array([L.tprh_uneven_offset(h, k) for k in range(self.nsys + 1)], np.int64)

    def offset(self, k):
        return int(self.offsets[k])

    def solve(self, a, b, c, d, x, stream=None):
        """x = A^-1 d for every system; b may be None (unit diagonal)."""
        t = self.equations
        _lib.call("tprh_uneven_solve", self._h, _dptr(a, "a", t), _dptr(b, "b", t), _dptr(c, "c", t),
                  _dptr(d, "d", t), _dptr(x, "x", t), _stream(stream))
        return x

    def generate(self, kind, seed, sys0, a, b, c, d, xt=None, stream=None):
        """Seeded synthetic systems sys0 + k in the plan's layout."""
        t = self.equations
        _lib.call("tprh_uneven_generate", self._h, GEN[kind], int(seed), int(sys0), _dptr(a, "a", t),
                  _dptr(b, "b", t), _dptr(c, "c", t), _dptr(d, "d", t), _dptr(xt, "xt", t), _stream(stream))

    def close(self):
        if self._h is not None and self._h.value:
            _lib.lib().tprh_uneven_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def log_uniform_sizes(nsys, lo_log2, hi_log2, seed):
    """n_k = 2^U{lo..hi}, drawn from a seeded numpy generator (host side)."""
    rng = np.random.default_rng(seed)
    return (1 << rng.integers(lo_log2, hi_log2 + 1, size=nsys)).astype(np.int32)
