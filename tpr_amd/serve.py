"""The serving loop (tprh_server_*): continuous batching of solve requests
against device-resident matrices under an HBM budget.

    srv = Server(budget_bytes=64 << 30)
    m = srv.matrix_generate(4096, seed=1, sys=17)     # or matrix_upload(a, b, c)
    r = srv.submit(m, nrhs=8)                          # right-hand sides generated on the device
    srv.wait(r); x = srv.result(r)                     # torch tensor (copy) of 8 x 4096 answers
    srv.release(r); srv.matrix_release(m)

Answers either stay in the server's arena until ``release`` (x=None at
submit) or are copied into a caller tensor when the request completes.
Timestamps (submit, launch, done) are on the server's steady clock.
"""
import ctypes

import numpy as np

from . import _lib
from .batched import ARITH, GEN, METHODS
from .uneven import schedule_array

STATES = {0: "queued", 1: "on device", 2: "done", 3: "failed"}


def _ptr(t, name, numel):
    if t is None:
        return None
    import torch
    if isinstance(t, np.ndarray):
        if t.dtype != np.float32 or not t.flags.c_contiguous or t.size < numel:
            raise TypeError("%s must be a contiguous float32 array of >= %d elements" % (name, numel))
        return t.ctypes.data_as(ctypes.c_void_p)
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_contiguous() or t.numel() < numel:
        raise TypeError("%s must be a contiguous float32 tensor of >= %d elements" % (name, numel))
    return ctypes.c_void_p(t.data_ptr())


class Server:
    """tprh_server: a request loop over one arena of ``budget_bytes``."""

    def __init__(self, budget_bytes, method="ptpr", arith="fast", max_inflight=4, max_batch_systems=16384,
                 schedule=None, min_batch_systems=0, max_wait_us=0.0, adaptive_window=False):
        cfg = _lib.ServerConfig()
        cfg.method, cfg.arith = METHODS[method], ARITH[arith]
        cfg.hbm_budget = int(budget_bytes)
        cfg.max_inflight, cfg.max_batch_systems = int(max_inflight), int(max_batch_systems)
        cfg.min_batch_systems, cfg.max_wait_us = int(min_batch_systems), float(max_wait_us)
        cfg.adaptive_window = 1 if adaptive_window else 0
        self._tab = schedule_array(schedule) if schedule else None
        cfg.schedule = self._tab.ctypes.data if self._tab is not None else None
        h = ctypes.c_void_p()
        _lib.call("tprh_server_create", ctypes.byref(h), ctypes.byref(cfg))
        self._h = h
        self._n = {}
        self._keep = {}  # host arrays that must outlive their pending copies

    # -- matrices ----------------------------------------------------------
    def matrix_generate(self, n, kind="dominant", seed=0, sys=0):
        mid = ctypes.c_longlong()
        _lib.call("tprh_server_matrix_generate", self._h, int(n), GEN[kind], int(seed), int(sys), ctypes.byref(mid))
        self._n[mid.value] = int(n)
        return mid.value

    def matrix_upload(self, a, b, c):
        n = int(a.numel() if hasattr(a, "numel") else a.size)
        mid = ctypes.c_longlong()
        _lib.call("tprh_server_matrix_upload", self._h, n, _ptr(a, "a", n), _ptr(b, "b", n), _ptr(c, "c", n),
                  ctypes.byref(mid))
        self._n[mid.value] = n
        return mid.value

    def matrix_release(self, mid):
        _lib.call("tprh_server_matrix_release", self._h, int(mid))

    # -- requests ----------------------------------------------------------
    def submit(self, mid, nrhs, d=None, x=None, rhs_seed=0, j0=0):
        """Queue nrhs right-hand sides against matrix mid.  d: None (generated
        on the device from (rhs_seed, the matrix's generator id, j0 + j)) or
        nrhs * n floats; x: None (kept in the arena) or a tensor to receive
        the answers.  Returns the request id."""
        n = self._n.get(mid, 0)  # unknown ids are refused by the library
        rid = ctypes.c_longlong()
        _lib.call("tprh_server_submit", self._h, int(mid), int(nrhs), _ptr(d, "d", nrhs * n), int(rhs_seed), int(j0),
                  _ptr(x, "x", nrhs * n), ctypes.byref(rid))
        if d is not None or x is not None:
            self._keep[rid.value] = (d, x)
        return rid.value

    def wait(self, rid, timeout=60.0):
        _lib.call("tprh_server_wait", self._h, int(rid), float(timeout))

    def info(self, rid):
        inf = _lib.RequestInfo()
        _lib.call("tprh_server_request_info", self._h, int(rid), ctypes.byref(inf))
        return {"state": STATES[inf.state], "n": inf.n, "nrhs": inf.nrhs, "t_submit": inf.t_submit,
                "t_launch": inf.t_launch, "t_done": inf.t_done, "x_ptr": inf.x or None}

    def result(self, rid):
        """A torch copy (nrhs x n) of a completed request's retained answers."""
        import torch
        inf = self.info(rid)
        if inf["x_ptr"] is None:
            raise RuntimeError("request %d holds no answers in the arena" % rid)
        out = torch.empty(inf["nrhs"] * inf["n"], dtype=torch.float32, device="cuda")
        _lib.call("tprh_server_copy_result", self._h, int(rid), ctypes.c_void_p(out.data_ptr()))
        return out.view(inf["nrhs"], inf["n"])

    def poll(self, max_ids=4096, timeout=None):
        """Ids of requests completed since the last poll, oldest first
        (requests already released are left out); with a timeout (s), sleeps
        until at least one has completed or the timeout passes."""
        buf = (ctypes.c_longlong * max_ids)()
        cnt = ctypes.c_int()
        if timeout is None:
            _lib.call("tprh_server_poll", self._h, buf, int(max_ids), ctypes.byref(cnt))
        else:
            _lib.call("tprh_server_wait_any", self._h, buf, int(max_ids), ctypes.byref(cnt), float(timeout))
        return list(buf[:cnt.value])

    def release(self, rid):
        _lib.call("tprh_server_request_release", self._h, int(rid))
        self._keep.pop(rid, None)

    def now(self):
        return _lib.lib().tprh_server_now(self._h)

    def stats(self):
        st = _lib.ServerStats()
        _lib.call("tprh_server_stats_get", self._h, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in _lib.ServerStats._fields_}

    def close(self):
        if self._h is not None and self._h.value:
            _lib.lib().tprh_server_destroy(self._h)
        self._h = None
        self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
