"""Host-side checks of the serving C ABI: configuration errors are reported
before any device call, and null handles are refused."""
import ctypes

import pytest

from tpr_amd import _lib
from tpr_amd.uneven import schedule_array


def _create(**kw):
    cfg = _lib.ServerConfig()
    cfg.method = kw.get("method", _lib.PTPR_M)
    cfg.arith = kw.get("arith", _lib.ARITH_FAST)
    cfg.hbm_budget = kw.get("budget", 1 << 20)
    cfg.max_inflight = kw.get("inflight", 2)
    cfg.max_batch_systems = kw.get("batch", 64)
    cfg.min_batch_systems = kw.get("min_batch", 0)
    cfg.max_wait_us = kw.get("max_wait", 0.0)
    tab = kw.get("schedule")
    cfg.schedule = tab.ctypes.data if tab is not None else None
    h = ctypes.c_void_p()
    st = _lib.lib().tprh_server_create(ctypes.byref(h), ctypes.byref(cfg))
    return st, h


@pytest.mark.parametrize("kw", [dict(method=_lib.PCR_M), dict(arith=9), dict(inflight=0), dict(inflight=17),
                                dict(batch=0), dict(budget=16), dict(min_batch=-1), dict(max_wait=-1.0),
                                dict(max_wait=float("nan"))])
def test_bad_config_refused_before_device(kw):
    st, h = _create(**kw)
    assert st == _lib.ERROR_INVALID_VALUE
    assert h.value is None


def test_bad_schedule_refused():
    st, h = _create(schedule=schedule_array({4096: (512, 16, 32)}))   # shape outside the kernel's menu
    assert st == _lib.ERROR_NOT_SUPPORTED
    st, h = _create(schedule=schedule_array({4096: (256, 8, 16)}))    # E * L != s
    assert st == _lib.ERROR_INVALID_VALUE


def test_null_handles_refused():
    L = _lib.lib()
    rid = ctypes.c_longlong()
    assert L.tprh_server_submit(None, 1, 1, None, 0, 0, None, ctypes.byref(rid)) == _lib.ERROR_INVALID_VALUE
    assert L.tprh_server_wait(None, 1, 0.0) == _lib.ERROR_INVALID_VALUE
    assert L.tprh_server_matrix_release(None, 1) == _lib.ERROR_INVALID_VALUE
    cnt = ctypes.c_int()
    assert L.tprh_server_poll(None, None, 0, ctypes.byref(cnt)) == _lib.ERROR_INVALID_VALUE
    buf = (ctypes.c_longlong * 4)()
    assert L.tprh_server_wait_any(None, buf, 4, ctypes.byref(cnt), 0.0) == _lib.ERROR_INVALID_VALUE
    assert L.tprh_server_destroy(None) == _lib.SUCCESS
    assert L.tprh_status_string(_lib.ERROR_TIMEOUT) == b"timed out"
    assert L.tprh_status_string(_lib.ERROR_BUDGET) == b"memory budget exceeded"
