"""ctypes binding of libtpr_hip.so (the C ABI declared in include/tpr_hip.h).

The library must have been built (``python -m tpr_amd.build``); there is no
fallback: every entry point raises if the shared object cannot be loaded.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtpr_hip.so")

SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_NOT_SUPPORTED = 2
ERROR_HIP = 3
ERROR_OUT_OF_MEMORY = 4
ERROR_BUDGET = 5
ERROR_TIMEOUT = 6

PTPR_M, TPR_M, PCR_M, CR_M = 0, 1, 2, 3
ARITH_FAST, ARITH_EXACT = 0, 1
GEN_DOMINANT, GEN_WEAK, GEN_FIXED, GEN_RANDOM_RHS = 0, 1, 2, 3

_vp = ctypes.c_void_p
_i = ctypes.c_int
_u64 = ctypes.c_uint64

# name -> (restype, argtypes); one line per exported function of tpr_hip.h
SIGNATURES = {
    "tprh_status_string": (ctypes.c_char_p, [_i]),
    "tprh_last_error": (ctypes.c_char_p, []),
    "tprh_version": (_i, []),
    "tprh_ptpr_cu": (_i, [_vp, _vp, _vp, _vp, _i, _i]),
    "tprh_tpr_cu": (_i, [_vp, _vp, _vp, _vp, _i, _i]),
    "tprh_pcr_cu": (_i, [_vp, _vp, _vp, _vp, _i]),
    "tprh_cr_cu": (_i, [_vp, _vp, _vp, _vp, _i]),
    "tprh_set_kernel_timing": (_i, [_i]),
    "tprh_last_kernel_ms": (ctypes.c_float, []),
    "tprh_solver_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i]),
    "tprh_solver_set_arith": (_i, [_vp, _i]),
    "tprh_solver_set_system": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "tprh_solver_solve": (_i, [_vp, ctypes.POINTER(_i)]),
    "tprh_solver_get_ans": (_i, [_vp, _vp]),
    "tprh_solver_destroy": (_i, [_vp]),
    "tprh_solve_batched": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "tprh_solve_batched_shape": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "tprh_shape_supported": (_i, [_i, _i]),
    "tprh_slice_shape": (_i, [_i, _i, _i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "tprh_generate": (_i, [_i, _u64, _u64, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
}



class SolveDesc(ctypes.Structure):
    """tprh_solve_desc of include/tpr_hip.h."""
    _fields_ = [
        ("method", ctypes.c_int), ("arith", ctypes.c_int),
        ("n", ctypes.c_int), ("s", ctypes.c_int), ("nsys", ctypes.c_int),
        ("coef", ctypes.c_int),
        ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
        ("d", ctypes.c_void_p), ("x", ctypes.c_void_p),
        ("refine", ctypes.c_int), ("E", ctypes.c_int), ("L", ctypes.c_int), ("chunk", ctypes.c_int),
        ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t),
        ("coef_shared", ctypes.c_int), ("split_form", ctypes.c_int), ("x_cached", ctypes.c_int),
    ]


SIGNATURES.update({
    "tprh_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(SolveDesc)]),
    "tprh_solve_ex": (_i, [ctypes.POINTER(SolveDesc), _vp]),
    "tprh_solve_ex_noalloc": (_i, [ctypes.POINTER(SolveDesc), _vp]),
    "tprh_uneven_plan_create": (_i, [ctypes.POINTER(_vp), _i, _i, _vp, _i, _vp]),
    "tprh_uneven_solve": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tprh_uneven_offset": (ctypes.c_longlong, [_vp, _i]),
    "tprh_uneven_equations": (ctypes.c_longlong, [_vp]),
    "tprh_uneven_generate": (_i, [_vp, _i, _u64, _u64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tprh_uneven_plan_destroy": (_i, [_vp]),
})



class ServerConfig(ctypes.Structure):
    """tprh_server_config of include/tpr_hip.h."""
    _fields_ = [("method", ctypes.c_int), ("arith", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("max_inflight", ctypes.c_int), ("max_batch_systems", ctypes.c_int),
                ("schedule", ctypes.c_void_p), ("min_batch_systems", ctypes.c_int),
                ("max_wait_us", ctypes.c_double), ("adaptive_window", ctypes.c_int)]


class RequestInfo(ctypes.Structure):
    """tprh_request_info of include/tpr_hip.h."""
    _fields_ = [("state", ctypes.c_int), ("n", ctypes.c_int), ("nrhs", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_launch", ctypes.c_double), ("t_done", ctypes.c_double),
                ("x", ctypes.c_void_p)]


class ServerStats(ctypes.Structure):
    """tprh_server_stats of include/tpr_hip.h."""
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("failed", ctypes.c_longlong),
                ("systems", ctypes.c_longlong), ("equations", ctypes.c_longlong), ("batches", ctypes.c_longlong),
                ("max_batch_systems_seen", ctypes.c_longlong), ("max_inflight_seen", ctypes.c_int),
                ("queued", ctypes.c_int), ("bytes_in_use", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t),
                ("budget", ctypes.c_size_t)]


_ll = ctypes.c_longlong
SIGNATURES.update({
    "tprh_server_create": (_i, [ctypes.POINTER(_vp), ctypes.POINTER(ServerConfig)]),
    "tprh_server_destroy": (_i, [_vp]),
    "tprh_server_matrix_upload": (_i, [_vp, _i, _vp, _vp, _vp, ctypes.POINTER(_ll)]),
    "tprh_server_matrix_generate": (_i, [_vp, _i, _i, _u64, _u64, ctypes.POINTER(_ll)]),
    "tprh_server_matrix_release": (_i, [_vp, _ll]),
    "tprh_server_submit": (_i, [_vp, _ll, _i, _vp, _u64, _i, _vp, ctypes.POINTER(_ll)]),
    "tprh_server_request_info": (_i, [_vp, _ll, ctypes.POINTER(RequestInfo)]),
    "tprh_server_wait": (_i, [_vp, _ll, ctypes.c_double]),
    "tprh_server_poll": (_i, [_vp, ctypes.POINTER(_ll), _i, ctypes.POINTER(_i)]),
    "tprh_server_wait_any": (_i, [_vp, ctypes.POINTER(_ll), _i, ctypes.POINTER(_i), ctypes.c_double]),
    "tprh_server_copy_result": (_i, [_vp, _ll, _vp]),
    "tprh_server_request_release": (_i, [_vp, _ll]),
    "tprh_server_now": (ctypes.c_double, [_vp]),
    "tprh_server_stats_get": (_i, [_vp, ctypes.POINTER(ServerStats)]),
})

STORE_F32, STORE_F16 = 0, 1

_lib = None


class TprError(RuntimeError):
    """A libtpr_hip call returned a status other than success."""

    def __init__(self, status, func, msg):
        self.status = status
        super().__init__("%s failed: %s (%s)" % (func, msg, _status_name(status)))


def _status_name(st):
    return {0: "success", 1: "invalid value", 2: "not supported", 3: "HIP error",
            4: "out of memory", 5: "budget exceeded", 6: "timed out"}.get(st, "status %d" % st)


def _load_torch_first():
    """PyTorch's wheel carries its own HIP runtime (soname libamdhip64.so.7,
    requested by torch as libamdhip64.so).  Loaded after ours it would be a
    second runtime in the process, which finds no GPU; loaded first, our
    library's libamdhip64.so.7 dependency resolves to it.  So torch, when
    installed, is imported before any library of this package is opened."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded library; raises RuntimeError if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("tpr_amd/libtpr_hip.so is not built: run `python -m tpr_amd.build`")
        _load_torch_first()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, func):
    if status != SUCCESS:
        msg = lib().tprh_last_error()
        raise TprError(status, func, msg.decode() if msg else "")
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)
