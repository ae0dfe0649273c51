"""The C ABI without a GPU: libtpr_hip.so loads, exports every function that
include/tpr_hip.h declares, and rejects invalid arguments before touching the
device.  (Compute calls are in the gpu-marked tests.)"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tpr_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tprh_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("tprh_ptpr_cu", "tprh_tpr_cu", "tprh_pcr_cu", "tprh_cr_cu", "tprh_solver_create",
                 "tprh_solver_set_system", "tprh_solver_solve", "tprh_solver_get_ans", "tprh_solve_batched",
                 "tprh_generate"):
        assert must in names


def test_library_exports_every_declared_function():
    from tpr_amd import _lib
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    # and the Python binding covers exactly the declared surface
    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_status_strings_and_version():
    from tpr_amd import _lib
    L = _lib.lib()
    assert L.tprh_version() >= 100
    assert L.tprh_status_string(0) == b"success"
    assert L.tprh_status_string(1) == b"invalid value"


@pytest.mark.parametrize("n,s", [(1000, 10), (1024, 3), (1024, 0), (96, 32), (8, 16), (-4, 2)])
def test_invalid_shapes_rejected_without_device(n, s):
    from tpr_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    st = L.tprh_solver_create(ctypes.byref(h), _lib.PTPR_M, n, s)
    assert st == _lib.ERROR_INVALID_VALUE
    assert h.value is None
    assert b"" != L.tprh_last_error()
    st = L.tprh_solve_batched(_lib.PTPR_M, _lib.ARITH_FAST, None, None, None, None, None, n, s, 1, None)
    assert st == _lib.ERROR_INVALID_VALUE


def test_unsupported_sizes_reported():
    from tpr_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    # s above the largest compiled slice
    assert L.tprh_solver_create(ctypes.byref(h), _lib.PTPR_M, 8192, 4096) == _lib.ERROR_NOT_SUPPORTED
    # reduced system too large for one workgroup at s = 16 (m * 16 lanes > 1024)
    assert L.tprh_solver_create(ctypes.byref(h), _lib.PTPR_M, 65536 * 4, 16) == _lib.ERROR_NOT_SUPPORTED
    # explicit shape that does not multiply to s
    assert L.tprh_solve_batched_shape(0, 0, None, None, None, None, None, 1024, 256, 1, 8, 16, None) == \
        _lib.ERROR_INVALID_VALUE
    assert L.tprh_shape_supported(16, 16) == 1
    assert L.tprh_shape_supported(3, 16) == 0


def test_solver_lifecycle_without_device():
    """create / set_system / get_ans / destroy do not touch the GPU."""
    from tpr_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.tprh_solver_create(ctypes.byref(h), _lib.TPR_M, 1024, 256) == 0
    a = np.zeros(1024, np.float32)
    p = a.ctypes.data_as(ctypes.c_void_p)
    assert L.tprh_solver_set_system(h, p, None, p, p) == 0
    assert L.tprh_solver_set_system(h, None, None, p, p) == _lib.ERROR_INVALID_VALUE
    assert L.tprh_solver_set_arith(h, 7) == _lib.ERROR_INVALID_VALUE
    assert L.tprh_solver_destroy(h) == 0


def test_python_mirror_validates_arrays():
    import tpr_amd
    t = tpr_amd.PTPR(64, 16)
    a = np.zeros(64, np.float32)
    with pytest.raises(TypeError):
        t.set_tridiagonal_system(a.astype(np.float64), a, a)
    with pytest.raises(ValueError):
        t.set_tridiagonal_system(a[:32].copy(), a, a)
    with pytest.raises(TypeError):
        t.set_tridiagonal_system(a, a)
    with pytest.raises(tpr_amd.TprError):
        tpr_amd.PTPR(100, 16)


def test_workspace_sizing_without_device():
    import ctypes
    from tpr_amd import _lib
    L = _lib.lib()
    d = _lib.SolveDesc()
    d.method, d.arith, d.n, d.s, d.nsys = 0, 0, 1024, 256, 10
    assert L.tprh_workspace_bytes(ctypes.byref(d)) == 0          # one workgroup per system
    d.n, d.nsys = 65536, 256
    ws = L.tprh_workspace_bytes(ctypes.byref(d))
    # rows, boundary x and x_ed per slice, and a 128-byte line of
    # synchronisation words per system and one more
    assert ws == 4 * 256 * 9 * 256 + 128 * (256 + 1) + 256
    d.chunk = 3
    assert L.tprh_workspace_bytes(ctypes.byref(d)) == 4 * 3 * 9 * 256 + 128 * (3 + 1) + 256
    d.s = 3
    assert L.tprh_workspace_bytes(ctypes.byref(d)) == ctypes.c_size_t(-1).value
    d.s, d.refine, d.coef = 256, 1, _lib.STORE_F16                # refinement needs fp32 coefficients
    assert L.tprh_workspace_bytes(ctypes.byref(d)) == ctypes.c_size_t(-1).value
    d.refine, d.coef, d.coef_shared = 0, 0, 2                      # coef_shared is a flag
    assert L.tprh_workspace_bytes(ctypes.byref(d)) == ctypes.c_size_t(-1).value
    d.coef_shared = 1                                             # one matrix: same workspace per system
    assert L.tprh_workspace_bytes(ctypes.byref(d)) == 4 * 3 * 9 * 256 + 128 * (3 + 1) + 256
    for form in (1, 2, 3):                                        # every form takes the same workspace
        d.split_form = form
        assert L.tprh_workspace_bytes(ctypes.byref(d)) == 4 * 3 * 9 * 256 + 128 * (3 + 1) + 256
    for form in (-1, 4):
        d.split_form = form
        assert L.tprh_workspace_bytes(ctypes.byref(d)) == ctypes.c_size_t(-1).value
    d.split_form = 0
    for flag in (-1, 2):                                          # x_cached is a flag
        d.x_cached = flag
        assert L.tprh_workspace_bytes(ctypes.byref(d)) == ctypes.c_size_t(-1).value
    d.x_cached = 1
    assert L.tprh_workspace_bytes(ctypes.byref(d)) == 4 * 3 * 9 * 256 + 128 * (3 + 1) + 256
    d.x_cached = 0
    # a large system without a workspace is refused before any launch
    d.refine, d.coef, d.coef_shared = 0, 0, 0
    d.a = d.b = d.c = d.d = d.x = 16
    assert L.tprh_solve_ex(ctypes.byref(d), None) == _lib.ERROR_INVALID_VALUE


def test_cpp_wrappers_compile_and_report_errors(tmp_path):
    """include/tpr/hip_batch.hpp and hip_solvers.hpp compile against the C ABI
    and turn a refused argument into tpr::hip::Error (no device needed)."""
    import subprocess
    src = tmp_path / "wrap.cpp"
    src.write_text('''
#include "tpr/hip_batch.hpp"
#include "tpr/hip_solvers.hpp"
int main() {
    int bad = 0;
    try { tpr::hip::UnevenPlan p({24}, TPRH_PTPR, TPRH_ARITH_FAST); }
    catch (const tpr::hip::Error& e) { bad += e.status == TPRH_ERROR_INVALID_VALUE; }
    try { tpr::hip::PTPR t(1000, 10); }
    catch (const tpr::hip::Error& e) { bad += e.status == TPRH_ERROR_INVALID_VALUE; }
    tprh_server_config cfg = {TPRH_PTPR, TPRH_ARITH_FAST, 16, 1, 1, nullptr};
    try { tpr::hip::Server s(cfg); }
    catch (const tpr::hip::Error& e) { bad += e.status == TPRH_ERROR_INVALID_VALUE; }
    return bad == 3 ? 0 : 1;
}
''')
    exe = tmp_path / "wrap"
    lib = os.path.join(ROOT, "tpr_amd")
    p = subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L" + lib, "-ltpr_hip", "-Wl,-rpath," + lib], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
