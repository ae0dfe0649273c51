"""The reference-side binding of INTEGRATION.md, compiled against the
reference's own declarations (src/ptpr.cuh:50-57, src/tpr.cuh:50-57) and
linked with libtpr_hip.so, in a tree where the reference is present (the
build container; skipped elsewhere, nothing of the reference is needed at
run time).  A signature mismatch between the shim and the reference's
PTPR_CU / TPR_CU declarations fails to compile or to link here.

tests/cpp/shim/ holds the two shim files (the text INTEGRATION.md shows,
checked below) and a caller written against the reference headers alone."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
SHIM = os.path.join(ROOT, "tests", "cpp", "shim")
LIB = os.path.join(ROOT, "tpr_amd", "libtpr_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "ptpr.cuh")),
                               reason="the reference tree is not present here")


def test_integration_shows_the_compiled_shim():
    """INTEGRATION.md's code block is the file this test compiles."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", doc, flags=re.S)
    shim = open(os.path.join(SHIM, "ptpr_hip_shim.cpp")).read()
    assert shim in blocks


@needs_ref
@pytest.mark.skipif(not os.path.exists(LIB), reason="libtpr_hip.so is not built")
def test_shim_compiles_and_links_against_reference_headers(tmp_path):
    cxx = shutil.which("g++") or "g++"
    objs = []
    for src in ("ptpr_hip_shim.cpp", "tpr_hip_shim.cpp", "shim_caller.cpp"):
        obj = str(tmp_path / (src + ".o"))
        p = subprocess.run([cxx, "-std=c++17", "-O2", "-Wall", "-Werror", "-c", os.path.join(SHIM, src),
                            "-I", REF_SRC, "-I", os.path.join(ROOT, "include"), "-o", obj],
                           capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr[-3000:]
        objs.append(obj)
    exe = str(tmp_path / "shim_caller")
    p = subprocess.run([cxx] + objs + ["-o", exe, "-L", os.path.dirname(LIB), "-ltpr_hip",
                                       "-L", os.path.join(ROCM, "lib"), "-Wl,-rpath-link," + os.path.join(ROCM, "lib"),
                                       "-Wl,-rpath," + os.path.dirname(LIB)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-3000:]
    # the shim defines exactly the reference's four entry points
    nm = subprocess.run(["nm", "-C", "--defined-only", objs[0], objs[1]], capture_output=True, text=True).stdout
    for sym in ("PTPR_CU::ptpr_cu(float*, float*, float*, float*, int, int)",
                "PTPR_CU::pcr_cu(float*, float*, float*, float*, int)",
                "TPR_CU::tpr_cu(float*, float*, float*, float*, int, int)",
                "TPR_CU::cr_cu(float*, float*, float*, float*, int)"):
        assert sym in nm, sym
