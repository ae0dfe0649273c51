"""Build libtpr_hip.so (gfx950) and the tpr_pm driver in-tree.

    python -m tpr_amd.build            # library + driver
    python -m tpr_amd.build --lib      # library only

Every .hip source under csrc/ is compiled to an object in parallel with
hipcc, then linked into tpr_amd/libtpr_hip.so.  Flags:
  --offload-arch=gfx950                       MI355X only
  -ffp-contract=off                           no implicit FMA: the exact
                                              arithmetic mode depends on it;
                                              the fast mode asks for FMA
                                              explicitly
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                              no packed FP32 instructions
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtpr_hip.so")
CLI_SRC = os.path.join(HERE, "cli", "tpr_pm.cpp")
CLI_BIN = os.path.join(HERE, "tpr_pm")
NSTAB_SRC = os.path.join(HERE, "cli", "tpr_nstab.cpp")
NSTAB_BIN = os.path.join(HERE, "tpr_nstab")
SERVE_SRC = os.path.join(HERE, "cli", "tpr_serve_bench.cpp")
SERVE_BIN = os.path.join(HERE, "tpr_serve_bench")
VENDOR_SRC = os.path.join(HERE, "comparators", "rocsparse_gtsv.cpp")
DIAG_SRC = [os.path.join(HERE, "diag", "stage_clock.hip"), os.path.join(HERE, "diag", "exact_div.hip")]
DIAG_LIB = os.path.join(HERE, "libtpr_diag.so")
VENDOR_LIB = os.path.join(HERE, "libtpr_rocsparse.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST_HIP = ["-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    return p


def build_lib(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(OBJ, f[:-4] + ".o") for f in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
                for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    shutil.rmtree(OBJ, ignore_errors=True)
    return LIB


def build_cli():
    if not os.path.exists(CLI_SRC):
        return None
    tmp = CLI_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17"] + HOST_HIP + ["-I" + os.path.join(ROOT, "include"), CLI_SRC, "-o", tmp,
          "-L" + HERE, "-ltpr_hip", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-ldl",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, CLI_BIN)
    return CLI_BIN


def build_vendor():
    """libtpr_rocsparse.so, the rocSPARSE comparator (host code only); skipped
    when the image has no rocSPARSE."""
    if not os.path.exists(os.path.join(ROCM, "include", "rocsparse", "rocsparse.h")):
        return None
    tmp = VENDOR_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC"] + HOST_HIP + [VENDOR_SRC, "-o", tmp,
          "-L" + os.path.join(ROCM, "lib"), "-lrocsparse", "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, VENDOR_LIB)
    return VENDOR_LIB


def build_diag():
    """libtpr_diag.so: the per-stage clock builds of the batched fast-PTPR
    solve and of the split path's pipelined form (tpr_amd/diag/stage_clock.hip,
    run by tpr_amd.stage_clock) and the exhaustive check of the exact
    kernels' reciprocal (tpr_amd/diag/exact_div.hip); a diagnostic, not part
    of the solver library."""
    tmp = DIAG_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared"] + DIAG_SRC + ["-o", tmp])
    os.replace(tmp, DIAG_LIB)
    return DIAG_LIB


def build_nstab():
    tmp = NSTAB_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17"] + HOST_HIP + ["-I" + os.path.join(ROOT, "include"), NSTAB_SRC, "-o", tmp,
          "-L" + HERE, "-ltpr_hip", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-ldl",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, NSTAB_BIN)
    return NSTAB_BIN


def build_serve_bench():
    tmp = SERVE_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17"] + HOST_HIP + ["-I" + os.path.join(ROOT, "include"), SERVE_SRC, "-o", tmp,
          "-L" + HERE, "-ltpr_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, SERVE_BIN)
    return SERVE_BIN


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args(argv)
    print("built", build_lib(args.j))
    if not args.lib:
        for b in (build_cli(), build_vendor(), build_nstab(), build_serve_bench(), build_diag()):
            if b:
                print("built", b)


if __name__ == "__main__":
    sys.exit(main())
