"""Build recipe for the native libraries (hipcc for gfx950, g++ for host C++).

    python -m trim_amd._build          # build everything in-tree

Outputs (git-ignored, shipped with the tree):
    trim_amd/libtn_hip.so      HIP kernels + the C ABI of include/tn_hip.h
    trim_amd/libuni10_amd.so   host C++ Uni10 API (include/uni10_amd/*.h) and
                               its C facade (include/uni10_capi.h), linked
                               against libtn_hip.so
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-I" + INC, "-I" + CSRC,
    # packed FP32 VALU ops stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-I" + INC, "-I" + os.path.join(CSRC, "uni10"),
             "-Wall", "-Wno-sign-compare"]

TN_LIB = os.path.join(PKG, "libtn_hip.so")
U10_LIB = os.path.join(PKG, "libuni10_amd.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile_hip(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", out])
    return out


def _compile_cxx(src):
    out = os.path.join(OBJ, "u10_" + os.path.basename(src) + ".o")
    _run([CXX] + CXX_FLAGS + ["-c", src, "-o", out])
    return out


def build(verbose: bool = True) -> None:
    os.makedirs(OBJ, exist_ok=True)
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    cxx_srcs = sorted(glob.glob(os.path.join(CSRC, "uni10", "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        hip_objs = list(ex.map(_compile_hip, hip_srcs))
        cxx_objs = list(ex.map(_compile_cxx, cxx_srcs))
    tmp = TN_LIB + ".tmp%d" % os.getpid()
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + hip_objs)
    os.replace(tmp, TN_LIB)
    if cxx_objs:
        tmp = U10_LIB + ".tmp%d" % os.getpid()
        _run([CXX, "-shared", "-fPIC", "-o", tmp] + cxx_objs +
             ["-L" + PKG, "-ltn_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, U10_LIB)
    if verbose:
        print("built", TN_LIB, "and", U10_LIB if cxx_objs else "(no host C++ sources)")


if __name__ == "__main__":
    build()
    sys.exit(0)
