"""Build recipe for the native libraries (hipcc for gfx950, g++ for host C++).

    python -m trim_amd._build          # build everything in-tree

Outputs (git-ignored, shipped with the tree):
    trim_amd/libtn_hip.so      HIP kernels + the C ABI of include/tn_hip.h
    trim_amd/libuni10_amd.so   host C++ Uni10 API (include/uni10_amd/*.h) and
                               its C facade (include/uni10_capi.h), linked
                               against libtn_hip.so
    trim_amd/libuni10_seam_amd.so  the reference's backend seam functions
                               (include/uni10_seam_amd.h) over libtn_hip.so
    trim_amd/bin/itebd_dropin  an iTEBD program written against the C++ API
    trim_amd/bin/seam_check    a driver of the backend seam functions
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
SEAM_LIB = os.path.join(PKG, "libuni10_seam_amd.so")
BIN = os.path.join(PKG, "bin")
EXAMPLES = {"itebd_dropin": "-luni10_amd", "seam_check": "-luni10_seam_amd"}


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


def _link_so(target, objs, libs):
    tmp = target + ".tmp%d" % os.getpid()
    _run([CXX, "-shared", "-fPIC", "-o", tmp] + objs + ["-L" + PKG] + libs + ["-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, target)


def _build_example(name):
    src = os.path.join(PKG, "examples", name + ".cpp")
    out = os.path.join(BIN, name)
    tmp = out + ".tmp%d" % os.getpid()
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-I" + INC, src, "-o", tmp, "-L" + PKG, EXAMPLES[name], "-ltn_hip",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, out)
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
    _link_so(U10_LIB, cxx_objs, ["-ltn_hip", "-ldl"])
    seam_objs = []
    for src in sorted(glob.glob(os.path.join(CSRC, "seam", "*.cpp"))):
        out = os.path.join(OBJ, "seam_" + os.path.basename(src) + ".o")
        _run([CXX, "-O2", "-fPIC", "-std=c++17", "-Wall", "-I" + INC, "-c", src, "-o", out])
        seam_objs.append(out)
    _link_so(SEAM_LIB, seam_objs, ["-ltn_hip"])
    os.makedirs(BIN, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        list(ex.map(_build_example, sorted(EXAMPLES)))
    if verbose:
        print("built", TN_LIB, U10_LIB, SEAM_LIB, "and", ", ".join(os.path.join(BIN, n) for n in sorted(EXAMPLES)))


if __name__ == "__main__":
    build()
    sys.exit(0)
