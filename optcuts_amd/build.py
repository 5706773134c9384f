"""Compiles the HIP library for gfx950 in-tree: optcuts_amd/_lib/liboptcuts_gpu.so.

    python -m optcuts_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "liboptcuts_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", f)
           for f in ("oc_kernels.hip", "oc_amg.hip", "oc_gnn.hip", "oc_ic0.hip", "oc_capi.hip", "oc_io.cpp")]
HEADERS = [os.path.join(HERE, "csrc", f) for f in ("oc_common.hpp", "oc_kernels.hpp", "oc_sd_element.hpp",
                                                   "oc_amg.hpp", "oc_gnn.hpp", "oc_ic0.hpp", "oc_precond.hpp",
                                                   "oc_io.hpp")] + [
    os.path.join(ROOT, "include", "optcuts_gpu.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off (they are formed only by the compiler)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
