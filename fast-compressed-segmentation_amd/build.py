"""Build the native library libfcs.so (HIP for gfx950 + host runtime + producer).

Usage: python fast-compressed-segmentation_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libfcs.so")

HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
    # FP32 ray setup must round exactly like the host restatement
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src, kind in (("fcs_kernels.hip", "hip"), ("fcs_host.cpp", "host"), ("fcs_producer.cpp", "host")):
        obj = os.path.join(objdir, src + ".o")
        if kind == "hip":
            _run(["hipcc", *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        else:
            _run(["hipcc", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", "--offload-arch=gfx950", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
