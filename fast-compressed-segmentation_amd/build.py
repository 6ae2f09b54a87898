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
    for src, kind in (("fcs_rans.hip", "hip"), ("fcs_ops.hip", "hip"), ("fcs_march.hip", "hip"),
                      ("fcs_mlp.hip", "hip"), ("fcs_scratch.hip", "hip"), ("fcs_host.cpp", "host"), ("fcs_serve.cpp", "host"),
                      ("fcs_producer.cpp", "host")):
        obj = os.path.join(objdir, src + ".o")
        if kind == "hip":
            _run(["hipcc", *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        else:
            _run(["hipcc", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", "--offload-arch=gfx950", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    build_node_addon()
    return LIB


def build_node_addon():
    """N-API addon (node/fcs.node) linking libfcs.so; skipped without node headers."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        print("node headers not found; N-API addon not built", flush=True)
        return None
    out = os.path.join(HERE, "node", "fcs.node")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-I", inc, "-I", os.path.join(ROOT, "include"),
          "-DNODE_GYP_MODULE_NAME=fcs", os.path.join(HERE, "node", "fcs_napi.c"), "-o", out, "-L", LIBDIR,
          "-lfcs", "-Wl,-rpath,$ORIGIN/../lib"])
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
