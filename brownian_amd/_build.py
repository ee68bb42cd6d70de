"""Build libbrownian (HIP, gfx950) and the active_brownian driver in-tree.

Outputs: brownian_amd/lib/libbrownian.so, brownian_amd/bin/active_brownian.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
OBJ = os.path.join(PKG, "lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: no FMA contraction, so the arithmetic is the written
# operation sequence (the reference's order); packed FP32 ops disabled.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fvisibility=hidden", "-Wall"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    objs = []
    for src, kind in [("kernels.hip", "hip"), ("mlp.hip", "hip"), ("api.cpp", "hip"), ("host_init.cpp", "host")]:
        out = os.path.join(OBJ, src + ".o")
        if kind == "hip":
            _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(SRC, src), "-o", out], verbose)
        else:
            _run(["g++", *HOST_FLAGS, "-c", os.path.join(SRC, src), "-o", out], verbose)
        objs.append(out)
    so = os.path.join(LIB, "libbrownian.so")
    tmp = so + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, so)
    drv = os.path.join(BIN, "active_brownian")
    _run(["g++", *HOST_FLAGS, "-fvisibility=default", os.path.join(SRC, "driver", "active_brownian.cpp"),
          "-I", os.path.join(os.path.dirname(PKG), "include"), "-L", LIB, "-lbrownian",
          "-Wl,-rpath,$ORIGIN/../lib", "-o", drv], verbose)
    return so


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
