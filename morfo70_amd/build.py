"""Builds libm70h.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

Usage: python -m morfo70_amd.build   (or morfo70_amd.build.build() from Python)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libm70h.so")

SOURCES = ["hydro_kernels.hip", "line_solvers.hip", "partition_solver.hip", "adjoint.hip", "adjoint_batch.hip", "sediment.hip",
           "wave.hip", "outer.hip", "netcdf3.cpp", "m70h_api.hip"]

# -ffp-contract=off: no fused multiply-adds, so every product and sum rounds as the
#   reference's x86-64 build does.
# -packed-fp32-ops off: the image's packed FP32 forms are avoided (FP64 code here anyway).
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


FACADE = os.path.join(HERE, "host", "facade_run")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # C++ facade (the reference's Hydro call surface) and its driver, plain g++
    host = os.path.join(HERE, "host")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", FACADE, os.path.join(host, "facade_run.cpp"),
           os.path.join(host, "Hydro.cpp"), os.path.join(host, "Sediment.cpp"), "-L" + HERE, "-lm70h", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
