"""Build libfdmamr.so (HIP for gfx950 + host C++) in-tree.

The library is the product: hand-written gfx950 kernels plus the host-side quadtree and
table builder behind the C-ABI declared in include/fdmamr.h. hipcc cross-compiles it
without a GPU present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libfdmamr.so")

SOURCES = ["fdmamr_kernels.hip", "fdmamr_fast.hip", "fdmamr_selector.hip", "fdmamr_api.cpp",
           "host_tree.cpp", "host_tables.cpp"]
# FP32 fast-path and selector kernels may contract multiply-adds into FMAs
PER_FILE_FLAGS = {"fdmamr_fast.hip": ["-ffp-contract=fast"],
                  "fdmamr_selector.hip": ["-ffp-contract=fast"]}

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every multiply and add rounds on its own (FP64 parity with the reference's arithmetic);
    # FP32 kernels opt back into contraction locally
    "-ffp-contract=off",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, *PER_FILE_FLAGS.get(src, []), "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
