"""Build recipe of libhybird_amd.so (hipcc, gfx950, in-tree).

Each source is compiled to an object in parallel, then the objects are linked
into the shared library (every translation unit carries its own device code).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhybird_amd.so")
SOURCES = ["lb.hip", "fs.hip", "adj.hip", "dem.hip", "api.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# FP64 kernels keep the reference's operation order: no FMA contraction unless a
# kernel opts in with `#pragma clang fp contract(fast)` (the FP32 path does).
# Packed FP32 instructions are disabled for all device code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
         f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    tag = os.path.splitext(os.path.basename(out))[0]
    defs = [f"-D{d}" for d in defines]

    def compile_one(src: str) -> str:
        obj = os.path.join(objdir, f"{tag}.{os.path.splitext(src)[0]}.o")
        cmd = [HIPCC, *FLAGS, *defs, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = out + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
