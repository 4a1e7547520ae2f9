"""Build recipe of libhybird_amd.so (hipcc, gfx950, in-tree)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhybird_amd.so")
SOURCES = ["lb.hip", "fs.hip", "adj.hip", "dem.hip", "api.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# FP64 kernels keep the reference's operation order: no FMA contraction unless a
# kernel opts in with `#pragma clang fp contract(fast)` (the FP32 path does).
# Packed FP32 instructions are disabled for all device code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
         f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
