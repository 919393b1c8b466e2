"""Builds libpot3d_hip.so (gfx950) in-tree with hipcc.

The library is plain HIP C++ behind a C ABI (include/pot3d_hip.h); no torch
extension is involved.  Every HIP source is compiled with packed-FP32
instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpot3d_hip.so")
SOURCES = ["p3d_kernels.hip", "p3d_sweep.hip", "p3d_solver.hip", "p3d_legacy.hip",
           "p3d_nn_kernels.hip", "p3d_nn.hip", "p3d_readbr.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("P3D_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into pot3d_amd/libpot3d_hip.so; returns its path."""
    cmd = ([HIPCC, *FLAGS, "-o", LIB + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES] +
           ["-lhsa-runtime64"])
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
