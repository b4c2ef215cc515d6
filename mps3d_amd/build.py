"""Builds the HIP library in-tree: mps3d_amd/lib/libmps3d_hip.so (gfx950).

    python -m mps3d_amd.build

The library is one translation unit (csrc/mps3d.hip) compiled by hipcc for
gfx950 only. FMA contraction is off so the FP64 operator reproduces the
reference's rounding; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "mps3d.hip")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libmps3d_hip.so")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, out=None, defines=()):
    """Compiles the library; `out` and `defines` (e.g. "MPS_SWEEP_TY=8") build
    a variant elsewhere for tuning."""
    target = out or LIB
    os.makedirs(os.path.dirname(target), exist_ok=True)
    tmp = target + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stderr)
        raise RuntimeError("hipcc failed building %s" % SRC)
    os.replace(tmp, target)
    return target


if __name__ == "__main__":
    print(build(verbose=True))
