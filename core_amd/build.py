"""Build the gfx950 solver library (core_amd/libirls2d.so) in-tree with hipcc.

    python -m core_amd.build          # or core_amd.build.build()

The library is a plain C-ABI shared object (include/irls2d.h); Python binds it with ctypes.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "irls2d.hip")
OUT = os.path.join(HERE, "libirls2d.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IRLS2D_ARCH", "gfx950")


def command(out: str = OUT, extra=()):
    return [
        HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        # no packed-FP32 instructions (v_pk_*_f32) in the device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
        *extra, SRC, "-o", out,
    ]


def build(verbose: bool = True) -> str:
    tmp = OUT + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
