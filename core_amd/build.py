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
CSRC = os.path.join(HERE, "csrc")
# translation units, compiled in parallel: the host code and small kernels, and the transform
# kernel instantiations (the slow ones to compile) in three units of their own
SOURCES = ["irls2d.hip", "fft_e12.hip", "fft_e16.hip", "fft_e24.hip", "long_passes.hip", "goldstein.hip"]
SRC = os.path.join(CSRC, "irls2d.hip")
OUT = os.path.join(HERE, "libirls2d.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IRLS2D_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions (v_pk_*_f32) in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def command(out: str = OUT, extra=()):
    """Single-invocation form (all sources at once), kept for reference builds."""
    return [HIPCC, *FLAGS, "-shared", *extra, *[os.path.join(CSRC, f) for f in SOURCES], "-o", out]


def _newest_header() -> float:
    """Modification time of the newest header (every unit includes them)."""
    paths = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    return max(os.path.getmtime(p) for p in paths + [os.path.join(ROOT, "include", "irls2d.h")])


def build(verbose: bool = True, incremental: bool = False) -> str:
    """Compile every unit and link the library; with `incremental`, units whose object file is
    newer than every source and header are not recompiled."""
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    headers = _newest_header() if incremental else None
    procs, objs = [], []
    for f in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(f)[0] + ".o")
        objs_ok = (incremental and os.path.exists(obj) and
                   os.path.getmtime(obj) > max(headers, os.path.getmtime(os.path.join(CSRC, f))))
        if objs_ok:
            objs.append(obj)
            continue
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, f), "-o", obj]
        if verbose:
            print(" ".join(shlex.quote(c) for c in cmd), flush=True)
        procs.append((f, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [f for f, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    tmp = OUT + ".tmp"
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(shlex.quote(c) for c in link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
