"""Builds libxlumina_hip.so in-tree for gfx950 with hipcc.

Usage: ``python -m xlumina_amd.build_lib`` (or ``__graft_entry__.build()``).
The library is one HIP translation unit (``csrc/xlumina.hip``) plus the
kernels it includes; the result lands in ``xlumina_amd/_lib/``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIBNAME = "libxlumina_hip.so"
ARCH = os.environ.get("XL_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-shared",
    "-fPIC",
    # packed-FP32 instructions stay off (plain FP32 only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def lib_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libxlumina_hip.so)")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, os.path.join(CSRC, "xlumina.hip"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    errs = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{errs}")
    if verbose and errs.strip():
        print(errs, file=sys.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
