"""Builds libmicpmd.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m cpmd_amd.build   (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmicpmd.so"

HIP_SOURCES = ["context.hip", "wfft.hip", "gemm.hip", "ortho.hip", "capi.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]


def hipcc() -> str:
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if p and Path(p).exists():
            return p
    return "hipcc"


def build(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in HIP_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(ROOT / "include"), *srcs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmicpmd.so")
    if verbose and (r.stdout or r.stderr):
        # the host pass warns that it does not know the packed-fp32 feature; harmless
        lines = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(f"built {LIB}")
