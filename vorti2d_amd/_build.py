"""Build step of libvorti2d.so: hipcc for gfx950, in-tree.

``python vorti2d_amd/_build.py`` compiles ``vorti2d_amd/csrc/*.hip`` into
``vorti2d_amd/libvorti2d.so``.  Called by ``__graft_entry__.build()``.  Run it as
a script (not ``-m``): importing the package first would load the library that
is about to be rebuilt.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvorti2d.so")
SOURCES = ("v2d_kernels.hip", "v2d_direct.hip", "v2d_adjoint.hip", "v2d_capi.hip", "v2d_batch.hip", "v2d_noalloc.hip", "v2d_mf.hip")

# -ffp-contract=off keeps the reference operation order in the metric and
# assembly arithmetic (bit-identical FP64); the dense kernels use explicit fma.
# Packed FP32 instructions are disabled for the device code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libvorti2d.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
