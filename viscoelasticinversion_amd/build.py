"""Build the gfx950 shared library ``libvi_mi355x.so`` in-tree with hipcc.

Flags:
  -ffp-contract=off      every multiply and add rounds separately (the reference's Julia
                         arithmetic has no implicit FMA), which also makes the kernels
                         reproduce the CPU oracle bit for bit;
  -packed-fp32-ops off   no v_pk_* FP32 instructions are generated;
  kernarg preload 8      the first 8 dwords of every kernel's arguments (the descriptor
                         pointer, the tile table, the step) arrive in SGPRs when the wave
                         starts (gfx950), so the first descriptor load does not wait for a
                         kernarg load: -1 % per step at 1024x512 (15.22-15.30 against
                         15.39-15.46 us, three runs each, same box).
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libvi_mi355x.so")
SOURCES = [os.path.join(CSRC, "vi_api.hip")]
HEADERS = [os.path.join(CSRC, "vi_kernels.hpp"), os.path.join(CSRC, "vi_adjoint.hpp"),
           os.path.join(CSRC, "vi_rates.hpp"), os.path.join(ROOT, "include", "vi_mi355x.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-kernarg-preload-count=8",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    """Compile the library (always recompiles). Returns the path of the .so."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed to build libvi_mi355x.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
