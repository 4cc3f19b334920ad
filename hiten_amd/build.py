"""Build the HIP library ``hiten_amd/libhiten_hip.so`` for gfx950.

    python -m hiten_amd.build

The library is the HIP kernels and C-ABI (hh_kernels.hip) plus host-side
native code (hh_bank.cpp), compiled by hipcc.  ``-ffp-contract=off``
is required for parity: it stops clang from fusing ``a*b + c`` into an FMA,
which numba (fastmath=False) never does in the reference.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libhiten_hip.so")
SOURCES = [os.path.join(CSRC, "hh_kernels.hip"), os.path.join(CSRC, "hh_bank.cpp")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("hh_device.h", "hh_wave.h", "hh_pow.h",
                                                  "hh_tables.h", "hh_pair.h", "hh_group6.h",
                                                  "hh_synodic.h")] + [
    os.path.join(ROOT, "include", "hiten_hip.h")]

ARCH = os.environ.get("HH_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build(verbose: bool = False) -> str:
    """Compile the library (always, even if an up-to-date build exists)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
