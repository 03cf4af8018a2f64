"""Builds libsemigpu.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m semigpu_amd.build          # or __graft_entry__.build()

hipcc compiles each translation unit to an object, then links one shared
library next to this file.  sg_integrals.hip is compiled without FMA
contraction so the integral formulas keep the reference's rounding.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsemigpu.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("SG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
UNITS = {
    "sg_integrals.hip": ["-ffp-contract=off"],
    "sg_fock.hip": [],
    "sg_linalg.hip": [],
    "sg_api.cpp": ["-x", "hip"],
    "sg_server.cpp": ["-x", "hip"],
}


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in UNITS.items():
        obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
