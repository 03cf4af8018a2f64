"""Builds libsemigpu.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m semigpu_amd.build          # or __graft_entry__.build()

hipcc compiles each translation unit to an object, then links one shared
library next to this file.  sg_integrals.hip is compiled without FMA
contraction so the integral formulas keep the reference's rounding.

The C++ classes with the reference's names (semigpu_amd/cpp) are host code
over the C ABI: g++ builds them into libsemigpu_cpp.so and links the semiGPU
command-line program against both libraries.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsemigpu.so")
CPP = os.path.join(HERE, "cpp")
OUT_CPP = os.path.join(HERE, "libsemigpu_cpp.so")
OUT_EXE = os.path.join(HERE, "semiGPU")
TEST_CPP = os.path.join(os.path.dirname(HERE), "tests", "cpp")  # test drivers of the classes
GXX = os.environ.get("CXX", "g++")
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
    "sg_large.hip": [],
    "sg_diis.hip": [],
    "sg_gradient.hip": ["-ffp-contract=off"],
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
    build_cpp(verbose)
    return OUT


def build_cpp(verbose: bool = False) -> str:
    """libsemigpu_cpp.so (reference-named classes) and the semiGPU program."""
    srcs = sorted(os.path.join(CPP, f) for f in os.listdir(CPP) if f.endswith(".cpp"))
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", CPP]
    link = ["-L", HERE, "-Wl,-rpath,$ORIGIN"]
    steps = [
        [GXX] + flags + ["-shared", "-o", OUT_CPP + ".tmp"] + srcs + link + ["-lsemigpu"],
        [GXX] + flags + ["-o", OUT_EXE + ".tmp", os.path.join(CPP, "semiGPU.cc")] + link +
        ["-lsemigpu_cpp", "-lsemigpu"],
    ]
    outs = [OUT_CPP, OUT_EXE]
    if os.path.isdir(TEST_CPP):
        tlink = ["-L", HERE, "-Wl,-rpath,$ORIGIN/../../semigpu_amd", "-lsemigpu_cpp", "-lsemigpu"]
        for f in sorted(os.listdir(TEST_CPP)):
            if f.endswith(".cc"):
                exe = os.path.join(TEST_CPP, f[:-3])
                steps.append([GXX] + flags + ["-o", exe + ".tmp", os.path.join(TEST_CPP, f)] +
                             tlink)
                outs.append(exe)
    for cmd, out in zip(steps, outs):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)
    return OUT_EXE


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
