"""Builds the MI355X engine library libmpm_mi355x.so in-tree with hipcc, and
the reference-signature host shim libmpm_refshim.so with g++.

Every HIP translation unit is compiled for gfx950 only. exact64.hip is built
with -ffp-contract=off (it reproduces the reference's FP64 arithmetic
operation by operation); all device code is built without packed FP32
instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmpm_mi355x.so")
BUILD = os.path.join(HERE, "_build")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
UNITS = [
    ("sort.hip", []),
    ("bin.hip", []),
    ("noalloc.hip", []),
    ("exact64.hip", ["-ffp-contract=off"]),
    ("fast32.hip", []),
    ("stencil32.hip", []),
    ("mlp.hip", []),
    ("surrogate.cpp", ["-x", "hip"]),
    ("engine.cpp", ["-x", "hip"]),
    ("case_io.cpp", ["-x", "c++"]),
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


SHIM_OUT = os.path.join(HERE, "libmpm_refshim.so")
SHIM_INC = os.path.join(os.path.dirname(HERE), "include", "mpm_ref")


def build_shim(verbose=False):
    """The reference-signature host shim (CPU FP64, include/mpm_ref/): plain
    C++17 with -ffp-contract=off, as the reference's own x86-64 build."""
    tmp = SHIM_OUT + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-I" + SHIM_INC,
           "-shared", "-o", tmp, os.path.join(HERE, "shim", "ref_shim.cpp")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, SHIM_OUT)
    return SHIM_OUT


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    build_shim(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
