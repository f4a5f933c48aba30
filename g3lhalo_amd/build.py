"""Builds libg3l_hip.so (gfx950, hipcc), libg3lhalo.so (the C++ API) and the
drop-in drivers (bin/*.x) in-tree.

Usage: python -m g3lhalo_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libg3l_hip.so")
ARCH = os.environ.get("G3L_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no packed FP32 instructions are generated (they have
# produced silent wrong results on these parts).
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]

SOURCES = ["kernels.hip", "engine.cpp", "server.cpp", "round_graph.cpp", "precompute.cpp",
           "noalloc.cpp"]


def hipcc() -> str:
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found")
    return h


CXX_SOURCES = ["params_hod_cosmo.cpp", "ingest.cpp", "models.cpp"]
DRIVERS = ["calculateNNMap", "calculateNNMap_twoPop", "calculateNMapMap", "doPrecomputations"]
CXX_LIB = os.path.join(HERE, "libg3lhalo.so")
BIN = os.path.join(HERE, "bin")


def build_cxx(verbose: bool = False) -> None:
    """The reference-compatible C++ API (libg3lhalo.so) and the drop-in drivers."""
    cxx = os.environ.get("CXX", "g++")
    inc = os.path.join(os.path.dirname(HERE), "include")
    srcs = [os.path.join(HERE, "cxx", s) for s in CXX_SOURCES]
    flags = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-I", inc]
    cmd = [cxx, *flags, "-shared", *srcs, "-o", CXX_LIB + ".tmp", "-L", HERE, "-l:libg3l_hip.so",
           "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(CXX_LIB + ".tmp", CXX_LIB)
    os.makedirs(BIN, exist_ok=True)
    for d in DRIVERS:
        out = os.path.join(BIN, d + ".x")
        cmd = [cxx, *flags, os.path.join(HERE, "cxx", "drivers", d + ".cpp"), "-o", out + ".tmp",
               "-L", HERE, "-l:libg3lhalo.so", "-l:libg3l_hip.so", "-Wl,-rpath,$ORIGIN/.."]
        _run(cmd, verbose)
        os.replace(out + ".tmp", out)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *COMMON, "-shared", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    build_cxx(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
