"""Builds the native libraries of the package in-tree.

* ``joltphysics_amd/_lib/libjolt_xpbd.so``: HIP kernels for gfx950 + the C ABI
  runtime (include/jolt_xpbd.h) + the host settings builder.

Compiled with ``-ffp-contract=off`` (no FMA contraction, so every float
operation rounds like the reference's SSE code), IEEE division / square root
(hipcc's default correctly-rounded fp32 div/sqrt) and packed-FP32 instruction
selection disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIBNAME = "libjolt_xpbd.so"

HIP_SOURCES = ["xpbd_kernels.hip", "xpbd_bundle.hip", "xpbd_team.hip", "xpbd_rollout.hip", "xpbd_runtime.cpp", "settings_builder.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP extension cannot be built")


def library_path():
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *HIPCC_FLAGS, "-x", "hip", "-c", path, "-o", obj, "-I", os.path.join(ROOT, "include")]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    out = library_path()
    tmp = out + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
