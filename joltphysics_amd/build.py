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

HIP_SOURCES = ["xpbd_kernels.hip", "xpbd_bundle.hip", "xpbd_team.hip", "xpbd_skin.hip", "xpbd_rollout.hip", "xpbd_runtime.cpp", "settings_builder.cpp"]

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


# Host code checked with AddressSanitizer and UndefinedBehaviorSanitizer: the runtime and the
# settings builder are rebuilt with the sanitizers on their host pass only (the device code of
# every source, and the kernel objects, are the regular build's).  Used by the CPU tests
# through JOLT_XPBD_LIBRARY with the sanitizer runtime preloaded.
HOST_SOURCES = ["xpbd_runtime.cpp", "settings_builder.cpp"]
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]


def sanitizer_runtime():
    out = subprocess.run([hipcc(), "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def asan_library_path():
    return os.path.join(LIBDIR, "libjolt_xpbd_asan.so")


def build_sanitized(verbose=False):
    """libjolt_xpbd_asan.so: host sources with ASan + UBSan, kernels from the regular build."""
    objdir = os.path.join(LIBDIR, "obj")
    cc = hipcc()
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        if src in HOST_SOURCES:
            obj = os.path.join(objdir, os.path.splitext(src)[0] + ".asan.o")
            cmd = [cc, *HIPCC_FLAGS, *SANITIZE, "-O1", "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj,
                   "-I", os.path.join(ROOT, "include")]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
        elif not os.path.exists(obj):
            raise RuntimeError("build the regular library first")
        objs.append(obj)
    out = asan_library_path()
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
           "-fsanitize=undefined", "-o", out + ".tmp", *objs]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    if "--sanitized" in sys.argv:
        print(build_sanitized(verbose="-v" in sys.argv))
