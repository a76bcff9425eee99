"""Builds msolve_amd/_lib/libmsolve_amd.so with hipcc for gfx950.

Usage: python -m msolve_amd.build  (or msolve_amd.build.build()).
Object files go to msolve_amd/_build/ (git-ignored); the shared library is
built in-tree so that it travels with the repository snapshot.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "libmsolve_amd.so")
SOURCES = ["kernels.hip", "capi.hip", "mesh.cpp", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-c", *FLAGS, os.path.join(SRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-c", "-x", "hip", *FLAGS, os.path.join(SRC, src), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (res.stderr or "").splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_host_check(verbose: bool = True) -> str:
    """Compile tests/cpp/host_api_check.cpp (the C++ host layer over the C ABI,
    include/msolve_amd.hpp) with the host compiler against the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "host_api_check.cpp")
    out_dir = os.path.join(root, "tests", "cpp", "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "host_api_check")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(root, "include"), src,
           "-o", out, "-L" + LIB_DIR, "-lmsolve_amd", "-Wl,-rpath,$ORIGIN/../../../msolve_amd/_lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("host check build failed:\n" + res.stderr)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build()
    build_host_check()
