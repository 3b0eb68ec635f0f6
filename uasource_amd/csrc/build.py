"""Builds ``uasource_amd/libuvh.so`` for gfx950 with hipcc (in-tree, no JIT cache).

Usage: ``python -m uasource_amd.csrc.build`` or ``build()`` from Python.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libuvh.so")
BUILD_DIR = os.path.join(HERE, "_build")

HIP_SOURCES = ["uvh_kernels.hip", "uvh_api.hip", "uvh_linsolve.hip"]
CXX_SOURCES = ["uvh_mesh.cpp", "uvh_schedule.cpp", "uvh_host_api.cpp"]
HEADERS = ["uvh_internal.h", os.path.join("..", "..", "include", "uvh.h")]

ARCH = os.environ.get("UVH_OFFLOAD_ARCH", "gfx950")
# -packed-fp32-ops: keep the compiler from emitting packed FP32 (v_pk_*_f32) instructions.
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]
CXXFLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build(verbose: bool = False) -> str:
    """Compile every source and link ``libuvh.so``; always rebuilds."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        cmd = [hipcc, *HIPFLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".cpp", ".o"))
        cmd = [os.environ.get("CXX", "g++"), *CXXFLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
