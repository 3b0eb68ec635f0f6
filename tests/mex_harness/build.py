"""Builds tests/mex_harness/libmexharness.so: the MEX gateway (uasource_amd/mex/uvhAssemblyGPU_mex.c)
with the test double of the MATLAB C API (mex_harness.c), linked to the in-tree libuvh.so.
Run after the library build (``__graft_entry__.build()`` does both)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "libmexharness.so")


def build() -> str:
    cmd = [os.environ.get("CC", "gcc"), "-std=c99", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra",
           "-I" + HERE, "-I" + os.path.join(REPO, "include"),
           os.path.join(HERE, "mex_harness.c"), os.path.join(REPO, "uasource_amd", "mex", "uvhAssemblyGPU_mex.c"),
           "-L" + os.path.join(REPO, "uasource_amd"), "-luvh", "-Wl,-rpath,$ORIGIN/../../uasource_amd",
           "-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("mex harness build failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
