"""Build libsemhip.so (HIP kernels for gfx950 + host code) in-tree.

Usage: python ma-shm_amd/build.py   (also called by __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "semhip" / "libsemhip.so"
BUILD = HERE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SEM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (hipcc emits them from plain float code otherwise)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _run(cmd):
    print(" ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    for src in ("sem_host.cpp", "sem_pzt.cpp"):
        obj = BUILD / src.replace(".cpp", ".o")
        _run(["g++", *HOST_FLAGS, "-c", CSRC / src, "-o", obj])
        objs.append(obj)
    # the HIP translation units compile in parallel (the kernel files hold many template
    # instances)
    procs = []
    for src in ("sem_kernels.hip", "sem_solid_line.hip", "sem_solid_stage.hip", "sem_shell_line.hip",
                "sem_shell_mfma.hip", "sem_grid.hip", "sem_api.hip"):
        obj = BUILD / (src.replace(".hip", ".o"))
        cmd = [str(c) for c in (HIPCC, *HIP_FLAGS, "-c", CSRC / src, "-o", obj)]
        print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
