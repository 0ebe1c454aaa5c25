"""Build libccd_amd.so (and the ccd driver) in-tree for gfx950.

    python -m coupled_cluster_doubles_amd.build

hipcc cross-compiles the device code without a GPU.  The host problem model
is compiled by g++ with -ffp-contract=off so its arithmetic matches the
reference's x86-64 evaluation bit for bit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libccd_amd.so")
EXE = os.path.join(PKG, "ccd")
INC = os.path.join(ROOT, "include")

ARCH = os.environ.get("CCD_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", f"-I{INC}", f"-I{SRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", f"-I{INC}", f"-I{SRC}"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# HIP translation units: the context and the small kernels, and the grouped
# pppp-ladder instances split by arithmetic type and V source (the largest
# part of the build), compiled side by side.
HIP_UNITS = ["ccd_device", "ladder_f32_gas", "ladder_f32_other", "ladder_f64_gas", "ladder_f64_other",
             "ladder_f32_gasb", "ladder_f64_gasb"]


def _run_all(cmds):
    procs = []
    for cmd in cmds:
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
    failed = [cmd for cmd, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])


def build(verbose: bool = True) -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = [os.path.join(OUT, "model.o")]
    cmds = [["g++", *CXX_FLAGS, "-c", os.path.join(SRC, "model.cpp"), "-o", objs[0]]]
    for u in HIP_UNITS:
        o = os.path.join(OUT, u + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(SRC, u + ".hip"), "-o", o])
        objs.append(o)
    _run_all(cmds)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    main = os.path.join(SRC, "ccd_main.cpp")
    if os.path.exists(main):
        _run(["g++", *CXX_FLAGS, main, "-o", EXE, f"-L{PKG}", "-lccd_amd",
              f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
