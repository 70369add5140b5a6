"""Builds libfvm2d_gpu.so in-tree: HIP kernels for gfx950 + host code.

    python -m fvm2d_amd.build

Flags that matter for results:
  -ffp-contract=off      no fused multiply-add, so the FP64 path evaluates
                         every expression exactly as written;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUTDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUTDIR, "libfvm2d_gpu.so")
ARCH = os.environ.get("FVM2D_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
FORTRAN = os.path.join(HERE, "fortran")
HOST_EXE = os.path.join(OUTDIR, "fvm2d_gpu")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-pthread"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[0])}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(OUTDIR, exist_ok=True)
    obj_geom = os.path.join(OUTDIR, "geometry.o")
    obj_part = os.path.join(OUTDIR, "partition.o")
    obj_hip = os.path.join(OUTDIR, "fvm2d_gpu.o")
    _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "geometry.cpp"), "-o", obj_geom])
    _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "partition.cpp"), "-o", obj_part])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "fvm2d_gpu.hip"), "-o", obj_hip])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, obj_hip, obj_geom, obj_part,
          "-pthread", "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    if os.path.exists(FLANG):
        build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> str:
    """The Fortran host program (fortran/fvm2d_gpu.f90) linked against the
    library, found next to it at run time."""
    mod = os.path.join(OUTDIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in ("fvm_gpu_mod.f90", "fvm2d_gpu.f90"):
        obj = os.path.join(mod, src.replace(".f90", ".o"))
        # no fused multiply-add: the host's own arithmetic (the state rebuilt
        # from SOL, WALL.DAT, PARAM.DAT) rounds as the reference's build does
        _run([FLANG, "-O2", "-ffp-contract=off", "-J", mod, "-c", os.path.join(FORTRAN, src), "-o", obj])
        objs.append(obj)
    tmp = HOST_EXE + ".tmp"
    _run([FLANG, "-o", tmp, *objs, "-L", OUTDIR, "-lfvm2d_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_EXE)
    if verbose:
        print("built", HOST_EXE)
    return HOST_EXE


if __name__ == "__main__":
    build(verbose=True)
