"""Build libimpactz_hip.so (HIP, gfx950) in-tree with hipcc.

The library has no PyTorch dependency: the C ABI in include/impactz_hip.h is
the whole interface, bound from Python with ctypes and from Fortran with
iso_c_binding.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # impact-z_amd/
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libimpactz_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["izh_api.hip", "izh_particles.hip", "izh_field.hip", "izh_fused.hip",
           "izh_wake.hip"]

# Packed FP32 instructions are disabled for all device code: they have given
# intermittently wrong results on this hardware generation's toolchain.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


FC = os.environ.get("IZH_FC", "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(PKG, "fortran")
FORTRAN_EXE = os.path.join(FORTRAN, "izh_lattice")


def build_fortran(verbose: bool = True) -> str:
    """The Fortran 90 host (iso_c_binding module + lattice driver), linked
    against the library with an rpath to it."""
    objdir = os.path.join(PKG, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in ("impactz_hip.f90", "izh_lattice.f90"):
        obj = os.path.join(objdir, src.replace(".f90", ".o"))
        cmd = [FC, "-O2", "-module-dir", objdir, "-c", os.path.join(FORTRAN, src),
               "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    cmd = [FC, "-O2", *objs, "-L", PKG, "-limpactz_hip", "-Wl,-rpath,$ORIGIN/..",
           "-o", FORTRAN_EXE]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return FORTRAN_EXE


def build(verbose: bool = True, jobs: int = 4) -> str:
    """Compile every translation unit, then link the shared library."""
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        while len([p for p in procs if p.poll() is None]) >= jobs:
            procs[0].wait()
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", LIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    if os.access(FC, os.X_OK):
        build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
