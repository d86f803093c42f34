"""Builds the HIP library (libadcgpu.so) in-tree for gfx950 with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libadcgpu.so")
DRIVER = os.path.join(HERE, "adcgpu_run_case")
SOURCES = ["kernels.hip", "jcg.hip", "mlp.hip", "api.hip", "noalloc.hip"]

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
    # parity: no fused multiply-add anywhere in the FP64 path
    "-ffp-contract=off",
    # no packed FP32 instructions (the target feature is disabled)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return "hipcc"


def build(verbose: bool = False) -> str:
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(CSRC, s.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the native driver (C ABI only), linked against the library beside it
    cmd = [hipcc(), *HIPCC_FLAGS, os.path.join(CSRC, "run_case.hip"), "-L", HERE,
           "-ladcgpu", "-Wl,-rpath,$ORIGIN", "-o", DRIVER]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    build_fortran(verbose)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_CHECK = os.path.join(FORTRAN, "jcg_dropin_check")
FORTRAN_RESIDENT = os.path.join(FORTRAN, "resident_check")


def flang() -> str:
    for c in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang"):
        if c and os.path.exists(c):
            return c
    return ""


def build_fortran(verbose: bool = False) -> None:
    """The ISO_C_BINDING interface module and the programs that call the
    library through it: the drop-in JCG, and the per-step path with the
    bind(C) mesh and parameter types (skipped, with a message, when no flang
    exists)."""
    fc = flang()
    if not fc:
        print("flang not found: Fortran interface check not built", file=sys.stderr)
        return
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    iface_o = os.path.join(bdir, "adcgpu_iface.o")
    cmds = [
        [fc, "-O2", "-c", os.path.join(FORTRAN, "adcgpu_iface.F90"), "-module-dir", bdir,
         "-o", iface_o],
        [fc, "-O2", os.path.join(FORTRAN, "jcg_dropin_check.F90"), iface_o, "-I", bdir,
         "-L", HERE, "-ladcgpu", "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_CHECK],
        [fc, "-O2", os.path.join(FORTRAN, "resident_check.F90"), iface_o, "-I", bdir,
         "-L", HERE, "-ladcgpu", "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_RESIDENT],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
