"""Compile libtis_hip.so (gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["tis_kernels.hip", "tis_capi.hip", "tis_bdhi.hip"]
OUT = os.path.join(HERE, "libtis_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # packed FP32 arithmetic off: ordinary scalar/float2 code only
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


FORTRAN = os.path.join(HERE, "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_EXE = os.path.join(FORTRAN, "tis_hip_check")


def build(verbose: bool = False) -> str:
    cmd = [HIPCC] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", OUT + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(OUT + ".tmp", OUT)
    build_fortran()
    return OUT


def build_fortran() -> str | None:
    """The Fortran host module (iso_c_binding) and its check program,
    linked against libtis_hip.so (rpath to the package directory)."""
    if not os.path.exists(FLANG):
        return None
    bld = os.path.join(FORTRAN, "_build")
    os.makedirs(bld, exist_ok=True)
    steps = [
        [FLANG, "-O2", "-module-dir", bld, "-c", os.path.join(FORTRAN, "tis_hip.F90"),
         "-o", os.path.join(bld, "tis_hip.o")],
        [FLANG, "-O2", "-I", bld, "-module-dir", bld, os.path.join(FORTRAN, "tis_hip_check.F90"),
         os.path.join(bld, "tis_hip.o"), "-L", HERE, "-ltis_hip", "-Wl,-rpath,$ORIGIN/..",
         "-o", FORTRAN_EXE],
    ]
    for cmd in steps:
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("flang failed:\n" + res.stdout + res.stderr)
    return FORTRAN_EXE


if __name__ == "__main__":
    print(build(verbose=True))
