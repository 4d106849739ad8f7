"""Builds libspbp.so (HIP, gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libspbp.so")
SOURCES = [os.path.join(HERE, "csrc", "spbp.hip")]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # fp32 arithmetic exactly as written: the sample geometry must round like
    # the reference's single-precision code (no contraction into FMA)
    "-ffp-contract=off",
    # hardware fp32 atomic add on global memory (no CAS loop)
    "-munsafe-fp-atomics",
    # no packed-fp32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_SOURCES = [os.path.join(HERE, "fortran", "spbp_mod.f90"), os.path.join(HERE, "fortran", "bp_drv.f90")]
DRIVER = os.path.join(HERE, "spbp_bp")


def build_fortran(verbose: bool = False) -> str:
    """The `BP 3F` / `BP 32F` Fortran driver (ISO_C_BINDING on libspbp)."""
    mod_dir = os.path.join(HERE, "fortran", "_build")
    os.makedirs(mod_dir, exist_ok=True)
    cmd = [FLANG, "-O2", "-J", mod_dir, "-o", DRIVER + ".tmp", *FORTRAN_SOURCES,
           "-L" + HERE, "-lspbp", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", LIB + ".tmp", *SOURCES,
           "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
