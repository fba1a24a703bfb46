"""Build the HIP shared library libsdpd.so in-tree for gfx950, and the
Fortran host mcf_amd/fortran/mcf_sdpd linked against it.

Run as ``python -m mcf_amd.build`` (or through __graft_entry__.build()).
hipcc cross-compiles for gfx950 without a GPU present; the Fortran host is
built with amdflang when it is installed.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsdpd.so")
FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_HOST = os.path.join(FORTRAN_DIR, "mcf_sdpd")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a * b + c is fused only inside one source expression, so every kernel
    # instantiation that inlines the same pair body rounds it the same way
    # (the list walk and the rescan give bitwise equal results)
    "-ffp-contract=on",
]


def hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def amdflang():
    for cand in ("/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran_host(verbose: bool = False):
    """Compile the iso_c_binding module and the host program; the binary
    finds libsdpd.so next to its own directory through its rpath."""
    fc = amdflang()
    if fc is None:
        if verbose:
            print("amdflang not found: Fortran host not built", flush=True)
        return None
    srcs = [os.path.join(FORTRAN_DIR, f) for f in ("sdpd_c.f90", "mcf_sdpd.f90")]
    cmd = [fc, "-O2", "-module-dir", FORTRAN_DIR, "-o", FORTRAN_HOST + ".tmp"] + srcs + \
          ["-L", HERE, "-lsdpd", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(FORTRAN_HOST + ".tmp", FORTRAN_HOST)
    return FORTRAN_HOST


def build(verbose: bool = False) -> str:
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", LIB + ".tmp"] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    build_fortran_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
