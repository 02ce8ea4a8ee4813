"""Build the native parts in-tree: the gfx950 HIP library and the Fortran shim.

    python -m aeros_amd.build            # library + Fortran module and test programs

hipcc cross-compiles gfx950 code on a machine without a GPU, so this runs
anywhere the ROCm toolchain is installed.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libaer_sht.so")

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
               # no packed-FP32 instructions in generated code
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose, cwd=None):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines() if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        print(out, file=sys.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    if verbose and out.strip():
        print(out)


def build_library(verbose=False):
    src = os.path.join(CSRC, "aer_sht.hip")
    tmp = LIB + ".tmp"
    _run(["hipcc", *HIPCC_FLAGS, "-shared", src, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    from aeros_amd.fortran import build_fortran as bf
    return bf.build(LIB, verbose=verbose)


def build(verbose=False):
    build_library(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
