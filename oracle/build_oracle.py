"""Build recipe for the FP64 CPU oracle (test infrastructure only).

Compiles oracle/sht_oracle.c, cpu_port.c (the CPU-baseline arrangement of its scalar
transforms), column_oracle.c and step_oracle.c with gcc and OpenMP into
oracle/liborc.so. There is no oracle/_ref target: the reference's
transform arithmetic is the third-party SHTns C library, which the reference
does not vendor, and its column and step code is Fortran that needs the absent
fesm-utils modules (nml, ncio, precision), so none of it builds here.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(HERE, f) for f in ("sht_oracle.c", "cpu_port.c", "column_oracle.c", "step_oracle.c")]
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c11",
           "-Wall", "-Wno-unused-function", *srcs, "-o", LIB, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
