"""Compile the Fortran drop-in modules and their test programs with amdflang.

Outputs go to aeros_amd/fortran/build/: the module files, libaeros_spectral_f.a
and the test executables, linked against ../libaer_sht.so with an rpath
relative to the executable, so the tree can move (as it does to a GPU host).
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")
MODULES = ["aeros_amd_kinds.f90", "aer_sht_c.f90", "aer_dyn_c.f90", "aeros_spectral.f90", "aeros_vordiv.f90"]
TESTS = ["test_host", "test_spectral", "test_vordiv", "test_step"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=BUILD, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        print(r.stdout)
        raise RuntimeError("Fortran build failed: " + " ".join(cmd))


def build(lib, verbose=False):
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    os.makedirs(BUILD, exist_ok=True)
    flags = ["-O2", "-fopenmp", "-J", BUILD]
    objs = []
    for m in MODULES:
        obj = os.path.join(BUILD, m.replace(".f90", ".o"))
        _run([fc, *flags, "-c", os.path.join(HERE, m), "-o", obj], verbose)
        objs.append(obj)
    _run(["ar", "rcs", os.path.join(BUILD, "libaeros_spectral_f.a"), *objs], verbose)
    libdir = os.path.dirname(lib)
    for t in TESTS:
        exe = os.path.join(BUILD, t + ".x")
        _run([fc, *flags, os.path.join(HERE, "tests", t + ".f90"), *objs, "-o", exe,
              "-L" + libdir, "-laer_sht", "-Wl,-rpath,$ORIGIN/../.."], verbose)
    return BUILD
