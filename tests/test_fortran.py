"""The Fortran drop-in (aeros_amd/fortran): the module builds with amdflang and its
test programs -- restatements of the reference's tests/test_spectral.f90 and
tests/test_vordiv.f90 -- pass. The executables are built by __graft_entry__.build()."""
import os
import subprocess

import pytest

BUILD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "aeros_amd", "fortran", "build")


def run(name, timeout=120):
    exe = os.path.join(BUILD, name + ".x")
    assert os.path.exists(exe), f"{exe} missing: run __graft_entry__.build()"
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    print(r.stdout)
    return r


def test_fortran_module_built_and_host_checks():
    for f in ("aeros_spectral.mod", "aeros_vordiv.mod", "aer_sht_c.mod", "libaeros_spectral_f.a"):
        assert os.path.exists(os.path.join(BUILD, f)), f
    r = run("test_host")
    assert r.returncode == 0 and "test_host:: PASS" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["test_spectral", "test_vordiv"])
def test_fortran_acceptance_on_gpu(name):
    r = run(name)
    assert r.returncode == 0 and f"{name}:: PASS" in r.stdout
