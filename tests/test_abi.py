"""The C ABI boundary: every symbol include/aer_sht.h declares is exported by the
built library and bound in aeros_amd._lib; host-only entry points work without
a GPU; transforms refuse to run without a gfx950 device (no CPU fallback)."""
import ctypes
import os
import re

import pytest

from aeros_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "aer_sht.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aer_\w+)\s*\(", text)))


def test_header_declares_the_replacement_surface():
    syms = declared_symbols()
    # the five SHTns transforms aeros calls (src/aeros_spectral.f90:496, 511, 539, 554, 580)
    for s in ("aer_spat_to_SH", "aer_SH_to_spat", "aer_spat_to_SHsphtor", "aer_SHsphtor_to_spat",
              "aer_SHsph_to_spat", "aer_sht_create", "aer_sht_set_grid", "aer_sht_gauss_wts",
              "aer_sht_destroy"):
        assert s in syms
    assert len(syms) >= 28


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIBPATH)
    for s in declared_symbols():
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, s
    assert set(_lib.SIGNATURES) == set(declared_symbols())


def test_host_only_entry_points():
    L = _lib.lib()
    a, b = ctypes.c_int(), ctypes.c_int()
    for T, expect in ((21, (64, 32)), (31, (96, 48)), (42, (128, 64)), (85, (256, 128)), (106, (320, 160))):
        assert L.aer_sht_grid_size(T, ctypes.byref(a), ctypes.byref(b)) == 0
        assert (a.value, b.value) == expect
    assert L.aer_sht_grid_size(-1, ctypes.byref(a), ctypes.byref(b)) == _lib.AER_EINVAL
    h = ctypes.c_void_p()
    norm = _lib.AER_SHT_ORTHONORMAL | _lib.AER_SHT_NO_CS_PHASE
    assert L.aer_sht_create(21, 21, 1, norm, ctypes.byref(h)) == 0
    info = _lib.ShtInfo()
    assert L.aer_sht_get_info(h, ctypes.byref(info)) == 0 and info.nlm == 253
    li = (ctypes.c_int * 253)()
    mi = (ctypes.c_int * 253)()
    assert L.aer_sht_get_lm(h, li, mi) == 0
    assert (li[0], mi[0], li[252], mi[252]) == (0, 0, 21, 21)
    # a transform before set_grid is a state error, not a crash
    buf = (ctypes.c_double * 4096)()
    assert L.aer_SH_to_spat(h, buf, buf) == _lib.AER_ESTATE
    assert L.aer_sht_destroy(h) == 0
    # unsupported conventions are refused
    assert L.aer_sht_create(21, 21, 1, _lib.AER_SHT_ORTHONORMAL, ctypes.byref(h)) == _lib.AER_EINVAL
    assert L.aer_sht_create(21, 20, 1, norm, ctypes.byref(h)) == _lib.AER_EINVAL
    assert L.aer_sht_destroy(None) == 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_set_grid_fails_loudly_without_gpu():
    L = _lib.lib()
    h = ctypes.c_void_p()
    norm = _lib.AER_SHT_ORTHONORMAL | _lib.AER_SHT_NO_CS_PHASE
    assert L.aer_sht_create(21, 21, 1, norm, ctypes.byref(h)) == 0
    rc = L.aer_sht_set_grid(h, _lib.AER_SHT_PHI_CONTIGUOUS, 1e-10, 32, 64)
    assert rc in (_lib.AER_ENODEV, _lib.AER_EHIP)
    assert L.aer_last_error()
    L.aer_sht_destroy(h)
