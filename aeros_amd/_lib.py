"""ctypes binding of libaer_sht.so (the C ABI in include/aer_sht.h).

The library is built in-tree by aeros_amd.build (or __graft_entry__.build()).
There is no fallback: if the shared object is missing or cannot be loaded,
every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libaer_sht.so")

AER_OK, AER_EINVAL, AER_ENOMEM, AER_EHIP, AER_ENODEV, AER_ESTATE = range(6)

AER_SHT_ORTHONORMAL = 0x1
AER_SHT_NO_CS_PHASE = 0x2
AER_SHT_PHI_CONTIGUOUS = 0x10
AER_SHT_GAUSS = 0x20
AER_SHT_QUICK_INIT = 0x40
AER_SHT_LOAD_SAVE_CFG = 0x80

AER_PREC_FP64, AER_PREC_FP32, AER_PREC_SPLIT = 0, 1, 2
AER_PATH_TILED, AER_PATH_STAGED = 0, 1
AER_NKERNEL_KIND = 5

# Every symbol include/aer_sht.h declares, with (restype, argtypes).
_c_int, _c_long, _c_double, _vp = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


class ShtInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("lmax", "mmax", "mres", "nlm", "nlat", "nphi", "nspat", "nlat_2")]


SIGNATURES = {
    "aer_strerror": (ctypes.c_char_p, [_c_int]),
    "aer_last_error": (ctypes.c_char_p, []),
    "aer_sht_grid_size": (_c_int, [_c_int, _ip, _ip]),
    "aer_sht_create": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.POINTER(_vp)]),
    "aer_sht_set_grid": (_c_int, [_vp, _c_int, _c_double, _c_int, _c_int]),
    "aer_sht_get_info": (_c_int, [_vp, ctypes.POINTER(ShtInfo)]),
    "aer_sht_get_lm": (_c_int, [_vp, _ip, _ip]),
    "aer_sht_get_cos_theta": (_c_int, [_vp, _dp]),
    "aer_sht_gauss_wts": (_c_int, [_vp, _dp]),
    "aer_sht_host_tables": (_c_int, [_c_int, _c_int, _dp, _dp, _dp, _dp, _dp]),
    "aer_sht_destroy": (_c_int, [_vp]),
    "aer_sht_set_precision": (_c_int, [_vp, _c_int]),
    "aer_sht_set_stream": (_c_int, [_vp, _vp]),
    "aer_sht_set_path": (_c_int, [_vp, _c_int]),
    "aer_sht_sync": (_c_int, [_vp]),
    "aer_sht_timing": (_c_int, [_vp, _c_int]),
    "aer_sht_timing_read": (_c_int, [_vp, _dp, _ip]),
    "aer_spat_to_SH": (_c_int, [_vp, _vp, _vp]),
    "aer_SH_to_spat": (_c_int, [_vp, _vp, _vp]),
    "aer_spat_to_SHsphtor": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "aer_SHsphtor_to_spat": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "aer_SHsph_to_spat": (_c_int, [_vp, _vp, _vp, _vp]),
    "aer_synth_d": (_c_int, [_vp, _c_int, _vp, _c_long, _vp, _c_long]),
    "aer_anal_d": (_c_int, [_vp, _c_int, _vp, _c_long, _vp, _c_long]),
    "aer_synth_vec_d": (_c_int, [_vp, _c_int, _vp, _vp, _c_long, _vp, _vp, _c_long]),
    "aer_anal_vec_d": (_c_int, [_vp, _c_int, _vp, _vp, _c_long, _vp, _vp, _c_long]),
    "aer_synth_grad_d": (_c_int, [_vp, _c_int, _vp, _c_long, _vp, _vp, _c_long]),
    "aer_uv_from_vordiv_d": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _vp]),
    "aer_vordiv_from_uv_d": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _vp]),
    "aer_gradient_d": (_c_int, [_vp, _c_int, _vp, _vp, _vp]),
    "aer_tendency_grid_transforms_d": (_c_int, [_vp, _c_int] + [_vp] * 14),
    "aer_tendency_spectral_d": (_c_int, [_vp, _c_int] + [_vp] * 9),
}


class AerJob(ctypes.Structure):
    _fields_ = [("kind", _c_int), ("nfield", _c_int),
                ("in0", _vp), ("in1", _vp), ("in2", _vp),
                ("out0", _vp), ("out1", _vp),
                ("in_stride", _c_long), ("out_stride", _c_long),
                ("mode0", _c_int), ("mode1", _c_int), ("mode2", _c_int),
                ("scale0", _c_double), ("scale1", _c_double),
                ("accumulate", _c_int), ("prec", _c_int)]


AER_JOB_SYNTH_SCALAR, AER_JOB_SYNTH_VECTOR, AER_JOB_SYNTH_GRAD, AER_JOB_ANAL_SCALAR, AER_JOB_ANAL_VECTOR = range(5)
(AER_MODE_ONE, AER_MODE_S_FROM_DIV, AER_MODE_T_FROM_VOR, AER_MODE_INV_LL1, AER_MODE_DIV_FROM_S,
 AER_MODE_VOR_FROM_T, AER_MODE_NEG_LAP, AER_MODE_ZERO) = range(8)
AER_MAX_JOBS = 8
AER_JOB_PREC_HANDLE, AER_JOB_PREC_FP64, AER_JOB_PREC_FP32, AER_JOB_PREC_SPLIT = 0, 1, 2, 3
SIGNATURES["aer_run_jobs_d"] = (_c_int, [_vp, _c_int, ctypes.POINTER(AerJob)])
SIGNATURES["aer_plan_create"] = (_c_int, [_c_int, ctypes.POINTER(_vp), ctypes.POINTER(AerJob), ctypes.POINTER(_vp)])
SIGNATURES["aer_plan_run"] = (_c_int, [_vp, _vp])
SIGNATURES["aer_plan_destroy"] = (_c_int, [_vp])
AER_TUNE_NCAND = 4
SIGNATURES["aer_autotune_jobs"] = (_c_int, [_vp, _c_int, ctypes.POINTER(AerJob), _c_int, _dp, _ip])
SIGNATURES["aer_plan_autotune"] = (_c_int, [_vp, _vp, _c_int, _dp, _ip])
SIGNATURES["aer_synth_adj_d"] = (_c_int, [_vp, _c_int, _vp, _c_long, _vp, _c_long])
SIGNATURES["aer_anal_adj_d"] = (_c_int, [_vp, _c_int, _vp, _c_long, _vp, _c_long])
SIGNATURES["aer_synth_vec_adj_d"] = (_c_int, [_vp, _c_int, _vp, _vp, _c_long, _vp, _vp, _c_long])
SIGNATURES["aer_anal_vec_adj_d"] = (_c_int, [_vp, _c_int, _vp, _vp, _c_long, _vp, _vp, _c_long])
SIGNATURES["aer_synth_grad_adj_d"] = (_c_int, [_vp, _c_int, _vp, _vp, _c_long, _vp, _c_long])
SIGNATURES["aer_vgrid_create"] = (_c_int, [_c_int, _dp, _dp, ctypes.POINTER(_vp)])
SIGNATURES["aer_vgrid_destroy"] = (_c_int, [_vp])
SIGNATURES["aer_column_terms_d"] = (_c_int, [_vp, _vp, _c_int] + [_vp] * 16)
SIGNATURES["aer_column_terms_dteq_d"] = (_c_int, [_vp, _vp] + [_vp] * 17)
SIGNATURES["aer_column_terms_adj_d"] = (_c_int, [_vp, _vp, _c_int] + [_vp] * 28)
SIGNATURES["aer_semiimp_create"] = (_c_int, [_vp, _vp, _c_double, _c_double, _c_double, _c_double, ctypes.POINTER(_vp)])
SIGNATURES["aer_semiimp_set_step"] = (_c_int, [_vp, _c_double])
SIGNATURES["aer_semiimp_set_diffusion"] = (_c_int, [_vp, _c_double, _c_double, _c_int])
SIGNATURES["aer_semiimp_gwspeed"] = (_c_int, [_vp, _dp])
SIGNATURES["aer_semiimp_destroy"] = (_c_int, [_vp])
SIGNATURES["aer_semiimp_linear_d"] = (_c_int, [_vp] + [_vp] * 6)
SIGNATURES["aer_semiimp_step_d"] = (_c_int, [_vp] + [_vp] * 12)
SIGNATURES["aer_leapfrog_d"] = (_c_int, [_vp, _c_int, _c_double, _c_double] + [_vp] * 16)
SIGNATURES["aer_mass_d"] = (_c_int, [_vp, _vp, _c_int, _c_double, _vp, _vp])
SIGNATURES["aer_leapfrog_adj_d"] = (_c_int, [_vp, _c_int, _c_double, _c_double] + [_vp] * 20)
AER_STEP_SEMI_IMPLICIT, AER_STEP_DIFFUSE, AER_STEP_FILTER = 1, 2, 4
SIGNATURES["aer_tendency_grid_transforms_batch_d"] = (_c_int, [_vp, _c_int, _c_int] + [_vp] * 14)
SIGNATURES["aer_tendency_spectral_batch_d"] = (_c_int, [_vp, _c_int, _c_int] + [_vp] * 9)
SIGNATURES["aer_column_terms_batch_d"] = (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _c_long] + [_vp] * 16)
SIGNATURES["aer_column_terms_adj_batch_d"] = (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _c_long] + [_vp] * 27)
SIGNATURES["aer_leapfrog_batch_d"] = (_c_int, [_vp, _c_int, _c_int, _c_double, _c_double] + [_vp] * 16)
SIGNATURES["aer_leapfrog_adj_batch_d"] = (_c_int, [_vp, _c_int, _c_int, _c_double, _c_double] + [_vp] * 20)
SIGNATURES["aer_mass_batch_d"] = (_c_int, [_vp, _c_int, _vp, _c_int, _vp, _vp, _vp])
SIGNATURES["aer_diag_create"] = (_c_int, [_vp, _c_int, ctypes.POINTER(_vp)])
SIGNATURES["aer_diag_accumulate_d"] = (_c_int, [_vp, _vp, _vp, _vp, _vp])
SIGNATURES["aer_diag_read"] = (_c_int, [_vp, _ip, _dp])
SIGNATURES["aer_diag_reset"] = (_c_int, [_vp])
SIGNATURES["aer_diag_destroy"] = (_c_int, [_vp])
SIGNATURES["aer_dev_malloc"] = (_c_int, [ctypes.c_size_t, ctypes.POINTER(_vp)])
SIGNATURES["aer_dev_free"] = (_c_int, [_vp])
SIGNATURES["aer_memcpy_h2d"] = (_c_int, [_vp, _vp, ctypes.c_size_t])
SIGNATURES["aer_memcpy_d2h"] = (_c_int, [_vp, _vp, ctypes.c_size_t])


class AerError(RuntimeError):
    def __init__(self, status, detail):
        self.status = status
        super().__init__(f"aer status {status}: {detail}")


_lib = None


def lib():
    """Load libaer_sht.so once; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != AER_OK:
        raise AerError(rc, lib().aer_last_error().decode())
    return rc
