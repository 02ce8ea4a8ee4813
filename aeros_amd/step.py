"""Python mirror of the spectral step completion over the C ABI: the reference's
aeros_semiimp module (src/dynamics/aeros_semiimp.f90) and the leapfrog pieces of
aeros_timestep_step (src/dynamics/aeros_timestep.f90:1043-1100, 1225-1276,
1673-1815), on the GPU in FP64.

Spectral fields are torch CUDA complex128 tensors: (nlev, nlm) for vorticity,
divergence and temperature, (nlm,) for ln p_s.
"""
import ctypes

from . import _lib
from ._lib import check, lib

P0 = 1.0e5   # reference src/aeros_defs.f90:103


def _ptrs(d, keys):
    return [d[k].data_ptr() for k in keys]


class SemiImplicit:
    """aeros_semiimp_class on the GPU, for one transform handle and vertical grid."""

    def __init__(self, sht, vg, dt_step, t_ref=300.0, ps_ref=P0, alpha=0.5):
        self.sht, self.vg = sht, vg
        h = ctypes.c_void_p()
        check(lib().aer_semiimp_create(sht._h, vg._h, t_ref, ps_ref, dt_step, alpha, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().aer_semiimp_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_step(self, dt_step):
        check(lib().aer_semiimp_set_step(self._h, dt_step))

    def set_diffusion(self, tau_diff, tau_diff_div=0.0, ndiff=6):
        check(lib().aer_semiimp_set_diffusion(self._h, tau_diff, tau_diff_div, ndiff))

    def gwspeed(self):
        c = ctypes.c_double()
        check(lib().aer_semiimp_gwspeed(self._h, ctypes.byref(c)))
        return c.value

    def linear(self, div, temp, lnps):
        import torch
        ld, lt, lp = torch.empty_like(div), torch.empty_like(temp), torch.empty_like(lnps)
        self.sht._bind_stream()
        check(lib().aer_semiimp_linear_d(self._h, div.data_ptr(), temp.data_ptr(), lnps.data_ptr(),
                                         ld.data_ptr(), lt.data_ptr(), lp.data_ptr()))
        return ld, lt, lp

    def step(self, old, now, tnd):
        """D, T, ln p_s at n+1 (dicts with div, temp, lnps)."""
        import torch
        new = {k: torch.empty_like(now[k]) for k in ("div", "temp", "lnps")}
        keys = ("div", "temp", "lnps")
        self.sht._bind_stream()
        check(lib().aer_semiimp_step_d(self._h, *_ptrs(old, keys), *_ptrs(now, keys), *_ptrs(tnd, keys),
                                       *_ptrs(new, keys)))
        return new

    def leapfrog(self, old, now, tnd, semi_implicit=True, diffuse=False, eps_filter=0.0, raw_alpha=1.0):
        """Sections 2-4 of aeros_timestep_step; `now` is filtered in place. Returns X^(n+1)."""
        import torch
        keys = ("vor", "div", "temp", "lnps")
        new = {k: torch.empty_like(now[k]) for k in keys}
        flags = ((_lib.AER_STEP_SEMI_IMPLICIT if semi_implicit else 0) |
                 (_lib.AER_STEP_DIFFUSE if diffuse else 0) |
                 (_lib.AER_STEP_FILTER if eps_filter > 0 else 0))
        self.sht._bind_stream()
        check(lib().aer_leapfrog_d(self._h, flags, eps_filter, raw_alpha, *_ptrs(old, keys), *_ptrs(now, keys),
                                   *_ptrs(tnd, keys), *_ptrs(new, keys)))
        return new


def mass(sht, lnps, fix=False, target=0.0):
    """Global dry-air mass of ln p_s on the device; with fix, the mass fixer applied
    to lnps in place. Returns (mass, ln r) as 0-d float64 CUDA tensors."""
    import torch
    m = torch.zeros((), dtype=torch.float64, device=lnps.device)
    r = torch.zeros((), dtype=torch.float64, device=lnps.device)
    sht._bind_stream()
    check(lib().aer_mass_d(sht._h, lnps.data_ptr(), 1 if fix else 0, float(target), m.data_ptr(), r.data_ptr()))
    return m, r
