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
KEYS = ("vor", "div", "temp", "lnps")


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

    def _members(self, d):
        """Member count of a state dict: vor (nlev, nlm) -> 1, (nmem, nlev, nlm) -> nmem."""
        v, p = d["vor"], d["lnps"]
        if v.dim() not in (2, 3) or tuple(p.shape) != tuple(v.shape[:-2]) + (v.shape[-1],):
            raise ValueError("state arrays must be ([nmem,] nlev, nlm) and ln p_s ([nmem,] nlm)")
        for k in KEYS:
            t = d[k]
            if not (t.is_cuda and t.is_contiguous() and t.dtype.is_complex):
                raise ValueError(f"{k}: expected a contiguous complex128 CUDA tensor")
            if tuple(t.shape) != tuple(p.shape if k == "lnps" else v.shape):
                raise ValueError(f"{k}: shape does not match the state")
        return v.shape[0] if v.dim() == 3 else 1

    @staticmethod
    def flags(semi_implicit, diffuse, eps_filter):
        return ((_lib.AER_STEP_SEMI_IMPLICIT if semi_implicit else 0) |
                (_lib.AER_STEP_DIFFUSE if diffuse else 0) |
                (_lib.AER_STEP_FILTER if eps_filter > 0 else 0))

    def leapfrog(self, old, now, tnd, semi_implicit=True, diffuse=False, eps_filter=0.0, raw_alpha=1.0):
        """Sections 2-4 of aeros_timestep_step; `now` is filtered in place. Returns X^(n+1).
        An ensemble of members of this size passes arrays with a leading member dimension."""
        import torch
        nmem = self._members(now)
        for d in (old, tnd):
            if self._members(d) != nmem:
                raise ValueError("old, now and tnd must hold the same members")
        new = {k: torch.empty_like(now[k]) for k in KEYS}
        self.sht._bind_stream()
        check(lib().aer_leapfrog_batch_d(self._h, nmem, self.flags(semi_implicit, diffuse, eps_filter), eps_filter,
                                         raw_alpha, *_ptrs(old, KEYS), *_ptrs(now, KEYS), *_ptrs(tnd, KEYS),
                                         *_ptrs(new, KEYS)))
        return new

    def leapfrog_adjoint(self, bn, bc, semi_implicit=True, diffuse=False, eps_filter=0.0, raw_alpha=1.0):
        """Transpose of leapfrog at this object's step settings: from the adjoints of X^(n+1)
        (bn) and of the filtered X^n (bc) to those of X^(n-1), X^n and the tendency."""
        import torch
        nmem = self._members(bn)
        if self._members(bc) != nmem:
            raise ValueError("bn and bc must hold the same members")
        go, gc, gt = ({k: torch.empty_like(bn[k]) for k in KEYS} for _ in range(3))
        self.sht._bind_stream()
        check(lib().aer_leapfrog_adj_batch_d(self._h, nmem, self.flags(semi_implicit, diffuse, eps_filter),
                                             eps_filter, raw_alpha, *_ptrs(bn, KEYS), *_ptrs(bc, KEYS),
                                             *_ptrs(go, KEYS), *_ptrs(gc, KEYS), *_ptrs(gt, KEYS)))
        return go, gc, gt


def mass(sht, lnps, fix=False, target=0.0):
    """Global dry-air mass of ln p_s on the device; with fix, the mass fixer applied
    to lnps in place. Returns (mass, ln r) as 0-d float64 CUDA tensors. An ensemble
    passes lnps (nmem, nlm) and, with fix, the targets as a (nmem,) float64 CUDA tensor;
    it gets (nmem,) tensors back and the host never waits."""
    import torch
    if not (lnps.is_cuda and lnps.is_contiguous() and lnps.dtype == torch.complex128 and lnps.shape[-1] == sht.nlm):
        raise ValueError("lnps: expected a contiguous complex128 CUDA tensor ([nmem,] nlm)")
    if lnps.dim() == 1:
        m = torch.zeros((), dtype=torch.float64, device=lnps.device)
        r = torch.zeros((), dtype=torch.float64, device=lnps.device)
        sht._bind_stream()
        check(lib().aer_mass_d(sht._h, lnps.data_ptr(), 1 if fix else 0, float(target), m.data_ptr(), r.data_ptr()))
        return m, r
    nmem = lnps.shape[0]
    m = torch.zeros((nmem,), dtype=torch.float64, device=lnps.device)
    r = torch.zeros((nmem,), dtype=torch.float64, device=lnps.device)
    tg = None
    if fix:
        if not (torch.is_tensor(target) and target.is_cuda and target.dtype == torch.float64
                and tuple(target.shape) == (nmem,) and target.is_contiguous()):
            raise ValueError("an ensemble's mass targets must be a contiguous (nmem,) float64 CUDA tensor")
        tg = target.data_ptr()
    sht._bind_stream()
    check(lib().aer_mass_batch_d(sht._h, nmem, lnps.data_ptr(), 1 if fix else 0, tg, m.data_ptr(), r.data_ptr()))
    return m, r
