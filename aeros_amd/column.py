"""Python mirror of the grid-space column work of an aeros dynamics step over
the C ABI: the hybrid vertical coordinate (reference aeros_vgrid_class,
src/dynamics/aeros_vertical.f90:46-55), the Simmons-Burridge column terms
(src/dynamics/aeros_tendency.f90:475-630) and the Held-Suarez forcing
(src/physics/aeros_held_suarez.f90:177-223), run on the GPU in one kernel.

Arrays are torch CUDA tensors in the layouts of the transform halves: 3-D grid
fields (nlev, nlat, nlon), 2-D (nlat, nlon), float64.
"""
import ctypes

import numpy as np

from ._lib import check, lib


def vgrid_init(nlev, stretch_a=0.4, stretch_r=2.0, sigma_t=0.2, p_top=1000.0, ps_ref=1.0e5):
    """Hybrid coefficients (A [Pa], B [-]) at the nlev+1 half levels, top first, from
    the stretching parameters of aeros_vgrid_init (src/dynamics/aeros_vertical.f90:126-298,
    the generator branch): level placement sigma = 1 - (a (1-x) + (1-a)(1-x)^r),
    x = j/nlev, squeezed onto [p_top/ps_ref, 1]; B = 0 above sigma_t, ramping linearly to
    1 at the surface (pure sigma for sigma_t = 0); A = (sigma - B) ps_ref; endpoints pinned.
    The defaults are the reference's production grid; Held-Suarez uses stretch_a = 1,
    sigma_t = 0 (uniform sigma)."""
    if nlev < 2 or not (0.0 < stretch_a <= 1.0) or not (0.0 <= sigma_t < 1.0):
        raise ValueError("nlev >= 2, stretch_a in (0, 1] and sigma_t in [0, 1) are required")
    A, B = np.zeros(nlev + 1), np.zeros(nlev + 1)
    sig_top = p_top / ps_ref
    for j in range(nlev + 1):
        x = j / nlev
        sh = sig_top + (1.0 - sig_top) * (1.0 - (stretch_a * (1.0 - x) + (1.0 - stretch_a) * (1.0 - x) ** stretch_r))
        if sigma_t <= 0.0:
            b = sh
        elif sh <= sigma_t:
            b = 0.0
        else:
            b = (sh - sigma_t) / (1.0 - sigma_t)
        A[j], B[j] = (sh - b) * ps_ref, b
    A[nlev], B[nlev], A[0], B[0] = 0.0, 1.0, p_top, 0.0
    return A, B


class VGrid:
    """Device copy of the hybrid coefficients A (Pa), B (-) at nlev+1 half levels, top first."""

    def __init__(self, a_half, b_half):
        a = np.ascontiguousarray(a_half, dtype=np.float64)
        b = np.ascontiguousarray(b_half, dtype=np.float64)
        if a.shape != b.shape or a.ndim != 1:
            raise ValueError("a_half and b_half must be 1-D of equal length nlev+1")
        self.nlev = a.size - 1
        self.a_half, self.b_half = a.copy(), b.copy()
        h = ctypes.c_void_p()
        dp = ctypes.POINTER(ctypes.c_double)
        check(lib().aer_vgrid_create(self.nlev, a.ctypes.data_as(dp), b.ctypes.data_as(dp), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().aer_vgrid_destroy(h)
            except Exception:
                pass
            self._h = None


GKEYS = ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy", "lnps_g", "dlnpsdx", "dlnpsdy")
CKEYS = ("ae", "an", "ek", "dtdt", "dlnpsdt")


def _members(sht, vg, grid, phis, dteq, held_suarez):
    """(nmem, dteq pointer, dteq member stride) after checking the grid dict's shapes:
    3-D fields ([nmem,] nlev, nlat, nlon), 2-D ([nmem,] nlat, nlon); dteq (nlev, nlat)
    shared by every member or (nmem, nlev, nlat)."""
    import torch
    u = grid["u"]
    lead = tuple(u.shape[:-3])
    if len(lead) > 1 or u.shape[-3:] != (vg.nlev, sht.nlat, sht.nlon):
        raise ValueError("grid fields do not match the vertical grid / transform grid")
    for k in GKEYS:
        t = grid[k]
        want = u.shape if k in GKEYS[:7] else lead + (sht.nlat, sht.nlon)
        if tuple(t.shape) != tuple(want):
            raise ValueError(f"{k}: expected shape {tuple(want)}")
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError(f"{k}: expected a contiguous float64 CUDA tensor")
    if phis is not None and not (phis.is_cuda and phis.dtype == torch.float64 and phis.is_contiguous()
                                 and tuple(phis.shape) == (sht.nlat, sht.nlon)):
        raise ValueError("phis: expected a contiguous float64 CUDA tensor (nlat, nlon)")
    nmem = lead[0] if lead else 1
    if dteq is None:
        return nmem, None, 0
    # Held-Suarez with T_eq = formula + dteq (the adjoint's control variable)
    if not held_suarez:
        raise ValueError("dteq needs the Held-Suarez forcing")
    if not (dteq.is_cuda and dteq.dtype == torch.float64 and dteq.is_contiguous()):
        raise ValueError("dteq: expected a contiguous float64 CUDA tensor")
    if tuple(dteq.shape) == (vg.nlev, sht.nlat):
        return nmem, dteq.data_ptr(), 0
    if lead and tuple(dteq.shape) == (nmem, vg.nlev, sht.nlat):
        return nmem, dteq.data_ptr(), vg.nlev * sht.nlat
    raise ValueError("dteq: expected shape (nlev, nlat) or (nmem, nlev, nlat)")


def column_terms(sht, vg, grid, phis=None, held_suarez=False, out=None, dteq=None):
    """Grid right-hand sides of every column from the dict that
    AerosSht.tendency_grid_transforms returns. Returns ae, an, ek, dtdt (3-D) and
    dlnpsdt (2-D) as a dict; inputs are left unchanged. An ensemble passes fields with a
    leading member dimension (one launch for all members)."""
    import torch
    nmem, dq, dqs = _members(sht, vg, grid, phis, dteq, held_suarez)
    u = grid["u"]
    if out is None:
        out = {k: torch.empty_like(u) for k in ("ae", "an", "ek", "dtdt")}
        out["dlnpsdt"] = torch.empty_like(grid["lnps_g"])
    sht._bind_stream()
    check(lib().aer_column_terms_batch_d(
        sht._h, vg._h, nmem, 1 if held_suarez else 0, dq, dqs, *[grid[k].data_ptr() for k in GKEYS],
        phis.data_ptr() if phis is not None else None, *[out[k].data_ptr() for k in CKEYS]))
    return out


def column_terms_adjoint(sht, vg, grid, cbar, phis=None, held_suarez=False, dteq=None, want_teq=False):
    """Adjoint of column_terms at the inputs `grid`: from the adjoints of ae, an, ek, dtdt,
    dlnpsdt (dict cbar) to those of the ten grid inputs, and, with want_teq, the sensitivity
    to dteq per (level, latitude) (summed over longitudes; for an ensemble per member,
    or summed over members when dteq is one shared field)."""
    import torch
    nmem, dq, dqs = _members(sht, vg, grid, phis, dteq, held_suarez)
    g = {k: torch.empty_like(grid[k]) for k in GKEYS}
    cb = {k: cbar[k].contiguous() for k in CKEYS}
    gteq = torch.empty_like(grid["u"]) if want_teq else None
    sht._bind_stream()
    check(lib().aer_column_terms_adj_batch_d(
        sht._h, vg._h, nmem, 1 if held_suarez else 0, dq, dqs, *[grid[k].data_ptr() for k in GKEYS],
        phis.data_ptr() if phis is not None else None,
        *[cb[k].data_ptr() for k in CKEYS],
        *[g[k].data_ptr() for k in GKEYS], gteq.data_ptr() if gteq is not None else None))
    if not want_teq:
        return g
    gq = gteq.sum(dim=-1)
    if dteq is not None and dteq.dim() == 2 and gq.dim() == 3:
        gq = gq.sum(dim=0)
    return g, gq
