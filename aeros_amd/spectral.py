"""Python mirror of the reference's aeros_spectral module over the C ABI.

`AerosSht` carries the same fields as the reference's aeros_sht_class
(src/aeros_spectral.f90:47-85) and the same operations (:128-143):
analysis, synthesis, analysis_vec, synthesis_vec, gradient, laplacian,
surface_integral, lm, plus aeros_vordiv's mapping (src/dynamics/aeros_vordiv.f90)
and the tendency transform halves (src/dynamics/aeros_tendency.f90:403-470).

Two kinds of arrays are accepted:
  * numpy arrays on the host: one field per call, through the per-call entry
    points (aer_SH_to_spat, ...), exactly like the reference's per-level calls;
  * torch CUDA tensors: any number of leading batch dimensions (levels,
    fields), one launch per stage, through the *_d entry points.

Shapes are C-order views of the reference's Fortran arrays: a grid field is
(nlat, nlon) (longitude contiguous, north -> south), a spectral field is
(nlm,) complex128 (m-major, l fastest). Analyses leave their inputs intact.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

R_EARTH = 6.371e6  # reference src/aeros_defs.f90:84

PRECISIONS = {"fp64": _lib.AER_PREC_FP64, "fp32": _lib.AER_PREC_FP32, "split": _lib.AER_PREC_SPLIT}


def grid_size(trunc):
    """Smallest quadratically unaliased grid (src/aeros_spectral.f90:399-438)."""
    a, b = ctypes.c_int(), ctypes.c_int()
    check(lib().aer_sht_grid_size(trunc, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class AerosSht:
    """One transform configuration (truncation + Gaussian grid) on the current GPU."""

    def __init__(self, trunc, nlon=0, nlat=0, precision="fp64"):
        dlon, dlat = grid_size(trunc)
        nlon = nlon if nlon > 0 else dlon
        nlat = nlat if nlat > 0 else dlat
        h = ctypes.c_void_p()
        norm = _lib.AER_SHT_ORTHONORMAL | _lib.AER_SHT_NO_CS_PHASE
        check(lib().aer_sht_create(trunc, trunc, 1, norm, ctypes.byref(h)))
        self._h = h
        layout = _lib.AER_SHT_PHI_CONTIGUOUS | _lib.AER_SHT_QUICK_INIT
        check(lib().aer_sht_set_grid(h, layout, 1.0e-10, nlat, nlon))
        info = _lib.ShtInfo()
        check(lib().aer_sht_get_info(h, ctypes.byref(info)))
        self.trunc = self.lmax = self.mmax = trunc
        self.mres = 1
        self.nlon, self.nlat, self.nlm = info.nphi, info.nlat, info.nlm
        li = (ctypes.c_int * self.nlm)()
        mi = (ctypes.c_int * self.nlm)()
        check(lib().aer_sht_get_lm(h, li, mi))
        self.l_of_lm = np.array(li[:], dtype=np.int64)
        self.m_of_lm = np.array(mi[:], dtype=np.int64)
        ct = np.empty(self.nlat)
        check(lib().aer_sht_get_cos_theta(h, ct.ctypes.data_as(_lib._dp)))
        wh = np.empty(info.nlat_2)
        check(lib().aer_sht_gauss_wts(h, wh.ctypes.data_as(_lib._dp)))
        self.sinlat = ct
        self.colat = np.arccos(ct)
        self.lon = 2.0 * np.pi * np.arange(self.nlon) / self.nlon
        w = np.empty(self.nlat)
        nh = info.nlat_2
        w[:nh] = wh
        w[self.nlat - nh:][::-1] = wh
        self.gauss_w = w
        self.set_precision(precision)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().aer_sht_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ settings
    def set_precision(self, precision):
        self.precision = precision
        check(lib().aer_sht_set_precision(self._h, PRECISIONS[precision]))

    def set_path(self, path):
        """Kernel path: "tiled" (fused synthesis; default) or "staged" (HBM intermediate)."""
        self.path = path
        check(lib().aer_sht_set_path(self._h, {"tiled": _lib.AER_PATH_TILED, "staged": _lib.AER_PATH_STAGED}[path]))

    def _bind_stream(self):
        import torch
        check(lib().aer_sht_set_stream(self._h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def lm(self, l, m):
        """1-based index of (l, m), -1 outside the truncation (src/aeros_spectral.f90:460-482)."""
        if m < 0 or m > self.mmax or l < m or l > self.lmax:
            return -1
        return (m * (2 * self.lmax + 2 - (m + 1))) // 2 + l + 1

    # ------------------------------------------------------------ helpers
    def _spec_np(self, q):
        q = np.ascontiguousarray(q, dtype=np.complex128)
        if q.shape != (self.nlm,):
            raise ValueError(f"spectral field must be ({self.nlm},), got {q.shape}")
        return q

    def _grid_np(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        if g.shape != (self.nlat, self.nlon):
            raise ValueError(f"grid field must be ({self.nlat}, {self.nlon}), got {g.shape}")
        return g

    def _dev_spec(self, q):
        import torch
        if q.dtype != torch.complex128 or q.shape[-1] != self.nlm or not q.is_contiguous() or not q.is_cuda:
            raise ValueError("device spectral arrays must be contiguous cuda complex128 (..., nlm)")
        return q

    def _dev_grid(self, g):
        import torch
        if g.dtype != torch.float64 or tuple(g.shape[-2:]) != (self.nlat, self.nlon) or not g.is_contiguous() or not g.is_cuda:
            raise ValueError("device grid arrays must be contiguous cuda float64 (..., nlat, nlon)")
        return g

    def _new_grid(self, like, batch):
        import torch
        return torch.empty(tuple(batch) + (self.nlat, self.nlon), dtype=torch.float64, device=like.device)

    def _new_spec(self, like, batch):
        import torch
        return torch.empty(tuple(batch) + (self.nlm,), dtype=torch.complex128, device=like.device)

    @staticmethod
    def _n(t, trailing):
        n = 1
        for s in t.shape[:t.dim() - trailing]:
            n *= s
        return n

    # ------------------------------------------------------------ scalar
    def synthesis(self, q):
        """Spectral (nlm) -> grid (nlat, nlon) (SH_to_spat, src/aeros_spectral.f90:502-515)."""
        if _is_torch(q):
            q = self._dev_spec(q)
            out = self._new_grid(q, q.shape[:-1])
            self._bind_stream()
            check(lib().aer_synth_d(self._h, self._n(q, 1), q.data_ptr(), 2 * self.nlm,
                                    out.data_ptr(), self.nlon * self.nlat))
            return out
        q = self._spec_np(q)
        g = np.empty((self.nlat, self.nlon))
        check(lib().aer_SH_to_spat(self._h, q.ctypes.data, g.ctypes.data))
        return g

    def analysis(self, g):
        """Grid -> spectral (spat_to_SH, src/aeros_spectral.f90:486-500); input preserved."""
        if _is_torch(g):
            g = self._dev_grid(g)
            out = self._new_spec(g, g.shape[:-2])
            self._bind_stream()
            check(lib().aer_anal_d(self._h, self._n(g, 2), g.data_ptr(), self.nlon * self.nlat,
                                   out.data_ptr(), 2 * self.nlm))
            return out
        g = self._grid_np(g)
        q = np.empty(self.nlm, dtype=np.complex128)
        check(lib().aer_spat_to_SH(self._h, g.ctypes.data, q.ctypes.data))
        return q

    # ------------------------------------------------------------ vector
    def synthesis_vec(self, s, t):
        """(S, T) -> (V_theta, V_phi) (SHsphtor_to_spat, src/aeros_spectral.f90:545-558)."""
        if _is_torch(s):
            s, t = self._dev_spec(s), self._dev_spec(t)
            vt, vp = self._new_grid(s, s.shape[:-1]), self._new_grid(s, s.shape[:-1])
            self._bind_stream()
            check(lib().aer_synth_vec_d(self._h, self._n(s, 1), s.data_ptr(), t.data_ptr(), 2 * self.nlm,
                                        vt.data_ptr(), vp.data_ptr(), self.nlon * self.nlat))
            return vt, vp
        s, t = self._spec_np(s), self._spec_np(t)
        vt, vp = np.empty((self.nlat, self.nlon)), np.empty((self.nlat, self.nlon))
        check(lib().aer_SHsphtor_to_spat(self._h, s.ctypes.data, t.ctypes.data, vt.ctypes.data, vp.ctypes.data))
        return vt, vp

    def analysis_vec(self, vt, vp):
        """(V_theta, V_phi) -> (S, T) (spat_to_SHsphtor, src/aeros_spectral.f90:529-543); inputs preserved."""
        if _is_torch(vt):
            vt, vp = self._dev_grid(vt), self._dev_grid(vp)
            s, t = self._new_spec(vt, vt.shape[:-2]), self._new_spec(vt, vt.shape[:-2])
            self._bind_stream()
            check(lib().aer_anal_vec_d(self._h, self._n(vt, 2), vt.data_ptr(), vp.data_ptr(),
                                       self.nlon * self.nlat, s.data_ptr(), t.data_ptr(), 2 * self.nlm))
            return s, t
        vt, vp = self._grid_np(vt), self._grid_np(vp)
        s, t = np.empty(self.nlm, dtype=np.complex128), np.empty(self.nlm, dtype=np.complex128)
        check(lib().aer_spat_to_SHsphtor(self._h, vt.ctypes.data, vp.ctypes.data, s.ctypes.data, t.ctypes.data))
        return s, t

    # ------------------------------------------------------------ adjoints (device tensors only)
    def synthesis_adjoint(self, g):
        """synthesis^T: grid (..., nlat, nlon) -> spectral (..., nlm)."""
        g = self._dev_grid(g)
        out = self._new_spec(g, g.shape[:-2])
        self._bind_stream()
        check(lib().aer_synth_adj_d(self._h, self._n(g, 2), g.data_ptr(), self.nlon * self.nlat,
                                    out.data_ptr(), 2 * self.nlm))
        return out

    def analysis_adjoint(self, q):
        """analysis^T: spectral (..., nlm) -> grid (..., nlat, nlon)."""
        q = self._dev_spec(q)
        out = self._new_grid(q, q.shape[:-1])
        self._bind_stream()
        check(lib().aer_anal_adj_d(self._h, self._n(q, 1), q.data_ptr(), 2 * self.nlm,
                                   out.data_ptr(), self.nlon * self.nlat))
        return out

    def synthesis_vec_adjoint(self, vt, vp):
        vt, vp = self._dev_grid(vt), self._dev_grid(vp)
        s, t = self._new_spec(vt, vt.shape[:-2]), self._new_spec(vt, vt.shape[:-2])
        self._bind_stream()
        check(lib().aer_synth_vec_adj_d(self._h, self._n(vt, 2), vt.data_ptr(), vp.data_ptr(),
                                        self.nlon * self.nlat, s.data_ptr(), t.data_ptr(), 2 * self.nlm))
        return s, t

    def analysis_vec_adjoint(self, s, t):
        s, t = self._dev_spec(s), self._dev_spec(t)
        vt, vp = self._new_grid(s, s.shape[:-1]), self._new_grid(s, s.shape[:-1])
        self._bind_stream()
        check(lib().aer_anal_vec_adj_d(self._h, self._n(s, 1), s.data_ptr(), t.data_ptr(), 2 * self.nlm,
                                       vt.data_ptr(), vp.data_ptr(), self.nlon * self.nlat))
        return vt, vp

    def gradient_adjoint(self, gt, gp):
        gt, gp = self._dev_grid(gt), self._dev_grid(gp)
        s = self._new_spec(gt, gt.shape[:-2])
        self._bind_stream()
        check(lib().aer_synth_grad_adj_d(self._h, self._n(gt, 2), gt.data_ptr(), gp.data_ptr(),
                                         self.nlon * self.nlat, s.data_ptr(), 2 * self.nlm))
        return s

    def gradient(self, q):
        """Unit-sphere (dF/dtheta, (1/sin) dF/dphi) (SHsph_to_spat, src/aeros_spectral.f90:560-584)."""
        if _is_torch(q):
            q = self._dev_spec(q)
            gt, gp = self._new_grid(q, q.shape[:-1]), self._new_grid(q, q.shape[:-1])
            self._bind_stream()
            check(lib().aer_synth_grad_d(self._h, self._n(q, 1), q.data_ptr(), 2 * self.nlm,
                                         gt.data_ptr(), gp.data_ptr(), self.nlon * self.nlat))
            return gt, gp
        q = self._spec_np(q)
        gt, gp = np.empty((self.nlat, self.nlon)), np.empty((self.nlat, self.nlon))
        check(lib().aer_SHsph_to_spat(self._h, q.ctypes.data, gt.ctypes.data, gp.ctypes.data))
        return gt, gp

    # ------------------------------------------------------------ operators (exact, host)
    def laplacian(self, q, inverse=False):
        """c * (-l(l+1)/a^2), or its inverse with l = 0 -> 0 (src/aeros_spectral.f90:588-630)."""
        q = np.array(q, dtype=np.complex128, copy=True)
        a2 = R_EARTH * R_EARTH
        for lm in range(self.nlm):
            l = int(self.l_of_lm[lm])
            if inverse:
                if l == 0:
                    q[lm] = 0.0
                    continue
                fac = -a2 / float(l * (l + 1))
            else:
                fac = -float(l * (l + 1)) / a2
            q[lm] = q[lm] * fac
        return q

    def surface_integral(self, g):
        """a^2 sum_j w_j (2pi/nlon) sum_i f_ij, row sums first (src/aeros_spectral.f90:632-666)."""
        g = self._grid_np(g)
        dlon = 2.0 * np.pi / self.nlon
        total = 0.0
        for j in range(self.nlat):
            row = 0.0
            for i in range(self.nlon):
                row += g[j, i]
            total += self.gauss_w[j] * row * dlon
        return total * R_EARTH * R_EARTH

    # ------------------------------------------------------------ aeros_vordiv (device)
    def uv_from_vordiv(self, vor, div):
        """(zeta, D) -> (u, v), src/dynamics/aeros_vordiv.f90:82-122."""
        vor, div = self._dev_spec(vor), self._dev_spec(div)
        u, v = self._new_grid(vor, vor.shape[:-1]), self._new_grid(vor, vor.shape[:-1])
        self._bind_stream()
        check(lib().aer_uv_from_vordiv_d(self._h, self._n(vor, 1), vor.data_ptr(), div.data_ptr(),
                                         u.data_ptr(), v.data_ptr()))
        return u, v

    def vordiv_from_uv(self, u, v):
        """(u, v) -> (zeta, D), src/dynamics/aeros_vordiv.f90:124-168."""
        u, v = self._dev_grid(u), self._dev_grid(v)
        vor, div = self._new_spec(u, u.shape[:-2]), self._new_spec(u, u.shape[:-2])
        self._bind_stream()
        check(lib().aer_vordiv_from_uv_d(self._h, self._n(u, 2), u.data_ptr(), v.data_ptr(),
                                         vor.data_ptr(), div.data_ptr()))
        return vor, div

    def grad_phys(self, f):
        """(dfdx east, dfdy north), src/dynamics/aeros_vordiv.f90:170-205."""
        f = self._dev_spec(f)
        dx, dy = self._new_grid(f, f.shape[:-1]), self._new_grid(f, f.shape[:-1])
        self._bind_stream()
        check(lib().aer_gradient_d(self._h, self._n(f, 1), f.data_ptr(), dx.data_ptr(), dy.data_ptr()))
        return dx, dy

    # ------------------------------------------------------------ tendency halves (device)
    def tendency_grid_transforms(self, vor, div, temp, lnps, out=None):
        """Spectral -> grid half of aeros_tendency_grid (src/dynamics/aeros_tendency.f90:403-419).
        vor/div/temp: (nlev, nlm); lnps: (nlm,). Returns a dict of grid tensors. An ensemble
        of members of this size passes (nmem, nlev, nlm) and (nmem, nlm): one fused call."""
        lead = tuple(vor.shape[:-2])
        if len(lead) > 1 or tuple(lnps.shape) != lead + (self.nlm,):
            raise ValueError("vor/div/temp must be ([nmem,] nlev, nlm) and lnps ([nmem,] nlm)")
        nmem, nlev = (lead[0] if lead else 1), vor.shape[-2]
        for a in (vor, div, temp):
            self._dev_spec(a)
            if a.shape != vor.shape:
                raise ValueError("vor, div and temp must have the same shape")
        self._dev_spec(lnps)
        if out is None:
            out = {k: self._new_grid(vor, lead + (nlev,)) for k in ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy")}
            out.update({k: self._new_grid(vor, lead) for k in ("lnps_g", "dlnpsdx", "dlnpsdy")})
        self._bind_stream()
        check(lib().aer_tendency_grid_transforms_batch_d(
            self._h, nmem, nlev, vor.data_ptr(), div.data_ptr(), temp.data_ptr(), lnps.data_ptr(),
            *[out[k].data_ptr() for k in ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy",
                                          "lnps_g", "dlnpsdx", "dlnpsdy")]))
        return out

    def tendency_spectral(self, ae, an, ek, dtdt, dlnpsdt, out=None):
        """aeros_tendency_spectral (src/dynamics/aeros_tendency.f90:428-473); inputs preserved.
        ae/an/ek/dtdt: ([nmem,] nlev, nlat, nlon); dlnpsdt ([nmem,] nlat, nlon)."""
        lead = tuple(ae.shape[:-3])
        if len(lead) > 1 or tuple(dlnpsdt.shape) != lead + (self.nlat, self.nlon):
            raise ValueError("ae/an/ek/dtdt must be ([nmem,] nlev, nlat, nlon) and dlnpsdt ([nmem,] nlat, nlon)")
        nmem, nlev = (lead[0] if lead else 1), ae.shape[-3]
        for a in (ae, an, ek, dtdt):
            self._dev_grid(a)
            if a.shape != ae.shape:
                raise ValueError("ae, an, ek and dtdt must have the same shape")
        self._dev_grid(dlnpsdt)
        if out is None:
            out = {k: self._new_spec(ae, lead + (nlev,)) for k in ("vor", "div", "temp")}
            out["lnps"] = self._new_spec(ae, lead)
        self._bind_stream()
        check(lib().aer_tendency_spectral_batch_d(
            self._h, nmem, nlev, ae.data_ptr(), an.data_ptr(), ek.data_ptr(), dtdt.data_ptr(), dlnpsdt.data_ptr(),
            out["vor"].data_ptr(), out["div"].data_ptr(), out["temp"].data_ptr(), out["lnps"].data_ptr()))
        return out

    def sync(self):
        check(lib().aer_sht_sync(self._h))
