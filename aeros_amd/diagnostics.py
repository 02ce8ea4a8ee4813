"""Online zonal-mean statistics on the GPU: the reference's aeros_diag_class
(src/aeros_diagnostics.f90) over the C ABI. Sums are FP64 on the device
(aer_diag_*); finalize divides them through and builds the pressure axis at
the time-mean zonal-mean surface pressure, as aeros_diag_finalize does
(src/aeros_diagnostics.f90:235-276); summary() reduces them to the scalars of
aeros_diag_summary (src/aeros_diagnostics.f90:278-372).
"""
import ctypes

import numpy as np

from ._lib import check, lib

FIELDS = ("u", "v", "temp", "uv", "vt", "eke", "tvar")


class ZonalStats:
    """aeros_diag_class for one transform handle and vertical grid size."""

    def __init__(self, sht, nlev):
        self.sht, self.nlev = sht, nlev
        h = ctypes.c_void_p()
        check(lib().aer_diag_create(sht._h, nlev, ctypes.byref(h)))
        self._h = h
        self.time_start = self.time_end = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().aer_diag_destroy(h)
            except Exception:
                pass
            self._h = None

    def accumulate(self, grid, time=0.0):
        """One sample from grid fields u, v, t_g (nlev, nlat, nlon) and lnps_g (nlat, nlon):
        the dict AerosSht.tendency_grid_transforms returns (src/aeros_diagnostics.f90:153-233)."""
        for k in ("u", "v", "t_g"):
            if tuple(grid[k].shape) != (self.nlev, self.sht.nlat, self.sht.nlon):
                raise ValueError(f"{k}: expected shape (nlev, nlat, nlon)")
        if self.time_start is None:
            self.time_start = time
        self.time_end = time
        self.sht._bind_stream()
        check(lib().aer_diag_accumulate_d(self._h, grid["u"].data_ptr(), grid["v"].data_ptr(),
                                          grid["t_g"].data_ptr(), grid["lnps_g"].data_ptr()))

    def reset(self):
        self.sht._bind_stream()
        check(lib().aer_diag_reset(self._h))
        self.time_start = self.time_end = None

    def finalize(self, a_half, b_half):
        """Time means as (nlat, nlev) arrays (the reference's index order), ps and u_sfc
        (nlat,), and the full-level pressure at the mean [p_s] of each latitude."""
        nlat, nlev = self.sht.nlat, self.nlev
        buf = np.empty((7 * nlev + 2) * nlat)
        n = ctypes.c_int()
        check(lib().aer_diag_read(self._h, ctypes.byref(n), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        if n.value < 1:
            raise RuntimeError("no samples were accumulated")
        out = {"nsample": n.value, "time_start": self.time_start, "time_end": self.time_end}
        f = buf[: 7 * nlev * nlat].reshape(7, nlev, nlat)
        for i, name in enumerate(FIELDS):
            out[name] = f[i].T.copy()
        out["ps"] = buf[7 * nlev * nlat: 7 * nlev * nlat + nlat].copy()
        out["u_sfc"] = buf[7 * nlev * nlat + nlat:].copy()
        A, B = np.asarray(a_half), np.asarray(b_half)
        ph = A[None, :] + B[None, :] * out["ps"][:, None]
        out["pressure"] = 0.5 * (ph[:, 1:] + ph[:, :-1])   # full level = mean of its half levels
        return out


def summary(d, lat_deg):
    """The scalars of aeros_diag_summary (src/aeros_diagnostics.f90:278-372). lat_deg is the
    grid latitude, north to south."""
    u = d["u"]
    jmax, kmax = np.unravel_index(np.argmax(u.T), u.T.shape)[::-1]
    jm2 = int(np.argmax(d["u_sfc"]))
    vt = d["vt"]
    jv, kv = np.unravel_index(np.argmax(vt.T), vt.T.shape)[::-1]
    nlat, nlev = u.shape
    num = float(((u - u[::-1, :]) ** 2).sum())
    den = float((u ** 2).sum())
    return {"u_max": float(u[jmax, kmax]), "u_max_lat": float(lat_deg[jmax]),
            "u_max_hpa": float(d["pressure"][jmax, kmax] / 100.0),
            "u_sfc_max": float(d["u_sfc"][jm2]), "u_sfc_max_lat": float(lat_deg[jm2]),
            "vt_max": float(vt[jv, kv]), "vt_max_lat": float(lat_deg[jv]),
            "uv_max": float(d["uv"].max()), "eke_max": float(d["eke"].max()),
            "t_sfc_equator": float(d["temp"][nlat // 2 - 1, nlev - 1]),
            "t_sfc_pole": float(d["temp"][0, nlev - 1]),
            "u_asymmetry": float(np.sqrt(num / max(den, np.finfo(float).tiny)))}
