"""numpy front end of the FP64 CPU oracle (oracle/sht_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py. The product path (aeros_amd) never imports it.

Arrays: spectral fields are complex128 of shape (nlm,), m-major, l fastest;
grid fields are float64 of shape (nlat, nlon), i.e. the reference's Fortran
(nlon, nlat) column-major layout seen from C order (longitude contiguous,
latitude north -> south).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liborc.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _Cfg(ctypes.Structure):
    _fields_ = [("trunc", ctypes.c_int), ("lmax", ctypes.c_int), ("mmax", ctypes.c_int),
                ("nlon", ctypes.c_int), ("nlat", ctypes.c_int), ("nlm", ctypes.c_int),
                ("mu", _dp), ("w", _dp), ("sint", _dp), ("P", _dp), ("D", _dp),
                ("l_of_lm", _ip), ("m_of_lm", _ip), ("fft_direct", ctypes.c_int),
                ("tw_fwd", _dp), ("tw_inv", _dp)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            from oracle import build_oracle
            build_oracle.build()
        L = ctypes.CDLL(_LIBPATH)
        cp = ctypes.POINTER(_Cfg)
        L.orc_create.restype = cp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [cp]
        L.orc_set_fft_direct.argtypes = [cp, ctypes.c_int]
        L.orc_grid_size.argtypes = [ctypes.c_int, _ip, _ip]
        L.orc_lm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_gauss.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_legendre.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_rfft_forward.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_rfft_inverse.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        for name, n in [("orc_synth", 2), ("orc_anal", 2), ("orc_synth_grad", 3),
                        ("orc_synth_vec", 4), ("orc_anal_vec", 4),
                        ("orc_uv_from_vordiv", 4), ("orc_vordiv_from_uv", 4),
                        ("orc_gradient", 3)]:
            getattr(L, name).argtypes = [cp] + [_dp] * n
        L.orc_laplacian.argtypes = [cp, _dp, ctypes.c_int]
        L.orc_surface_integral.argtypes = [cp, _dp]
        L.orc_surface_integral.restype = ctypes.c_double
        L.orc_tendency_grid_transforms.argtypes = [cp, ctypes.c_int] + [_dp] * 14
        L.orc_tendency_spectral.argtypes = [cp, ctypes.c_int] + [_dp] * 9
        L.orc_bench_set.argtypes = [cp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_bench_set.restype = ctypes.c_double
        L.orc_port_create.restype = ctypes.c_void_p
        L.orc_port_create.argtypes = [cp]
        L.orc_port_destroy.argtypes = [ctypes.c_void_p]
        L.orc_port_work_doubles.restype = ctypes.c_size_t
        L.orc_port_work_doubles.argtypes = [ctypes.c_void_p]
        L.orc_port_synth.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_port_anal.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_port_cfft.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_port_cfft.restype = ctypes.c_int
        L.orc_port_bench_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_port_bench_set.restype = ctypes.c_double
        d = ctypes.c_double
        L.orc_vgrid_init.argtypes = [ctypes.c_int, d, d, d, d, d, _dp, _dp]
        L.orc_vgrid_init.restype = ctypes.c_int
        L.orc_vgrid_pressure.argtypes = [ctypes.c_int, _dp, _dp, d, _dp, _dp, _dp]
        L.orc_vgrid_alpha.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_hydrostatic.argtypes = [ctypes.c_int, d, _dp, _dp, _dp]
        for name, n in [("orc_hs_teq", 3), ("orc_hs_kt", 2), ("orc_hs_kv", 1)]:
            getattr(L, name).argtypes = [d] * n
            getattr(L, name).restype = d
        L.orc_column_terms.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [_dp] * 11 + [_dp] * 5 + [ctypes.c_int]
        ci = ctypes.c_int
        L.orc_si_create.restype = ctypes.c_void_p
        L.orc_si_create.argtypes = [ci, _dp, _dp, d, d, ci, d, d]
        L.orc_si_destroy.argtypes = [ctypes.c_void_p]
        L.orc_si_set_step.argtypes = [ctypes.c_void_p, d]
        L.orc_si_set_step.restype = ci
        L.orc_si_linear.argtypes = [ctypes.c_void_p, ci, _ip] + [_dp] * 6
        L.orc_si_step.argtypes = [ctypes.c_void_p, ci, _ip] + [_dp] * 12
        L.orc_si_gwspeed.argtypes = [ctypes.c_void_p]
        L.orc_si_gwspeed.restype = d
        L.orc_diff_factors.argtypes = [ci, d, d, ci, _dp]
        L.orc_diffuse.argtypes = [ci, ci, ci, _ip, _dp, _dp, _dp, _dp, _dp]
        L.orc_raw_filter.argtypes = [ctypes.c_long, d, d, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def grid_size(trunc):
    a, b = ctypes.c_int(), ctypes.c_int()
    rc = lib().orc_grid_size(trunc, ctypes.byref(a), ctypes.byref(b))
    if rc:
        raise ValueError("invalid truncation")
    return a.value, b.value


def lm_index(lmax, l, m):
    """1-based (l, m) index, -1 outside the truncation."""
    return lib().orc_lm(lmax, l, m)


def gauss(n):
    mu = np.empty(n)
    w = np.empty(n)
    lib().orc_gauss(n, _p(mu), _p(w))
    return mu, w


def legendre(lmax, mu):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    nlm = (lmax + 1) * (lmax + 2) // 2
    P = np.empty((nlm, mu.size))
    D = np.empty((nlm, mu.size))
    lib().orc_legendre(lmax, mu.size, _p(mu), _p(P), _p(D))
    return P, D


def rfft_forward(x, mmax):
    x = np.ascontiguousarray(x, dtype=np.float64)
    X = np.empty(mmax + 1, dtype=np.complex128)
    lib().orc_rfft_forward(x.size, mmax, _p(x), X.ctypes.data_as(_dp))
    return X


def rfft_inverse(X, n):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    x = np.empty(n)
    lib().orc_rfft_inverse(n, X.size - 1, X.ctypes.data_as(_dp), _p(x))
    return x


class Oracle:
    """One truncation and grid, with the reference's per-call transform API."""

    def __init__(self, trunc, nlon=0, nlat=0, fft_direct=False):
        self._c = lib().orc_create(trunc, nlon, nlat)
        if not self._c:
            raise ValueError("invalid truncation")
        c = self._c.contents
        self.trunc, self.lmax, self.mmax = c.trunc, c.lmax, c.mmax
        self.nlon, self.nlat, self.nlm = c.nlon, c.nlat, c.nlm
        self.mu = np.ctypeslib.as_array(c.mu, (self.nlat,)).copy()
        self.w = np.ctypeslib.as_array(c.w, (self.nlat,)).copy()
        self.P = np.ctypeslib.as_array(c.P, (self.nlm, self.nlat)).copy()
        self.D = np.ctypeslib.as_array(c.D, (self.nlm, self.nlat)).copy()
        self.l_of_lm = np.ctypeslib.as_array(c.l_of_lm, (self.nlm,)).copy()
        self.m_of_lm = np.ctypeslib.as_array(c.m_of_lm, (self.nlm,)).copy()
        if fft_direct:
            lib().orc_set_fft_direct(self._c, 1)

    def __del__(self):
        if getattr(self, "_c", None) and _lib is not None:
            _lib.orc_destroy(self._c)
            self._c = None

    # --- helpers
    def _spec(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == (self.nlm,), a.shape
        return a

    def _grid(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.nlat, self.nlon), a.shape
        return a

    def _g(self):
        return np.empty((self.nlat, self.nlon))

    def _s(self):
        return np.empty(self.nlm, dtype=np.complex128)

    # --- transforms
    def synth(self, q):
        q = self._spec(q)
        g = self._g()
        lib().orc_synth(self._c, q.ctypes.data_as(_dp), _p(g))
        return g

    def anal(self, g):
        g = self._grid(g)
        q = self._s()
        lib().orc_anal(self._c, _p(g), q.ctypes.data_as(_dp))
        return q

    def synth_vec(self, s, t):
        s, t = self._spec(s), self._spec(t)
        vt, vp = self._g(), self._g()
        lib().orc_synth_vec(self._c, s.ctypes.data_as(_dp), t.ctypes.data_as(_dp), _p(vt), _p(vp))
        return vt, vp

    def anal_vec(self, vt, vp):
        vt, vp = self._grid(vt), self._grid(vp)
        s, t = self._s(), self._s()
        lib().orc_anal_vec(self._c, _p(vt), _p(vp), s.ctypes.data_as(_dp), t.ctypes.data_as(_dp))
        return s, t

    def synth_grad(self, s):
        s = self._spec(s)
        gt, gp = self._g(), self._g()
        lib().orc_synth_grad(self._c, s.ctypes.data_as(_dp), _p(gt), _p(gp))
        return gt, gp

    def laplacian(self, q, inverse=False):
        q = self._spec(q).copy()
        lib().orc_laplacian(self._c, q.ctypes.data_as(_dp), int(bool(inverse)))
        return q

    def surface_integral(self, g):
        return lib().orc_surface_integral(self._c, _p(self._grid(g)))

    def uv_from_vordiv(self, vor, div):
        vor, div = self._spec(vor), self._spec(div)
        u, v = self._g(), self._g()
        lib().orc_uv_from_vordiv(self._c, vor.ctypes.data_as(_dp), div.ctypes.data_as(_dp), _p(u), _p(v))
        return u, v

    def vordiv_from_uv(self, u, v):
        u, v = self._grid(u), self._grid(v)
        vor, div = self._s(), self._s()
        lib().orc_vordiv_from_uv(self._c, _p(u), _p(v), vor.ctypes.data_as(_dp), div.ctypes.data_as(_dp))
        return vor, div

    def gradient(self, f):
        f = self._spec(f)
        dx, dy = self._g(), self._g()
        lib().orc_gradient(self._c, f.ctypes.data_as(_dp), _p(dx), _p(dy))
        return dx, dy

    def tendency_grid_transforms(self, vor, div, temp, lnps):
        """Returns dict of grid fields; 3-D inputs are (nlev, nlm)."""
        nlev = vor.shape[0]
        vor, div, temp = (np.ascontiguousarray(a, dtype=np.complex128) for a in (vor, div, temp))
        lnps = self._spec(lnps)
        sh3 = (nlev, self.nlat, self.nlon)
        out = {k: np.empty(sh3) for k in ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy")}
        out.update({k: self._g() for k in ("lnps_g", "dlnpsdx", "dlnpsdy")})
        lib().orc_tendency_grid_transforms(
            self._c, nlev, *[a.ctypes.data_as(_dp) for a in (vor, div, temp, lnps)],
            *[_p(out[k]) for k in ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy",
                                   "lnps_g", "dlnpsdx", "dlnpsdy")])
        return out

    def tendency_spectral(self, ae, an, ek, dtdt, dlnpsdt):
        nlev = ae.shape[0]
        ae, an, ek, dtdt = (np.ascontiguousarray(a, dtype=np.float64) for a in (ae, an, ek, dtdt))
        dlnpsdt = self._grid(dlnpsdt)
        tv = np.empty((nlev, self.nlm), dtype=np.complex128)
        td = np.empty_like(tv)
        tt = np.empty_like(tv)
        tl = self._s()
        lib().orc_tendency_spectral(self._c, nlev, _p(ae), _p(an), _p(ek), _p(dtdt), _p(dlnpsdt),
                                    tv.ctypes.data_as(_dp), td.ctypes.data_as(_dp),
                                    tt.ctypes.data_as(_dp), tl.ctypes.data_as(_dp))
        return {"vor": tv, "div": td, "temp": tt, "lnps": tl}

    def bench_set(self, spec_in, nrep, nthreads):
        """Seconds per (synthesis + analysis) unit over an OpenMP loop of units."""
        spec_in = np.ascontiguousarray(spec_in, dtype=np.complex128)
        nlev = spec_in.shape[0]
        gw = np.empty((nlev, self.nlat, self.nlon))
        so = np.empty_like(spec_in)
        return lib().orc_bench_set(self._c, nlev, nrep, nthreads, spec_in.ctypes.data_as(_dp),
                                   _p(gw), so.ctypes.data_as(_dp))


class Port:
    """The oracle's scalar transforms arranged for CPU speed (oracle/cpu_port.c): what the
    CPU baseline times. Same API as Oracle.synth / Oracle.anal."""

    def __init__(self, oracle):
        self._o = oracle   # keeps the tables' owner alive
        self._p = lib().orc_port_create(oracle._c)
        if not self._p:
            raise ValueError("the CPU port needs an even nlat and nlon/2 with factors 2, 3, 5 only")
        self.nlat, self.nlon, self.nlm = oracle.nlat, oracle.nlon, oracle.nlm
        self._work = np.empty(lib().orc_port_work_doubles(self._p))

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.orc_port_destroy(self._p)
            self._p = None

    def synth(self, q):
        q = np.ascontiguousarray(q, dtype=np.complex128)
        g = np.empty((self.nlat, self.nlon))
        lib().orc_port_synth(self._p, q.ctypes.data_as(_dp), _p(g), _p(self._work))
        return g

    def anal(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        q = np.empty(self.nlm, dtype=np.complex128)
        lib().orc_port_anal(self._p, _p(g), q.ctypes.data_as(_dp), _p(self._work))
        return q

    def bench_set(self, spec_in, nrep, nthreads):
        """Seconds per (synthesis + analysis) unit over an OpenMP loop of the rows of spec_in."""
        spec_in = np.ascontiguousarray(spec_in, dtype=np.complex128)
        nunit = spec_in.shape[0]
        gw = np.empty((nunit, self.nlat, self.nlon))
        so = np.empty_like(spec_in)
        return lib().orc_port_bench_set(self._p, nunit, nrep, nthreads, spec_in.ctypes.data_as(_dp),
                                        _p(gw), so.ctypes.data_as(_dp))


def port_cfft(x, sign):
    """Complex FFT of the CPU port: sum_j x_j e^{sign 2 pi i jk/n}."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    if lib().orc_port_cfft(x.size, sign, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp)):
        raise ValueError("length with a prime factor other than 2, 3, 5")
    return y


# ------------------------------------------------------------- column physics

R_D, CP_D, OMEGA, P0 = 287.05, 1004.64, 7.292e-5, 1.0e5
KAPPA = R_D / CP_D


def vgrid_init(nlev, stretch_a=0.4, stretch_r=2.0, sigma_t=0.2, p_top=1000.0, ps_ref=P0):
    """Hybrid coefficients (A, B) at nlev+1 half levels, top first; the reference
    generator's defaults are its production grid (src/dynamics/aeros_vertical.f90:178-186)."""
    A, B = np.zeros(nlev + 1), np.zeros(nlev + 1)
    rc = lib().orc_vgrid_init(nlev, stretch_a, stretch_r, sigma_t, p_top, ps_ref, _p(A), _p(B))
    if rc != 0:
        raise ValueError("invalid vertical-grid parameters")
    return A, B


def vgrid_pressure(A, B, ps):
    n = len(A) - 1
    ph, pf, dp = np.zeros(n + 1), np.zeros(n), np.zeros(n)
    lib().orc_vgrid_pressure(n, _p(np.ascontiguousarray(A)), _p(np.ascontiguousarray(B)), ps, _p(ph), _p(pf), _p(dp))
    return ph, pf, dp


def vgrid_alpha(p_half):
    n = len(p_half) - 1
    al, dl = np.zeros(n), np.zeros(n)
    lib().orc_vgrid_alpha(n, _p(np.ascontiguousarray(p_half)), _p(al), _p(dl))
    return al, dl


def hydrostatic(phis, temp, p_half):
    n = len(temp)
    phi = np.zeros(n)
    lib().orc_hydrostatic(n, phis, _p(np.ascontiguousarray(temp, dtype=np.float64)),
                          _p(np.ascontiguousarray(p_half)), _p(phi))
    return phi


def hs_teq(p, sinlat2, coslat2):
    return lib().orc_hs_teq(p, sinlat2, coslat2)


def hs_kt(sigma, coslat4):
    return lib().orc_hs_kt(sigma, coslat4)


def hs_kv(sigma):
    return lib().orc_hs_kv(sigma)


def column_terms(A, B, sinlat, u, v, vor_g, div_g, t_g, dtdx, dtdy, lnps_g, dlnpsdx, dlnpsdy,
                 phis, hs=False):
    """Grid right-hand sides (ae, an, ek, dtdt, dlnpsdt) of every column; 3-D inputs
    are (nlev, nlat, nlon), 2-D (nlat, nlon)."""
    nlev, nlat, nlon = u.shape
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    ins3 = [c(x) for x in (u, v, vor_g, div_g, t_g, dtdx, dtdy)]
    ins2 = [c(x) for x in (lnps_g, dlnpsdx, dlnpsdy, phis)]
    ae, an, ek, dtdt = (np.zeros((nlev, nlat, nlon)) for _ in range(4))
    dlnpsdt = np.zeros((nlat, nlon))
    lib().orc_column_terms(nlon, nlat, nlev, _p(c(A)), _p(c(B)), _p(c(sinlat)),
                           *[_p(x) for x in ins3 + ins2],
                           _p(ae), _p(an), _p(ek), _p(dtdt), _p(dlnpsdt), 1 if hs else 0)
    return ae, an, ek, dtdt, dlnpsdt


# ------------------------------------------------------------- step completion

GRAV = 9.80665   # src/aeros_defs.f90:85


def _cv(a):
    """complex (..., nlm) -> contiguous complex128 viewed as interleaved doubles."""
    return np.ascontiguousarray(a, dtype=np.complex128)


def _cp(a):
    return a.ctypes.data_as(_dp)


class SemiImplicit:
    """The reference's aeros_semiimp_class on the CPU (src/dynamics/aeros_semiimp.f90)."""

    def __init__(self, A, B, lmax, dt_step, t_ref=300.0, ps_ref=P0, alpha=0.5):
        self.nlev, self.lmax = len(A) - 1, lmax
        self._s = lib().orc_si_create(self.nlev, _p(np.ascontiguousarray(A, dtype=np.float64)),
                                      _p(np.ascontiguousarray(B, dtype=np.float64)), t_ref, ps_ref,
                                      lmax, dt_step, alpha)
        if not self._s:
            raise ValueError("invalid semi-implicit parameters")
        self.l_of_lm = np.array([l for m in range(lmax + 1) for l in range(m, lmax + 1)], dtype=np.int32)

    def __del__(self):
        if getattr(self, "_s", None):
            lib().orc_si_destroy(self._s)
            self._s = None

    def set_step(self, dt_step):
        if lib().orc_si_set_step(self._s, dt_step) != 0:
            raise ValueError("step must be > 0")

    def gwspeed(self):
        return lib().orc_si_gwspeed(self._s)

    def _li(self):
        return self.l_of_lm.ctypes.data_as(_ip)

    def linear(self, div, temp, lnps):
        div, temp, lnps = _cv(div), _cv(temp), _cv(lnps)
        ld, lt, lp = np.empty_like(div), np.empty_like(temp), np.empty_like(lnps)
        lib().orc_si_linear(self._s, lnps.size, self._li(), _cp(div), _cp(temp), _cp(lnps),
                            _cp(ld), _cp(lt), _cp(lp))
        return ld, lt, lp

    def step(self, old, now, tnd):
        """old/now/tnd: dicts with div, temp (nlev, nlm) and lnps (nlm,). Returns new (div, temp, lnps)."""
        ins = [_cv(x[k]) for x in (old, now, tnd) for k in ("div", "temp", "lnps")]
        nd, nt, nl = np.empty_like(ins[0]), np.empty_like(ins[1]), np.empty_like(ins[2])
        lib().orc_si_step(self._s, ins[2].size, self._li(), *[_cp(a) for a in ins], _cp(nd), _cp(nt), _cp(nl))
        return {"div": nd, "temp": nt, "lnps": nl}


def diff_factors(lmax, h, tau, ndiff=6):
    f = np.empty(lmax + 1)
    lib().orc_diff_factors(lmax, h, tau, ndiff, _p(f))
    return f


def diffuse(l_of_lm, lmax, fac_vt, fac_div, vor, div, temp):
    """In place on copies; fac_* are (nlev, lmax+1). Returns (vor, div, temp)."""
    vor, div, temp = _cv(vor).copy(), _cv(div).copy(), _cv(temp).copy()
    li = np.ascontiguousarray(l_of_lm, dtype=np.int32)
    lib().orc_diffuse(vor.shape[-1], vor.shape[0], lmax, li.ctypes.data_as(_ip),
                      _p(np.ascontiguousarray(fac_vt, dtype=np.float64)),
                      _p(np.ascontiguousarray(fac_div, dtype=np.float64)), _cp(vor), _cp(div), _cp(temp))
    return vor, div, temp


def raw_filter(eps, raw_alpha, old, cur, new):
    """Returns filtered (cur, new) copies of one field."""
    old, cur, new = _cv(old), _cv(cur).copy(), _cv(new).copy()
    lib().orc_raw_filter(2 * old.size, eps, raw_alpha, _cp(old), _cp(cur), _cp(new))
    return cur, new


def mass_of(o, lnps):
    """Global dry-air mass [kg] of a spectral ln p_s (src/dynamics/aeros_timestep.f90:1250-1276)."""
    return o.surface_integral(np.exp(o.synth(lnps))) / GRAV


def c00(o):
    """The (0,0) coefficient of the constant field 1, read back by analysis (:573-588)."""
    return o.anal(np.ones((o.nlat, o.nlon)))[0].real


def mass_fix(o, lnps, target):
    """p_s <- r p_s through the (0,0) coefficient alone (:1225-1248). Returns (lnps, ln r)."""
    lnps = _cv(lnps).copy()
    lnr = np.log(target / mass_of(o, lnps))
    lnps[0] += lnr * c00(o)
    return lnps, lnr


def zonal_sample(u, v, temp, lnps_g):
    """One sample of aeros_diag_accumulate (src/aeros_diagnostics.f90:153-233) in plain loops
    over longitude, FP64: per (level, latitude) the zonal means [u], [v], [T] and the
    covariances about the instantaneous means [u'v'], [v'T'], [(u'^2+v'^2)/2], [T'^2];
    per latitude [p_s] (p_s = exp(ln p_s)) and the lowest level's [u]. Arrays are
    (nlev, nlat, nlon) and (nlat, nlon); returns a dict of (nlev, nlat) / (nlat,) arrays."""
    nlev, nlat, nlon = u.shape
    rn = 1.0 / nlon
    out = {k: np.zeros((nlev, nlat)) for k in ("u", "v", "temp", "uv", "vt", "eke", "tvar")}
    for k in range(nlev):
        for j in range(nlat):
            um = vm = tm = 0.0
            for i in range(nlon):
                um += u[k, j, i]
                vm += v[k, j, i]
                tm += temp[k, j, i]
            um *= rn
            vm *= rn
            tm *= rn
            uvm = vtm = ekem = tvm = 0.0
            for i in range(nlon):
                up, vp, tp = u[k, j, i] - um, v[k, j, i] - vm, temp[k, j, i] - tm
                uvm += up * vp
                vtm += vp * tp
                ekem += 0.5 * (up * up + vp * vp)
                tvm += tp * tp
            out["u"][k, j], out["v"][k, j], out["temp"][k, j] = um, vm, tm
            out["uv"][k, j], out["vt"][k, j] = uvm * rn, vtm * rn
            out["eke"][k, j], out["tvar"][k, j] = ekem * rn, tvm * rn
    out["ps"] = np.array([sum(np.exp(lnps_g[j, i]) for i in range(nlon)) * rn for j in range(nlat)])
    out["u_sfc"] = np.array([sum(u[nlev - 1, j, i] for i in range(nlon)) * rn for j in range(nlat)])
    return out
