"""The zonal-statistics restatement (aeros_diag_accumulate) against closed forms, and the
summary and pressure-axis host logic of aeros_amd.diagnostics."""
import numpy as np

from oracle.oracle import zonal_sample


def _fields(nlev=3, nlat=4, nlon=16):
    lam = 2 * np.pi * np.arange(nlon) / nlon
    U = np.linspace(-5, 20, nlev * nlat).reshape(nlev, nlat)
    V = np.linspace(-1, 2, nlev * nlat).reshape(nlev, nlat)
    T = np.linspace(210, 300, nlev * nlat).reshape(nlev, nlat)
    a, b, c = 3.0, 2.0, 1.5
    u = U[..., None] + a * np.cos(3 * lam)
    v = V[..., None] + b * np.cos(3 * lam) + 0.7 * np.sin(2 * lam)
    t = T[..., None] + c * np.cos(3 * lam + 0.25)
    lnps = np.log(1e5) + 0.01 * np.cos(lam)[None, :] + np.zeros((nlat, 1))
    return u, v, t, lnps, (U, V, T, a, b, c)


def test_zonal_sample_closed_forms():
    """Means of a wave are zero; the covariance of two waves of one wavenumber is half the
    product of amplitudes times the cosine of their phase difference; different
    wavenumbers are uncorrelated."""
    u, v, t, lnps, (U, V, T, a, b, c) = _fields()
    d = zonal_sample(u, v, t, lnps)
    assert np.allclose(d["u"], U, atol=1e-13) and np.allclose(d["v"], V, atol=1e-13)
    assert np.allclose(d["temp"], T, atol=1e-12)
    assert np.allclose(d["uv"], a * b / 2, atol=1e-13)
    assert np.allclose(d["vt"], b * c / 2 * np.cos(0.25), atol=1e-13)
    assert np.allclose(d["eke"], 0.5 * (a * a / 2 + b * b / 2 + 0.7 ** 2 / 2), atol=1e-13)
    assert np.allclose(d["tvar"], c * c / 2, atol=1e-12)
    # mean of exp(ln p0 + e cos) = p0 I0(e)
    from scipy.special import i0
    assert np.allclose(d["ps"], 1e5 * i0(0.01), rtol=1e-14)
    assert np.allclose(d["u_sfc"], U[-1], atol=1e-13)


def test_zonal_sample_primes_are_instantaneous():
    """Covariances are about the instantaneous zonal mean: adding a zonally uniform
    offset leaves them unchanged (the total-eddy definition of the reference)."""
    u, v, t, lnps, _ = _fields()
    d0 = zonal_sample(u, v, t, lnps)
    d1 = zonal_sample(u + 7.0, v - 3.0, t + 11.0, lnps)
    for k in ("uv", "vt", "eke", "tvar"):
        assert np.allclose(d0[k], d1[k], atol=1e-12)


def test_summary_locates_the_jet():
    from aeros_amd.diagnostics import summary
    nlat, nlev = 6, 4
    lat = np.array([70.0, 40.0, 10.0, -10.0, -40.0, -70.0])
    u = np.zeros((nlat, nlev))
    u[1, 1] = 30.0
    u[4, 1] = 29.0
    vt = np.zeros((nlat, nlev))
    vt[2, 3] = 15.0
    temp = np.full((nlat, nlev), 250.0)
    temp[nlat // 2 - 1, nlev - 1] = 305.0
    temp[0, nlev - 1] = 260.0
    d = {"u": u, "v": u * 0, "temp": temp, "uv": u * 2, "vt": vt, "eke": u * 3, "tvar": u,
         "ps": np.full(nlat, 1e5), "u_sfc": np.array([0, 9.0, 1, 1, 8, 0]),
         "pressure": np.tile(np.linspace(5e3, 9.5e4, nlev), (nlat, 1))}
    s = summary(d, lat)
    assert s["u_max"] == 30.0 and s["u_max_lat"] == 40.0
    assert abs(s["u_max_hpa"] - d["pressure"][1, 1] / 100) < 1e-12
    assert s["u_sfc_max"] == 9.0 and s["u_sfc_max_lat"] == 40.0
    assert s["vt_max"] == 15.0 and s["vt_max_lat"] == 10.0
    assert s["uv_max"] == 60.0 and s["eke_max"] == 90.0
    assert s["t_sfc_equator"] == 305.0 and s["t_sfc_pole"] == 260.0
    expect = np.sqrt(((u - u[::-1]) ** 2).sum() / (u ** 2).sum())
    assert abs(s["u_asymmetry"] - expect) < 1e-15


def test_hs_perturb_matches_the_reference_formula():
    """aeros_hs_perturb's seed, written out for one coefficient."""
    import math
    from aeros_amd.held_suarez import hs_perturb

    class S:
        lmax = mmax = 21

        def lm(self, l, m):
            return m * (2 * 21 + 3 - m) // 2 + (l - m) + 1

    s = S()
    t = np.zeros((3, 253), complex)
    hs_perturb(t, s, 1.0)
    l, m, k = 7, 5, 2
    want = complex(math.cos(3 * l + 5 * m + k), math.sin(7 * l + 2 * m + 3 * k))
    assert t[k - 1, s.lm(l, m) - 1] == want
    assert t[0, s.lm(10, 5) - 1] == 0 and t[0, s.lm(3, 0) - 1] == 0 and t[0, s.lm(9, 9) - 1] == 0
