"""Deterministic dynamics states of the reference's tendency tests, and the
measures those tests apply (tests/test_tendency.f90:188-300, 303-398), shared by
the oracle tests and the GPU tests."""
import numpy as np

from oracle.oracle import P0, R_D, lm_index

A_EARTH = 6.371e6
SQ4PI = np.sqrt(4.0 * np.pi)


def _set(q, T, l, m, val):
    q[..., lm_index(T, l, m) - 1] = val


def rest_state(T, nlm, nlev, topography, tiso=280.0):
    """Isothermal atmosphere at rest in discrete hydrostatic balance with its own
    surface geopotential (tests/test_tendency.f90:188-235). Returns the spectral
    state; phis is built from the synthesized ln p_s by the caller."""
    vor = np.zeros((nlev, nlm), complex)
    div = np.zeros((nlev, nlm), complex)
    temp = np.zeros((nlev, nlm), complex)
    lnps = np.zeros(nlm, complex)
    _set(temp, T, 0, 0, tiso * SQ4PI)
    _set(lnps, T, 0, 0, np.log(P0) * SQ4PI)
    if topography:
        _set(lnps, T, 2, 1, -0.06 + 0.03j)
        _set(lnps, T, 3, 0, 0.05)
        _set(lnps, T, 5, 3, 0.02 - 0.01j)
    return vor, div, temp, lnps


def rest_phis(lnps_g, tiso=280.0):
    return R_D * tiso * (np.log(P0) - lnps_g)


def moving_state(T, nlm, nlev, tiso=280.0):
    """A sheared rotational flow (tests/test_tendency.f90:323-354)."""
    vor = np.zeros((nlev, nlm), complex)
    div = np.zeros((nlev, nlm), complex)
    temp = np.zeros((nlev, nlm), complex)
    lnps = np.zeros(nlm, complex)
    _set(lnps, T, 0, 0, np.log(P0) * SQ4PI)
    _set(lnps, T, 2, 1, -0.03 + 0.01j)
    for k in range(nlev):
        kk = k + 1
        _set(temp[k], T, 0, 0, tiso * SQ4PI)
        _set(temp[k], T, 2, 0, -20.0 * kk / nlev)
        for m in range(0, T + 1):
            for l in range(max(m, 1), min(6, T) + 1):
                amp = 1.0e-5 / (l * l) * (0.5 + 0.5 * kk / nlev)
                zv = amp * np.cos(3 * l + m) + 1j * amp * np.sin(l + 2 * m)
                dv = 0.2 * amp * np.sin(l + m) + 1j * 0.2 * amp * np.cos(2 * l + m)
                if m == 0:
                    zv, dv = zv.real, dv.real
                _set(vor[k], T, l, m, zv)
                _set(div[k], T, l, m, dv)
    return vor, div, temp, lnps


def rest_errors(ek, tvor, tdiv, ttemp, tlnps, kmin=1):
    """Scaled tendency magnitudes of a resting state (tests/test_tendency.f90:240-292):
    momentum against max|Phi+K| at the lowest level / a^2, temperature against 280 K."""
    scale = max(np.abs(ek[-1]).max(), 1.0)
    dscale = scale / A_EARTH ** 2
    k0 = kmin - 1
    return {"vor": np.abs(tvor[k0:]).max() / dscale,
            "div": np.abs(tdiv[k0:]).max() / dscale,
            "temp": np.abs(ttemp[k0:]).max() / 280.0,
            "lnps": np.abs(tlnps).max()}


def mass_identity(dlnpsdt, lnps_g, w, nlon):
    """|global mass tendency| / its absolute scale (tests/test_tendency.f90:376-396)."""
    area = w[:, None] * (2.0 * np.pi / nlon) * A_EARTH ** 2
    x = dlnpsdt * np.exp(lnps_g) * area
    return abs(x.sum()) / np.abs(x).sum()
