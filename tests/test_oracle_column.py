"""CPU: the column-physics restatement in oracle/ (vertical coordinate, column
terms, Held-Suarez forcing) pinned by the reference's own known answers."""
import numpy as np
import pytest

from oracle import oracle as O
from dyn_states import (mass_identity, moving_state, rest_errors, rest_phis, rest_state)

K_A, K_S, K_F = 1 / (40 * 86400.0), 1 / (4 * 86400.0), 1 / 86400.0


# ------------------------------------------------------------ Held-Suarez

def test_hs_point_values():
    """tests/test_held_suarez.f90:88-172."""
    assert abs(O.hs_teq(1.0e5, 0.0, 1.0) - 315.0) < 1e-12          # equator, surface
    assert abs(O.hs_teq(1.0e5, 1.0, 0.0) - 255.0) < 1e-12          # pole, surface
    expect = (315.0 - 60.0 * 0.5 - 10.0 * np.log(0.5) * 0.5) * 0.5 ** O.KAPPA
    assert abs(O.hs_teq(0.5e5, 0.5, 0.5) - expect) < 1e-12         # 45N, 500 hPa
    assert abs(O.hs_teq(1.0e3, 0.0, 1.0) - 200.0) < 1e-12          # the 200 K floor
    assert abs(O.hs_kt(1.0, 1.0) - K_S) < 1e-18
    assert abs(O.hs_kt(0.5, 1.0) - K_A) < 1e-18
    assert abs(O.hs_kt(1.0, 0.0) - K_A) < 1e-18
    assert O.hs_kv(0.5) == 0.0 and O.hs_kv(0.7) == 0.0
    assert abs(O.hs_kv(1.0) - K_F) < 1e-18
    assert abs(O.hs_kv(0.85) - 0.5 * K_F) < 1e-18


# ------------------------------------------------------------ vertical grid

def test_vgrid_pure_sigma_and_endpoints():
    """sigma_t = 0 is pure sigma exactly (A = 0 below the top); uniform placement for
    stretch_a = 1 (src/dynamics/aeros_vertical.f90:197-290)."""
    A, B = O.vgrid_init(20, stretch_a=1.0, sigma_t=0.0, p_top=0.0)
    assert np.all(A == 0.0)
    assert np.abs(B - np.arange(21) / 20.0).max() < 1e-15
    A, B = O.vgrid_init(20)                                   # production default
    assert B[0] == 0.0 and A[0] == 1000.0 and B[-1] == 1.0 and A[-1] == 0.0
    ph, pf, dp = O.vgrid_pressure(A, B, 1.0e5)
    assert np.all(dp > 0) and abs(dp.sum() - (1.0e5 - 1000.0)) < 1e-9
    assert np.all(B[1:] >= B[:-1]) and np.any(B[1:-1] == 0.0)   # hybrid: B = 0 aloft
    with pytest.raises(ValueError):
        O.vgrid_init(1)


def test_alpha_and_hydrostatic_closed_forms():
    A, B = O.vgrid_init(10, stretch_a=1.0, sigma_t=0.0, p_top=0.0)
    ph, pf, dp = O.vgrid_pressure(A, B, 9.0e4)
    al, dl = O.vgrid_alpha(ph)
    assert al[0] == np.log(2.0) and dl[0] == -9999.0            # zero-pressure top
    k = np.arange(1, 10)
    expect = 1.0 - ph[k] / dp[k] * np.log(ph[k + 1] / ph[k])
    assert np.abs(al[1:] - expect).max() < 1e-15
    # isothermal column: Phi at the half levels is Phi_s + R T ln(p_s/p)
    T = np.full(10, 250.0)
    phi = O.hydrostatic(100.0, T, ph)
    half = 100.0 + O.R_D * 250.0 * np.log(ph[-1] / ph[1:])     # Phi_{k+1/2}, k = 0..9
    assert np.abs(phi - (half + al * O.R_D * 250.0)).max() < 1e-9


# ------------------------------------------------------------ column terms

def _tendency(o, A, B, vor, div, temp, lnps, phis=None, hs=False, tiso=280.0, rest=False):
    g = o.tendency_grid_transforms(vor, div, temp, lnps)
    if phis is None:
        phis = rest_phis(g["lnps_g"], tiso) if rest else np.zeros_like(g["lnps_g"])
    ae, an, ek, dtdt, dps = O.column_terms(A, B, o.mu, g["u"], g["v"], g["vor_g"], g["div_g"], g["t_g"],
                                           g["dtdx"], g["dtdy"], g["lnps_g"], g["dlnpsdx"], g["dlnpsdy"],
                                           phis, hs=hs)
    t = o.tendency_spectral(ae, an, ek, dtdt, dps)
    return g, (ae, an, ek, dtdt, dps), (t["vor"], t["div"], t["temp"], t["lnps"])


@pytest.mark.parametrize("case,vgrid,topo,tol,kmin", [
    ("flat, pure sigma", dict(sigma_t=0.0, p_top=0.0), False, 1e-12, 1),
    ("topography, hybrid default", dict(), True, 1e-9, 1),
    ("topography, pure sigma, p_top = 0", dict(sigma_t=0.0, p_top=0.0), True, 1e-9, 2),
])
def test_rest_state_balance(case, vgrid, topo, tol, kmin):
    """tests/test_tendency.f90:77-107: a resting isothermal state in discrete
    hydrostatic balance has vanishing tendencies. T21 L12, the reference's sizes."""
    T, nlev = 21, 12
    o = O.Oracle(T)
    A, B = O.vgrid_init(nlev, **vgrid)
    vor, div, temp, lnps = rest_state(T, o.nlm, nlev, topo)
    g, (ae, an, ek, dtdt, dps), (tv, td, tt, tl) = _tendency(o, A, B, vor, div, temp, lnps, rest=True)
    e = rest_errors(ek, tv, td, tt, tl, kmin)
    if topo:
        assert max(e.values()) < tol, (case, e)
        assert np.abs(g["dlnpsdx"]).max() > 1e-8   # not vacuous: the terms really cancel
    else:
        # Flat: the column terms are exact -- no momentum, heating or mass tendency,
        # and Phi+K constant on every level -- so vorticity, temperature and ln p_s
        # tendencies are exactly zero. The divergence tendency is then -lap of the
        # analysis of a constant up to 2.5e5 m2/s2: FP64 quadrature rounding of one
        # ulp of c_00 at l = 20 is l(l+1)/a^2 * 1e-10 against a scale of 3.4e3/a^2,
        # i.e. ~1e-11 -- the floor of any FP64 analysis; the reference's 1e-12 for
        # this case is below it, so the bound here is 1e-10.
        assert np.all(ae == 0) and np.all(an == 0) and np.all(dtdt == 0) and np.all(dps == 0)
        assert all(np.all(ek[k] == ek[k].flat[0]) for k in range(nlev))
        assert e["vor"] == 0 and e["temp"] == 0 and e["lnps"] == 0
        assert e["div"] < 1e-10, e


def test_moving_state_mass_identity():
    """tests/test_tendency.f90:303-398 (T21 L12)."""
    T, nlev = 21, 12
    o = O.Oracle(T)
    A, B = O.vgrid_init(nlev, sigma_t=0.0, p_top=0.0)
    vor, div, temp, lnps = moving_state(T, o.nlm, nlev)
    g, (ae, an, ek, dtdt, dps), (tv, td, tt, tl) = _tendency(o, A, B, vor, div, temp, lnps)
    assert 1.0 < np.abs(g["u"]).max() < 150.0
    assert all(np.isfinite(x).all() for x in (tv, td, tt, tl))
    assert mass_identity(dps, g["lnps_g"], o.w, o.nlon) < 1e-12


def test_held_suarez_is_applied_as_specified():
    """tests/test_held_suarez.f90:176-281: the forcing adds exactly -k_T (T - T_eq) and
    -k_v v, and leaves the free atmosphere (sigma < 0.7) undamped."""
    T, nlev = 21, 20
    o = O.Oracle(T)
    A, B = O.vgrid_init(nlev)
    vor, div, temp, lnps = moving_state(T, o.nlm, nlev)
    g = o.tendency_grid_transforms(vor, div, temp, lnps)
    args = [g[k] for k in ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy", "lnps_g", "dlnpsdx", "dlnpsdy")]
    phis = np.zeros_like(g["lnps_g"])
    a0 = O.column_terms(A, B, o.mu, *args, phis, hs=False)
    a1 = O.column_terms(A, B, o.mu, *args, phis, hs=True)
    assert np.array_equal(a0[2], a1[2]) and np.array_equal(a0[4], a1[4])   # ek, dlnps/dt untouched
    j, i = 10, 7
    ps = np.exp(g["lnps_g"][j, i])
    _, pf, _ = O.vgrid_pressure(A, B, ps)
    s2 = o.mu[j] ** 2
    for k in range(nlev):
        sig = pf[k] / ps
        kt, kv = O.hs_kt(sig, (1 - s2) ** 2), O.hs_kv(sig)
        teq = O.hs_teq(pf[k], s2, 1 - s2)
        assert a1[3][k, j, i] == a0[3][k, j, i] - kt * (g["t_g"][k, j, i] - teq)
        assert a1[0][k, j, i] == a0[0][k, j, i] - kv * g["u"][k, j, i]
        assert a1[1][k, j, i] == a0[1][k, j, i] - kv * g["v"][k, j, i]
    upper = O.vgrid_pressure(A, B, 1.0e5)[1] / 1.0e5 < 0.7 * 0.9
    assert np.array_equal(a1[0][upper], a0[0][upper])


def test_planted_mistakes_are_caught():
    """Mistakes planted in the oracle's outputs (on the CPU) fail the balance checks."""
    T, nlev = 21, 20
    o = O.Oracle(T)
    A, B = O.vgrid_init(nlev)
    vor, div, temp, lnps = rest_state(T, o.nlm, nlev, True)
    g = o.tendency_grid_transforms(vor, div, temp, lnps)
    args = [g[k] for k in ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy", "lnps_g", "dlnpsdx", "dlnpsdy")]
    ae, an, ek, dtdt, dps = O.column_terms(A, B, o.mu, *args, rest_phis(g["lnps_g"]))
    # pressure-gradient term with the wrong sign: the balance is lost
    pgf_x = ae.copy()
    bad_ae = ae - 2 * pgf_x
    t = o.tendency_spectral(bad_ae, an, ek, dtdt, dps)
    assert rest_errors(ek, t["vor"], t["div"], t["temp"], t["lnps"])["div"] > 1e-3
    # geopotential of the wrong level (shifted by one)
    t = o.tendency_spectral(ae, an, np.roll(ek, 1, axis=0), dtdt, dps)
    assert rest_errors(ek, t["vor"], t["div"], t["temp"], t["lnps"])["div"] > 1e-3


@pytest.mark.parametrize("kw", [dict(), dict(stretch_a=1.0, sigma_t=0.0, p_top=0.0), dict(sigma_t=0.35, p_top=500.0)])
def test_product_vgrid_generator_matches_oracle(kw):
    """The host-side generator of aeros_amd (used by callers without an A/B table)
    against the oracle's restatement."""
    from aeros_amd.column import vgrid_init
    for nlev in (12, 20, 31):
        a, b = vgrid_init(nlev, **kw)
        A, B = O.vgrid_init(nlev, **kw)
        assert np.abs(a - A).max() <= 1e-9 and np.abs(b - B).max() <= 1e-15
