"""CPU: the step-completion restatement in oracle/ (semi-implicit solve,
diffusion, RAW filter, mass fixer) pinned by the reference's acceptance checks
(tests/test_semiimp.f90, tests/test_timestep.f90) and closed forms."""
import numpy as np
import pytest

from oracle import oracle as O
from dyn_states import SQ4PI


def reldiff(a, b):
    """max|a-b| / max|b| on complex moduli (tests/test_semiimp.f90:411-448)."""
    s = np.abs(b).max()
    return np.abs(a).max() if s <= 0 else np.abs(a - b).max() / s


def base_state(T, nlm, nlev, t_ref=300.0, ps_ref=O.P0):
    """Isothermal, uniform p_s, at rest (tests/test_semiimp.f90:138-165)."""
    s = {"vor": np.zeros((nlev, nlm), complex), "div": np.zeros((nlev, nlm), complex),
         "temp": np.zeros((nlev, nlm), complex), "lnps": np.zeros(nlm, complex)}
    s["temp"][:, 0] = t_ref * SQ4PI
    s["lnps"][0] = np.log(ps_ref) * SQ4PI
    return s


def tendency(o, A, B, st):
    """The oracle's whole adiabatic right-hand side (transforms, column terms, transforms)."""
    g = o.tendency_grid_transforms(st["vor"], st["div"], st["temp"], st["lnps"])
    ae, an, ek, dtdt, dps = O.column_terms(A, B, o.mu, *[g[k] for k in (
        "u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy", "lnps_g", "dlnpsdx", "dlnpsdy")],
        np.zeros_like(g["lnps_g"]))
    return o.tendency_spectral(ae, an, ek, dtdt, dps)


def test_gravity_wave_speed():
    """tests/test_semiimp.f90:72-111: the external mode is sqrt(gamma R T) within 5%, and
    does not depend on the coordinate (1%)."""
    c_lamb = np.sqrt(O.CP_D / (O.CP_D - O.R_D) * O.R_D * 300.0)
    si = O.SemiImplicit(*O.vgrid_init(20, stretch_a=1.0, sigma_t=0.0, p_top=1000.0), 21, 1800.0)
    c_sigma = si.gwspeed()
    si = O.SemiImplicit(*O.vgrid_init(20), 21, 1800.0)
    c_hyb = si.gwspeed()
    assert abs(c_hyb - c_lamb) / c_lamb < 0.05
    assert abs(c_hyb - c_sigma) / c_sigma < 0.01


def test_linearization_against_finite_differences():
    """tests/test_semiimp.f90:167-289: L reproduces the finite-differenced nonlinear
    tendency for T, ln p_s and D perturbations (T21 L20, hybrid default)."""
    T, nlev = 21, 20
    o = O.Oracle(T)
    A, B = O.vgrid_init(nlev)
    si = O.SemiImplicit(A, B, T, 1800.0)
    base = base_state(T, o.nlm, nlev)
    t0 = tendency(o, A, B, base)
    l0 = si.linear(base["div"], base["temp"], base["lnps"])
    assert np.abs(t0["div"]).max() < 1e-18
    lm = O.lm_index(T, 3, 2) - 1
    k = np.arange(1, nlev + 1)
    # temperature
    p = {key: v.copy() for key, v in base.items()}
    p["temp"][:, lm] += 1.0e-2 * np.cos(k) + 0.6e-2j * np.sin(2 * k)
    t1, l1 = tendency(o, A, B, p), si.linear(p["div"], p["temp"], p["lnps"])
    assert reldiff(t1["div"] - t0["div"], l1[0] - l0[0]) < 1e-6
    assert np.abs(t1["temp"] - t0["temp"]).max() < 1e-15
    assert np.abs(t1["lnps"] - t0["lnps"]).max() < 1e-18
    # surface pressure: the coefficient test (R T_ref, not R T_ref c_k)
    p = {key: v.copy() for key, v in base.items()}
    p["lnps"][lm] += 2.0e-5 - 1.0e-5j
    t1, l1 = tendency(o, A, B, p), si.linear(p["div"], p["temp"], p["lnps"])
    assert reldiff(t1["div"] - t0["div"], l1[0] - l0[0]) < 1e-6
    assert np.abs(t1["temp"] - t0["temp"]).max() < 1e-15
    # divergence: tau and nu
    p = {key: v.copy() for key, v in base.items()}
    p["div"][:, lm] += 1.0e-9 * np.cos(k) + 0.7e-9j * np.sin(3 * k)
    t1, l1 = tendency(o, A, B, p), si.linear(p["div"], p["temp"], p["lnps"])
    assert reldiff(t1["temp"] - t0["temp"], l1[1] - l0[1]) < 1e-6
    assert reldiff(t1["lnps"] - t0["lnps"], l1[2] - l0[2]) < 1e-9


def solve_inputs(T, nlm, nlev, l_of_lm):
    """tests/test_semiimp.f90:323-357."""
    old, now = base_state(T, nlm, nlev), base_state(T, nlm, nlev)
    tnd = {"div": np.zeros((nlev, nlm), complex), "temp": np.zeros((nlev, nlm), complex),
           "lnps": np.zeros(nlm, complex)}
    for kk in range(1, nlev + 1):
        k = kk - 1
        for lm in range(nlm):
            l = int(l_of_lm[lm])
            if l < 1 or l > 8:
                continue
            amp = 1.0e-6 / (l * l)
            old["div"][k, lm] = amp * np.cos(l + 2 * kk) + 1j * amp * np.sin(3 * l + kk)
            now["div"][k, lm] = amp * np.cos(2 * l + kk) + 1j * amp * np.sin(l + 3 * kk)
            old["temp"][k, lm] += 0.5 * np.cos(l * kk)
            now["temp"][k, lm] += 0.4 * np.sin(l + kk)
            tnd["div"][k, lm] = 1.0e-11 * np.sin(l + kk)
            tnd["temp"][k, lm] = 1.0e-4 * np.cos(l - kk)
    for lm in range(nlm):
        l = int(l_of_lm[lm])
        if l < 1 or l > 8:
            continue
        old["lnps"][lm] = 1.0e-3 * np.cos(l)
        now["lnps"][lm] = 0.9e-3 * np.sin(l)
        tnd["lnps"][lm] = 1.0e-8 * np.cos(2 * l)
    return old, now, tnd


@pytest.mark.parametrize("alpha", [0.5, 1.0])
def test_solve_satisfies_its_defining_equation(alpha):
    """tests/test_semiimp.f90:291-409: X^new = X^old + h R + h [L(X_a) - L(X^now)],
    X_a = alpha X^new + (1 - alpha) X^old, at machine precision."""
    T, nlev, h = 21, 20, 3600.0
    A, B = O.vgrid_init(nlev)
    si = O.SemiImplicit(A, B, T, h, alpha=alpha)
    nlm = len(si.l_of_lm)
    old, now, tnd = solve_inputs(T, nlm, nlev, si.l_of_lm)
    new = si.step(old, now, tnd)
    bar = {k: alpha * new[k] + (1 - alpha) * old[k] for k in ("div", "temp", "lnps")}
    lb, ln_ = si.linear(bar["div"], bar["temp"], bar["lnps"]), si.linear(now["div"], now["temp"], now["lnps"])
    for i, (k, tol) in enumerate([("div", 1e-10), ("temp", 1e-12), ("lnps", 1e-12)]):
        r = np.abs(new[k] - old[k] - h * (tnd[k] + lb[i] - ln_[i])).max() / np.abs(new[k] - old[k]).max()
        assert r < tol, (k, r)


def test_solve_planted_mistake_is_caught():
    """A slip in the elimination (the explicit half-step weight) fails the defining equation."""
    T, nlev, h = 21, 20, 3600.0
    A, B = O.vgrid_init(nlev)
    si = O.SemiImplicit(A, B, T, h)
    nlm = len(si.l_of_lm)
    old, now, tnd = solve_inputs(T, nlm, nlev, si.l_of_lm)
    new = si.step(old, now, {k: v * 1.01 for k, v in tnd.items()})   # planted: wrong tendency
    bar = {k: 0.5 * (new[k] + old[k]) for k in ("div", "temp", "lnps")}
    lb, ln_ = si.linear(bar["div"], bar["temp"], bar["lnps"]), si.linear(now["div"], now["temp"], now["lnps"])
    r = np.abs(new["temp"] - old["temp"] - h * (tnd["temp"] + lb[1] - ln_[1])).max() / np.abs(new["temp"] - old["temp"]).max()
    assert r > 1e-6


def test_diffusion_factors_and_application():
    lmax, h, tau = 21, 3600.0, 0.25 * 86400.0
    f = O.diff_factors(lmax, h, tau, 6)
    assert f[0] == 1.0 and abs(f[lmax] - 1.0 / (1.0 + h / tau)) < 1e-16
    x = 12 * 13 / (21 * 22.0)
    assert abs(f[12] - 1.0 / (1.0 + h / tau * x ** 3)) < 1e-16
    nlev, nlm = 3, (lmax + 1) * (lmax + 2) // 2
    li = np.array([l for m in range(lmax + 1) for l in range(m, lmax + 1)], dtype=np.int32)
    fv = np.stack([f] * nlev)
    fd = np.stack([f ** 2] * nlev)
    rng = np.random.default_rng(3)
    z = rng.standard_normal((nlev, nlm)) + 1j * rng.standard_normal((nlev, nlm))
    v, d, t = O.diffuse(li, lmax, fv, fd, z, 2 * z, 3 * z)
    assert np.array_equal(v, f[li] * z) and np.array_equal(d, f[li] ** 2 * (2 * z))
    assert np.array_equal(t, f[li] * (3 * z))


def test_raw_filter():
    """d = nu/2 (X^- - 2X + X^+); raw_alpha = 1 is Robert-Asselin (new untouched);
    raw_alpha = 1/2 conserves the three-level mean (Williams 2009)."""
    rng = np.random.default_rng(5)
    a, b, c = (rng.standard_normal(50) + 1j * rng.standard_normal(50) for _ in range(3))
    cur, new = O.raw_filter(0.1, 1.0, a, b, c)
    assert np.array_equal(new, c) and np.allclose(cur, b + 0.05 * (a - 2 * b + c), rtol=0, atol=1e-15)
    cur, new = O.raw_filter(0.1, 0.5, a, b, c)
    assert np.allclose(cur + new, b + c, rtol=0, atol=1e-15)
    cur, new = O.raw_filter(0.0, 0.5, a, b, c)
    assert np.array_equal(cur, b) and np.array_equal(new, c)


def test_mass_fixer_restores_the_target_exactly():
    """src/dynamics/aeros_timestep.f90:1225-1276: after the fix the state's mass equals
    the target; only the (0,0) coefficient moves; c00 = sqrt(4 pi)."""
    T = 21
    o = O.Oracle(T)
    assert abs(O.c00(o) - SQ4PI) < 1e-14
    lnps = np.zeros(o.nlm, complex)
    lnps[0] = np.log(1.0e5) * SQ4PI
    lnps[O.lm_index(T, 2, 1) - 1] = -0.03 + 0.01j
    m0 = O.mass_of(o, lnps)
    # closed form for uniform p_s: 4 pi a^2 p_s / g
    uni = np.zeros(o.nlm, complex)
    uni[0] = np.log(1.0e5) * SQ4PI
    assert abs(O.mass_of(o, uni) / (4 * np.pi * 6.371e6 ** 2 * 1.0e5 / O.GRAV) - 1) < 1e-13
    target = m0 * (1 + 3e-4)
    fixed, lnr = O.mass_fix(o, lnps, target)
    assert abs(O.mass_of(o, fixed) / target - 1) < 1e-14
    assert np.array_equal(fixed[1:], lnps[1:]) and abs(lnr - np.log(1 + 3e-4)) < 1e-13
