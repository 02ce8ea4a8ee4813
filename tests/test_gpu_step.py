"""GPU: the step completion (semi-implicit solve, leapfrog, diffusion, RAW filter,
mass fixer) against the oracle on the same inputs, and the reference's
defining-equation check on the GPU's own output."""
import numpy as np
import pytest

from helpers import rel
from test_oracle_step import solve_inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def _dev(torch, d):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in d.items()}


def _host(d):
    return {k: v.cpu().numpy() for k, v in d.items()}


def _setup(T, nlev, h, alpha, vgrid=None):
    from aeros_amd.column import VGrid
    from aeros_amd.spectral import AerosSht
    from aeros_amd.step import SemiImplicit
    from oracle import oracle as O
    A, B = O.vgrid_init(nlev, **(vgrid or {}))
    s = AerosSht(T)
    vg = VGrid(A, B)
    return s, vg, SemiImplicit(s, vg, h, alpha=alpha), O.SemiImplicit(A, B, T, h, alpha=alpha)


@pytest.mark.parametrize("T,nlev,alpha", [(21, 20, 0.5), (21, 20, 1.0), (42, 20, 0.5), (31, 64, 0.5)])
def test_semiimp_step_matches_oracle(torch, T, nlev, alpha):
    h = 3600.0
    s, vg, si, osi = _setup(T, nlev, h, alpha)
    old, now, tnd = solve_inputs(T, s.nlm, nlev, osi.l_of_lm)
    new = _host(si.step(_dev(torch, old), _dev(torch, now), _dev(torch, tnd)))
    ref = osi.step(old, now, tnd)
    for k in ("div", "temp", "lnps"):
        # FP64 both sides; the GPU applies the inverse of the LU factors the oracle
        # back-substitutes through, so they agree to a few ulps times the
        # Helmholtz matrix's condition number
        assert rel(new[k], ref[k]) < 1e-12, (k, rel(new[k], ref[k]))
    # the scheme's defining equation on the GPU's own output (tests/test_semiimp.f90:359-378)
    bar = {k: alpha * new[k] + (1 - alpha) * old[k] for k in ("div", "temp", "lnps")}
    lb = osi.linear(bar["div"], bar["temp"], bar["lnps"])
    ln_ = osi.linear(now["div"], now["temp"], now["lnps"])
    for i, (k, tol) in enumerate([("div", 1e-10), ("temp", 1e-12), ("lnps", 1e-12)]):
        r = np.abs(new[k] - old[k] - h * (tnd[k] + lb[i] - ln_[i])).max() / np.abs(new[k] - old[k]).max()
        assert r < tol, (k, r)


def test_linear_and_gwspeed_match_oracle(torch):
    T, nlev = 21, 20
    s, vg, si, osi = _setup(T, nlev, 1800.0, 0.5)
    old, now, tnd = solve_inputs(T, s.nlm, nlev, osi.l_of_lm)
    d = _dev(torch, now)
    got = [x.cpu().numpy() for x in si.linear(d["div"], d["temp"], d["lnps"])]
    ref = osi.linear(now["div"], now["temp"], now["lnps"])
    for g, r in zip(got, ref):
        assert rel(g, r) < 1e-13
    assert abs(si.gwspeed() / osi.gwspeed() - 1) < 1e-12


@pytest.mark.parametrize("semi", [True, False])
def test_leapfrog_with_diffusion_and_filter_matches_oracle(torch, semi):
    from oracle import oracle as O
    T, nlev, h = 42, 20, 2 * 1200.0
    s, vg, si, osi = _setup(T, nlev, h, 0.5)
    tau, eps, ra = 0.1 * 86400.0, 0.06, 0.53
    si.set_diffusion(tau, 0.5 * tau, 6)
    old, now, tnd = solve_inputs(T, s.nlm, nlev, osi.l_of_lm)
    rng = np.random.default_rng(11)
    z = lambda: (rng.standard_normal((nlev, s.nlm)) + 1j * rng.standard_normal((nlev, s.nlm))) * 1e-6
    old["vor"], now["vor"], tnd["vor"] = z(), z(), z() * 1e-5
    dnow = _dev(torch, now)
    new = _host(si.leapfrog(_dev(torch, old), dnow, _dev(torch, tnd), semi_implicit=semi,
                            diffuse=True, eps_filter=eps, raw_alpha=ra))
    cur = _host(dnow)
    # the oracle, stage by stage in the reference's order
    ref = osi.step(old, now, tnd) if semi else {k: old[k] + h * tnd[k] for k in ("div", "temp", "lnps")}
    ref["vor"] = old["vor"] + h * tnd["vor"]
    fv = np.stack([O.diff_factors(T, h, tau, 6)] * nlev)
    fd = np.stack([O.diff_factors(T, h, 0.5 * tau, 6)] * nlev)
    ref["vor"], ref["div"], ref["temp"] = O.diffuse(osi.l_of_lm, T, fv, fd, ref["vor"], ref["div"], ref["temp"])
    rcur = {}
    for k in ("vor", "div", "temp", "lnps"):
        rcur[k], ref[k] = O.raw_filter(eps, ra, old[k], now[k], ref[k])
    for k in ("vor", "div", "temp", "lnps"):
        assert rel(new[k], ref[k]) < 1e-12, ("new", k, rel(new[k], ref[k]))
        assert rel(cur[k], rcur[k]) < 1e-12, ("now", k, rel(cur[k], rcur[k]))


def test_mass_and_fixer_match_oracle(torch):
    from aeros_amd.step import mass
    from oracle import oracle as O
    T = 42
    s, vg, si, osi = _setup(T, 10, 1800.0, 0.5)
    o = O.Oracle(T)
    lnps = np.zeros(s.nlm, complex)
    lnps[0] = np.log(1.0e5) * np.sqrt(4 * np.pi)
    lnps[O.lm_index(T, 2, 1) - 1] = -0.03 + 0.01j
    lnps[O.lm_index(T, 5, 3) - 1] = 0.004 - 0.002j
    d = torch.from_numpy(lnps.copy()).cuda()
    m, _ = mass(s, d)
    m0 = O.mass_of(o, lnps)
    assert abs(m.item() / m0 - 1) < 1e-13
    target = m0 * (1 - 2e-4)
    m, lnr = mass(s, d, fix=True, target=target)
    fixed = d.cpu().numpy()
    ref, rlnr = O.mass_fix(o, lnps, target)
    assert abs(lnr.item() - rlnr) < 1e-15
    assert np.array_equal(fixed[1:], lnps[1:]) and abs(fixed[0] - ref[0]) < 1e-13
    m2, _ = mass(s, d)
    assert abs(m2.item() / target - 1) < 1e-13
