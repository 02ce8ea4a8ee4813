"""GPU: the hand-written adjoints of the dynamics step. Linear pieces pass the dot-product
test <A x, y> = <x, A^T y> (Euclidean inner products: grid values, real and imaginary
parts of spectral coefficients); nonlinear pieces and the whole checkpointed run match
central finite differences of the forward code along random directions; and the
gradient does not depend on the checkpoint spacing."""
import numpy as np
import pytest

from tests.dyn_states import moving_state

pytestmark = pytest.mark.gpu

GKEYS = ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy", "lnps_g", "dlnpsdx", "dlnpsdy")
CKEYS = ("ae", "an", "ek", "dtdt", "dlnpsdt")
SKEYS = ("vor", "div", "temp", "lnps")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a visible MI355X"
    return t


def _dot(a, b):
    from aeros_amd.adjoint import dot
    return dot(a, b)


def _scale(a, b):
    """sum |a||b| over every pair: the size of the terms an identity adds up."""
    import torch
    if isinstance(a, dict):
        return sum(_scale(a[k], b[k]) for k in a)
    ra = torch.view_as_real(a) if a.is_complex() else a
    rb = torch.view_as_real(b) if b.is_complex() else b
    return float((ra.abs() * rb.abs()).sum())


def _stepper(T=21, nlev=8, hybrid=False, **kw):
    from aeros_amd.column import vgrid_init
    from aeros_amd.dynamics import Stepper
    A, B = vgrid_init(nlev) if hybrid else vgrid_init(nlev, stretch_a=1.0, sigma_t=0.0, p_top=1000.0)
    args = dict(tau_diff=6 * 3600.0, eps_filter=0.06, raw_alpha=0.53, held_suarez=True, mass_fixer=True)
    args.update(kw)
    return Stepper(T, A, B, 900.0, **args)


def _moving(st, torch):
    s = st.sht
    vor, div, temp, lnps = moving_state(s.trunc, s.nlm, st.nlev)
    c = lambda a: torch.as_tensor(a, dtype=torch.complex128).cuda().contiguous()
    return {"vor": c(vor), "div": c(div), "temp": c(temp), "lnps": c(lnps)}


def _smooth_spec(rng, s, shape, amp, lmax=10, torch=None):
    """Random spectral field with degrees 1..lmax (real at m = 0)."""
    l = s.l_of_lm
    m = s.m_of_lm
    q = (rng.standard_normal(shape + (s.nlm,)) + 1j * rng.standard_normal(shape + (s.nlm,))) * amp
    q = np.where((l >= 1) & (l <= lmax), q, 0.0)
    q = np.where(m == 0, q.real + 0j, q)
    return torch.as_tensor(q, dtype=torch.complex128).cuda().contiguous()


def test_column_adjoint_matches_finite_differences(torch):
    """column_terms_adjoint against central differences of column_terms, Held-Suarez forcing
    with an additive T_eq and a hybrid vertical grid, for the ten grid inputs and dteq."""
    from aeros_amd.column import column_terms, column_terms_adjoint
    st = _stepper(hybrid=True)
    s = st.sht
    x = _moving(st, torch)
    g0 = s.tendency_grid_transforms(x["vor"], x["div"], x["temp"], x["lnps"])
    rng = np.random.default_rng(5)
    dteq = torch.as_tensor(rng.standard_normal((st.nlev, s.nlat)), dtype=torch.float64).cuda()
    out0 = column_terms(s, st.vg, g0, held_suarez=True, dteq=dteq)
    cbar = {k: torch.as_tensor(rng.standard_normal(tuple(out0[k].shape)), dtype=torch.float64).cuda()
            for k in CKEYS}
    gbar, gteq = column_terms_adjoint(s, st.vg, g0, cbar, held_suarez=True, dteq=dteq, want_teq=True)
    # direction: each input scaled by its own spread
    d = {}
    for k in GKEYS:
        sd = float(g0[k].std()) + 1e-30
        d[k] = torch.as_tensor(rng.standard_normal(tuple(g0[k].shape)), dtype=torch.float64).cuda() * sd
    dq = torch.as_tensor(rng.standard_normal((st.nlev, s.nlat)), dtype=torch.float64).cuda()

    def J(e):
        g = {k: (g0[k] + e * d[k]).contiguous() for k in GKEYS}
        c = column_terms(s, st.vg, g, held_suarez=True, dteq=(dteq + e * dq).contiguous())
        return _dot(cbar, c)

    eps = 1e-6
    fd = (J(eps) - J(-eps)) / (2 * eps)
    ad = _dot(gbar, d) + _dot(gteq, dq)
    print(f"column adjoint: fd {fd:.12e} adjoint {ad:.12e} rel {abs(fd - ad) / abs(ad):.2e}", flush=True)
    assert abs(fd - ad) <= 1e-6 * abs(ad)


@pytest.mark.parametrize("semi,diffuse,eps", [(True, True, 0.06), (False, False, 0.0), (True, False, 0.1)])
def test_leapfrog_adjoint_dot_product(torch, semi, diffuse, eps):
    from aeros_amd import _lib
    from aeros_amd._lib import check, lib
    st = _stepper(tau_diff=6 * 3600.0 if diffuse else 0.0)
    s = st.sht
    st.si.set_step(1800.0)
    rng = np.random.default_rng(11)
    shp = {"vor": (st.nlev,), "div": (st.nlev,), "temp": (st.nlev,), "lnps": ()}
    rs = lambda: {k: _smooth_spec(rng, s, shp[k], 1.0, lmax=s.trunc, torch=torch) for k in SKEYS}
    old, now, tnd, bn, bc = rs(), rs(), rs(), rs(), rs()
    nowf = {k: v.clone() for k, v in now.items()}
    new = st.si.leapfrog(old, nowf, tnd, semi_implicit=semi, diffuse=diffuse, eps_filter=eps, raw_alpha=0.53)
    flags = ((_lib.AER_STEP_SEMI_IMPLICIT if semi else 0) | (_lib.AER_STEP_DIFFUSE if diffuse else 0) |
             (_lib.AER_STEP_FILTER if eps > 0 else 0))
    go, gc, gt = ({k: torch.empty_like(v) for k, v in now.items()} for _ in range(3))
    P = lambda d: [d[k].data_ptr() for k in SKEYS]
    s._bind_stream()
    check(lib().aer_leapfrog_adj_d(st.si._h, flags, eps, 0.53, *P(bn), *P(bc), *P(go), *P(gc), *P(gt)))
    lhs = _dot(bn, new) + _dot(bc, nowf)
    rhs = _dot(go, old) + _dot(gc, now) + _dot(gt, tnd)
    scale = _scale(bn, new) + _scale(bc, nowf)
    assert abs(lhs - rhs) <= 1e-13 * scale, (lhs, rhs, abs(lhs - rhs) / scale)


def test_tendency_halves_transpose(torch):
    """The spectral->grid half and the grid->spectral half against their adjoints."""
    from aeros_amd.adjoint import HeldSuarezAdjoint
    st = _stepper()
    s = st.sht
    adj = HeldSuarezAdjoint(st)
    rng = np.random.default_rng(3)
    shp = {"vor": (st.nlev,), "div": (st.nlev,), "temp": (st.nlev,), "lnps": ()}
    x = {k: _smooth_spec(rng, s, shp[k], 1.0, lmax=s.trunc, torch=torch) for k in SKEYS}
    g = s.tendency_grid_transforms(x["vor"], x["div"], x["temp"], x["lnps"])
    gbar = {k: torch.as_tensor(rng.standard_normal(tuple(g[k].shape)), dtype=torch.float64).cuda() for k in GKEYS}
    lam = adj.tg_adjoint(gbar)
    lhs, rhs = _dot(gbar, g), _dot(lam, x)
    assert abs(lhs - rhs) <= 1e-13 * _scale(gbar, g), (lhs, rhs)
    c = {k: torch.as_tensor(rng.standard_normal(tuple(g["u" if k != "dlnpsdt" else "lnps_g"].shape)),
                            dtype=torch.float64).cuda() for k in CKEYS}
    t = s.tendency_spectral(c["ae"], c["an"], c["ek"], c["dtdt"], c["dlnpsdt"])
    tbar = {k: _smooth_spec(rng, s, shp[k], 1.0, lmax=s.trunc, torch=torch) for k in SKEYS}
    cb = adj.ts_adjoint(tbar)
    lhs, rhs = _dot(tbar, t), _dot(cb, c)
    assert abs(lhs - rhs) <= 1e-13 * _scale(tbar, t), (lhs, rhs)


def _directions(st, rng, torch):
    s = st.sht
    d = {"vor": _smooth_spec(rng, s, (st.nlev,), 2e-6, torch=torch),
         "div": _smooth_spec(rng, s, (st.nlev,), 5e-7, torch=torch),
         "temp": _smooth_spec(rng, s, (st.nlev,), 0.5, torch=torch),
         "lnps": _smooth_spec(rng, s, (), 2e-3, torch=torch)}
    d["lnps"][0] = 1e-3       # the global mean too (the mass fixer's target moves with it)
    dq = torch.as_tensor(rng.standard_normal((st.nlev, s.nlat)), dtype=torch.float64).cuda()
    return d, dq


def test_whole_run_gradient_matches_finite_differences(torch):
    """24 steps of the Held-Suarez configuration (semi-implicit, diffusion, RAW filter, mass
    fixer) from a sheared flow: <grad J, d> against central differences of J, for a
    direction in the initial state and one in dteq, with checkpoints every 5 steps."""
    from aeros_amd.adjoint import HeldSuarezAdjoint
    st = _stepper()
    x0 = _moving(st, torch)
    st.dteq = torch.zeros((st.nlev, st.sht.nlat), dtype=torch.float64, device="cuda")
    adj = HeldSuarezAdjoint(st)
    nsteps = 24
    J0, lam0, gq = adj.gradient([x0[k] for k in SKEYS], nsteps, every=5)
    rng = np.random.default_rng(17)
    d, dq = _directions(st, rng, torch)
    h = 1e-3
    Jp = adj.run([x0[k] + h * d[k] for k in SKEYS], nsteps)
    Jm = adj.run([x0[k] - h * d[k] for k in SKEYS], nsteps)
    fd, ad = (Jp - Jm) / (2 * h), _dot(lam0, d)
    print(f"state direction: J {J0:.10f} fd {fd:.10e} adjoint {ad:.10e} rel {abs(fd - ad) / abs(ad):.2e}",
          flush=True)
    assert abs(fd - ad) <= 1e-5 * abs(ad)
    hq = 0.1
    st.dteq = hq * dq
    Jp = adj.run([x0[k] for k in SKEYS], nsteps)
    st.dteq = -hq * dq
    Jm = adj.run([x0[k] for k in SKEYS], nsteps)
    st.dteq = torch.zeros_like(dq)
    fd, ad = (Jp - Jm) / (2 * hq), _dot(gq, dq)
    print(f"dteq direction: fd {fd:.10e} adjoint {ad:.10e} rel {abs(fd - ad) / abs(ad):.2e}", flush=True)
    assert abs(fd - ad) <= 1e-5 * abs(ad)


def test_gradient_independent_of_checkpoint_spacing(torch):
    from aeros_amd.adjoint import HeldSuarezAdjoint
    st = _stepper()
    x0 = _moving(st, torch)
    adj = HeldSuarezAdjoint(st)
    Ja, la, qa = adj.gradient([x0[k] for k in SKEYS], 13, every=4)
    Jb, lb, qb = adj.gradient([x0[k] for k in SKEYS], 13, every=13)
    assert Ja == Jb
    assert Ja == adj.run([x0[k] for k in SKEYS], 13)
    for k in SKEYS:
        assert torch.equal(la[k], lb[k]), k
    assert torch.equal(qa, qb)


def hs_T31L20(torch, spinup_days):
    """The Held-Suarez protocol at T31L20 (uniform sigma, 1000 Pa top, dt 1800 s, del^6 with
    a 6 h e-folding, RAW filter, mass fixer), spun up from the seeded isothermal start."""
    from aeros_amd.column import vgrid_init
    from aeros_amd.dynamics import Stepper
    from aeros_amd.held_suarez import initial_state
    A, B = vgrid_init(20, stretch_a=1.0, stretch_r=2.0, sigma_t=0.0, p_top=1000.0)
    st = Stepper(31, A, B, 1800.0, tau_diff=6 * 3600.0, ndiff=6, eps_filter=0.06, raw_alpha=0.53,
                 held_suarez=True, mass_fixer=True)
    st.set_state(*initial_state(st.sht, 20))
    for _ in range(int(spinup_days * 48)):
        st.step()
    x0 = [st.now[k].clone() for k in SKEYS]
    st.dteq = torch.zeros((20, st.sht.nlat), dtype=torch.float64, device="cuda")
    return st, x0


def test_config3_gradient_T31L20_96_steps(torch):
    """96 steps at T31L20 from a developed (60-day) Held-Suarez flow, checkpoints every 12:
    the gradient w.r.t. the initial zeta, T and the 2-D T_eq field against central
    differences with eps = 1e-6 along 16 seeded directions."""
    from aeros_amd.adjoint import HeldSuarezAdjoint, directional_check, seeded_directions
    st, x0 = hs_T31L20(torch, 60)
    adj = HeldSuarezAdjoint(st)
    dirs = seeded_directions(st, 16)
    J, lam, gq, res = directional_check(adj, x0, 96, dirs, eps=1e-6, every=12)
    rels = [r for _, _, r in res]
    for i, (fd, ad, r) in enumerate(res):
        print(f"direction {i:2d}: fd {fd:+.10e} adjoint {ad:+.10e} rel {r:.2e}", flush=True)
    print(f"J {J:.12f}  max rel {max(rels):.2e}  median {sorted(rels)[8]:.2e}", flush=True)
    assert max(rels) <= 1e-5
