"""GPU parity of the dynamics-side transforms: aeros_vordiv and the tendency
transform halves (src/dynamics/aeros_vordiv.f90, src/dynamics/aeros_tendency.f90:403-470),
restating tests/test_vordiv.f90 and comparing with the FP64 oracle."""
import numpy as np
import pytest

from aeros_amd.seeding import red_spectra, red_spectrum
from helpers import rel

pytestmark = pytest.mark.gpu

A = 6.371e6
TOL_VORDIV = 1e3 * np.finfo(float).eps   # tests/test_vordiv.f90:36


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def t21():
    from aeros_amd.spectral import AerosSht
    from oracle.oracle import Oracle
    return AerosSht(21), Oracle(21)


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_solid_body_rotation(torch, t21):
    """tests/test_vordiv.f90:59-143: u = U0 cos(lat), v = 0 -> zeta_10 = (2 U0/a)/sqrt(3/4pi), D = 0."""
    s, _ = t21
    u0 = 40.0
    u = u0 * np.sin(s.colat)[:, None] * np.ones((1, s.nlon))
    v = np.zeros_like(u)
    vor, div = s.vordiv_from_uv(dev(torch, u[None]), dev(torch, v[None]))
    vor, div = vor.cpu().numpy()[0], div.cpu().numpy()[0]
    assert np.abs(div).max() / np.abs(vor).max() < TOL_VORDIV
    lm10 = s.lm(1, 0) - 1
    expect = (2 * u0 / A) / np.sqrt(3 / (4 * np.pi))
    assert abs(vor[lm10].real - expect) / expect < TOL_VORDIV
    rest = vor.copy()
    rest[lm10] = 0
    assert np.abs(rest).max() / expect < TOL_VORDIV
    zg = s.synthesis(dev(torch, vor[None])).cpu().numpy()[0]
    scale = 2 * u0 / A
    assert np.abs(zg - scale * s.sinlat[:, None]).max() / scale < TOL_VORDIV
    assert zg[0, 0] > 0 and zg[-1, 0] < 0   # cyclonic in the north


def test_pure_divergence_and_round_trip(torch, t21):
    """tests/test_vordiv.f90:145-243."""
    s, _ = t21
    vor = np.zeros(s.nlm, complex)
    div = np.zeros(s.nlm, complex)
    div[s.lm(2, 1) - 1] = 1e-6 - 4e-7j
    u, v = s.uv_from_vordiv(dev(torch, vor[None]), dev(torch, div[None]))
    vor2, div2 = (x.cpu().numpy()[0] for x in s.vordiv_from_uv(u, v))
    assert np.abs(vor2).max() / np.abs(div).max() < TOL_VORDIV
    assert np.abs(div2 - div).max() / np.abs(div).max() < TOL_VORDIV
    l, m = s.l_of_lm, s.m_of_lm
    amp = np.where(l > 0, 1e-5 / np.maximum(l, 1) ** 2, 0.0)
    vor = amp * np.cos(3.0 * l + m) + 1j * amp * np.sin(l + 2.0 * m)
    div = 0.3 * amp * np.sin(l + 1.0 * m) + 1j * 0.3 * amp * np.cos(2.0 * l + m)
    vor[m == 0] = vor[m == 0].real
    div[m == 0] = div[m == 0].real
    u, v = s.uv_from_vordiv(dev(torch, vor[None]), dev(torch, div[None]))
    vor2, div2 = (x.cpu().numpy()[0] for x in s.vordiv_from_uv(u, v))
    scale = max(np.abs(vor).max(), np.abs(div).max())
    assert max(np.abs(vor2 - vor).max(), np.abs(div2 - div).max()) / scale < TOL_VORDIV
    umax = np.abs(u.cpu().numpy()).max()
    assert 1.0 < umax < 200.0


def test_l0_discarded_bit_for_bit(torch, t21):
    """tests/test_vordiv.f90:245-289."""
    s, _ = t21
    vor = np.zeros(s.nlm, complex)
    div = np.zeros(s.nlm, complex)
    vor[s.lm(3, 2) - 1] = 1e-5 + 2e-6j
    u, v = s.uv_from_vordiv(dev(torch, vor[None]), dev(torch, div[None]))
    vrt, drt = s.vordiv_from_uv(u, v)
    assert vrt.cpu().numpy()[0, 0] == 0 and drt.cpu().numpy()[0, 0] == 0
    vor[0] = 1e-3
    div[0] = 1e-3
    u2, v2 = s.uv_from_vordiv(dev(torch, vor[None]), dev(torch, div[None]))
    assert torch.equal(u, u2) and torch.equal(v, v2)


@pytest.mark.parametrize("T", [21, 42])
def test_vordiv_and_gradient_match_oracle(torch, T):
    from aeros_amd.spectral import AerosSht
    from oracle.oracle import Oracle
    s, o = AerosSht(T), Oracle(T)
    nlev = 6
    vor = red_spectra(T, nlev, config=11) * 1e-5
    div = red_spectra(T, nlev, config=12) * 1e-6
    u, v = (x.cpu().numpy() for x in s.uv_from_vordiv(dev(torch, vor), dev(torch, div)))
    vor2, div2 = (x.cpu().numpy() for x in s.vordiv_from_uv(dev(torch, u), dev(torch, v)))
    dx, dy = (x.cpu().numpy() for x in s.grad_phys(dev(torch, vor)))
    for k in range(nlev):
        ru, rv = o.uv_from_vordiv(vor[k], div[k])
        assert rel(u[k], ru) < 1e-13 and rel(v[k], rv) < 1e-13
        rvo, rdi = o.vordiv_from_uv(u[k], v[k])
        assert rel(vor2[k], rvo) < 1e-13 and rel(div2[k], rdi) < 1e-13
        rdx, rdy = o.gradient(vor[k])
        assert rel(dx[k], rdx) < 1e-13 and rel(dy[k], rdy) < 1e-13


@pytest.mark.parametrize("T,nlev", [(31, 20), (42, 20)])
def test_tendency_halves_match_oracle(torch, T, nlev):
    from aeros_amd.spectral import AerosSht
    from oracle.oracle import Oracle
    s, o = AerosSht(T), Oracle(T)
    vor = red_spectra(T, nlev, config=21) * 1e-5
    div = red_spectra(T, nlev, config=22) * 1e-6
    temp = red_spectra(T, nlev, config=23) * 10.0
    lnps = red_spectrum(T, 24) * 0.01
    out = s.tendency_grid_transforms(dev(torch, vor), dev(torch, div), dev(torch, temp), dev(torch, lnps))
    ref = o.tendency_grid_transforms(vor, div, temp, lnps)
    for k, r in ref.items():
        g = out[k].cpu().numpy()
        if g.ndim == 3:
            assert max(rel(g[i], r[i]) for i in range(nlev)) < 1e-13, k
        else:
            assert rel(g, r) < 1e-13, k
    # grid -> spectral: use products of the synthesized fields as nonlinear grid inputs
    ae = ref["u"] * ref["vor_g"] * 1e3 + ref["dtdx"]
    an = ref["v"] * ref["vor_g"] * 1e3 - ref["dtdy"]
    ek = 0.5 * (ref["u"] ** 2 + ref["v"] ** 2) + 9.8 * ref["t_g"]
    dtdt = -(ref["u"] * ref["dtdx"] + ref["v"] * ref["dtdy"])
    dl = -ref["div_g"].sum(axis=0) * 1e-2
    tnd = s.tendency_spectral(*(dev(torch, a) for a in (ae, an, ek, dtdt, dl)))
    rtn = o.tendency_spectral(ae, an, ek, dtdt, dl)
    for k in ("vor", "div", "temp"):
        g = tnd[k].cpu().numpy()
        assert max(rel(g[i], rtn[k][i]) for i in range(nlev)) < 1e-13, k
    assert rel(tnd["lnps"].cpu().numpy(), rtn["lnps"]) < 1e-13
