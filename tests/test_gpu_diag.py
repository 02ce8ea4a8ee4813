"""GPU zonal-mean statistics (aer_diag_*) against the restatement of
aeros_diag_accumulate, and a short Held-Suarez integration through the driver."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a visible MI355X"
    return t


@pytest.mark.parametrize("T,nlev", [(21, 4), (42, 3)])
def test_zonal_stats_match_oracle(torch, T, nlev):
    from aeros_amd.diagnostics import ZonalStats
    from aeros_amd.spectral import AerosSht
    from oracle.oracle import zonal_sample
    s = AerosSht(T)
    rng = np.random.default_rng(T)
    zs = ZonalStats(s, nlev)
    samples = []
    for n in range(2):
        u = rng.normal(5.0, 10.0, (nlev, s.nlat, s.nlon))
        v = rng.normal(0.0, 5.0, (nlev, s.nlat, s.nlon)) + 0.3 * u
        t = rng.normal(260.0, 15.0, (nlev, s.nlat, s.nlon)) + 0.1 * v
        lp = np.log(1e5) + rng.normal(0.0, 0.01, (s.nlat, s.nlon))
        g = {"u": torch.from_numpy(u).cuda(), "v": torch.from_numpy(v).cuda(), "t_g": torch.from_numpy(t).cuda(),
             "lnps_g": torch.from_numpy(lp).cuda()}
        zs.accumulate(g, float(n))
        samples.append(zonal_sample(u, v, t, lp))
    A = np.zeros(nlev + 1)
    B = np.linspace(0.0, 1.0, nlev + 1)
    d = zs.finalize(A, B)
    assert d["nsample"] == 2
    for k in ("u", "v", "temp", "uv", "vt", "eke", "tvar"):
        want = 0.5 * (samples[0][k] + samples[1][k]).T
        assert rel(d[k], want) < 1e-13, k
    for k in ("ps", "u_sfc"):
        assert rel(d[k], 0.5 * (samples[0][k] + samples[1][k])) < 1e-13, k
    assert np.allclose(d["pressure"][:, -1], 0.5 * (B[-1] + B[-2]) * d["ps"], rtol=1e-15)
    zs.reset()
    zs.accumulate(g, 9.0)
    d1 = zs.finalize(A, B)
    assert d1["nsample"] == 1 and rel(d1["u"], samples[1]["u"].T) < 1e-13


def test_held_suarez_short_run_is_sane(torch):
    """Ten days of the Held-Suarez protocol at T21L10: finite, the isothermal start warms
    toward the equilibrium at low latitudes, the zonal-mean wind stays far below 100 m/s."""
    from aeros_amd.held_suarez import run
    sm, d, info = run(trunc=21, nlev=10, dt=1800.0, days=10.0, avg_start=5.0, precision="fp64")
    print(sm, info)
    assert d["nsample"] == 6
    assert np.isfinite(d["temp"]).all() and np.isfinite(d["u"]).all()
    assert np.abs(d["u"]).max() < 100.0
    assert d["temp"][d["temp"].shape[0] // 2, -1] > 264.0   # the tropics warm from 264 K
