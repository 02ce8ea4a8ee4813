"""GPU: the column kernel (column terms + Held-Suarez forcing) against the oracle
on the same inputs, and the reference's rest-state and mass checks through the
whole GPU chain (transform half -> column kernel -> transform half)."""
import numpy as np
import pytest

from dyn_states import mass_identity, moving_state, rest_errors, rest_phis, rest_state
from helpers import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def _grid_from_gpu(torch, s, vor, div, temp, lnps):
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return s.tendency_grid_transforms(dev(vor), dev(div), dev(temp), dev(lnps))


@pytest.mark.parametrize("T,nlev,vgrid,hs", [
    (21, 12, dict(), True),
    (29, 20, dict(sigma_t=0.0, p_top=0.0), False),   # 90 x 45 grid: a partial last block
    (42, 20, dict(), True),
    (21, 64, dict(), True),                           # the level limit
])
def test_column_terms_match_oracle(torch, T, nlev, vgrid, hs):
    from aeros_amd.column import VGrid, column_terms
    from aeros_amd.spectral import AerosSht
    from oracle import oracle as O
    s = AerosSht(T)
    A, B = O.vgrid_init(nlev, **vgrid)
    vg = VGrid(A, B)
    vor, div, temp, lnps = moving_state(T, s.nlm, nlev)
    g = _grid_from_gpu(torch, s, vor, div, temp, lnps)
    phis = (torch.rand((s.nlat, s.nlon), dtype=torch.float64, device="cuda") * 2000.0).contiguous()
    out = column_terms(s, vg, g, phis=phis, held_suarez=hs)
    torch.cuda.synchronize()
    h = {k: v.cpu().numpy() for k, v in g.items()}
    ref = O.column_terms(A, B, s.sinlat, h["u"], h["v"], h["vor_g"], h["div_g"], h["t_g"], h["dtdx"],
                         h["dtdy"], h["lnps_g"], h["dlnpsdx"], h["dlnpsdy"], phis.cpu().numpy(), hs=hs)
    for name, r in zip(("ae", "an", "ek", "dtdt", "dlnpsdt"), ref):
        got = out[name].cpu().numpy()
        # FP64 on both sides; they differ only in the last bits of exp/log/pow
        # and in fused multiply-adds
        assert rel(got, r) < 1e-13, (name, rel(got, r))


@pytest.mark.parametrize("case,vgrid,topo,tol,kmin", [
    ("flat, pure sigma", dict(sigma_t=0.0, p_top=0.0), False, 1e-10, 1),
    ("topography, hybrid default", dict(), True, 1e-9, 1),
    ("topography, pure sigma, p_top = 0", dict(sigma_t=0.0, p_top=0.0), True, 1e-9, 2),
])
def test_rest_state_balance_on_gpu(torch, case, vgrid, topo, tol, kmin):
    """tests/test_tendency.f90:77-107, every stage on the GPU (T21 L12). The flat case's
    bound is the FP64 analysis floor of a constant Phi+K (see test_oracle_column)."""
    from aeros_amd.column import VGrid, column_terms
    from aeros_amd.spectral import AerosSht
    from oracle import oracle as O
    T, nlev = 21, 12
    s = AerosSht(T)
    A, B = O.vgrid_init(nlev, **vgrid)
    vg = VGrid(A, B)
    vor, div, temp, lnps = rest_state(T, s.nlm, nlev, topo)
    g = _grid_from_gpu(torch, s, vor, div, temp, lnps)
    phis = torch.from_numpy(rest_phis(g["lnps_g"].cpu().numpy())).cuda()
    c = column_terms(s, vg, g, phis=phis)
    t = s.tendency_spectral(c["ae"], c["an"], c["ek"], c["dtdt"], c["dlnpsdt"])
    torch.cuda.synchronize()
    e = rest_errors(c["ek"].cpu().numpy(), t["vor"].cpu().numpy(), t["div"].cpu().numpy(),
                    t["temp"].cpu().numpy(), t["lnps"].cpu().numpy(), kmin)
    assert max(e.values()) < tol, (case, e)


def test_mass_identity_on_gpu(torch):
    """tests/test_tendency.f90:376-396 on the GPU chain (T42 L20, hybrid default, HS on)."""
    from aeros_amd.column import VGrid, column_terms
    from aeros_amd.spectral import AerosSht
    from oracle import oracle as O
    T, nlev = 42, 20
    s = AerosSht(T)
    vg = VGrid(*O.vgrid_init(nlev))
    vor, div, temp, lnps = moving_state(T, s.nlm, nlev)
    g = _grid_from_gpu(torch, s, vor, div, temp, lnps)
    c = column_terms(s, vg, g, held_suarez=True)
    torch.cuda.synchronize()
    assert mass_identity(c["dlnpsdt"].cpu().numpy(), g["lnps_g"].cpu().numpy(), s.gauss_w, s.nlon) < 1e-12


def test_vgrid_validation(torch):
    from aeros_amd._lib import AerError
    from aeros_amd.column import VGrid
    from oracle import oracle as O
    with pytest.raises(AerError):
        VGrid(*O.vgrid_init(65))              # beyond the level limit
    A, B = O.vgrid_init(10)
    with pytest.raises(AerError):
        VGrid(A[::-1].copy(), B[::-1].copy())  # pressures must increase downward
