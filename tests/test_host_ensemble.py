"""Host-side logic of the ensemble and climate drivers, on the CPU: the seeded rest states'
noise amplitude (checked against the oracle's synthesis, not against the Parseval formula
the code uses), the Held-Suarez seed's coefficient set, and the climate comparison's
measures."""
import math

import numpy as np

from oracle.oracle import Oracle, gauss, grid_size


class _Sht:
    """The attributes held_suarez.hs_perturb reads, for a triangular truncation."""

    def __init__(self, T):
        self.mmax = self.lmax = T
        self.nlm = (T + 1) * (T + 2) // 2

    def lm(self, l, m):
        return (m * (2 * self.lmax + 2 - (m + 1))) // 2 + l + 1


def test_rest_states_noise_rms_by_synthesis():
    """rest_states scales the vorticity noise with sum_m w_m |c|^2 / 4 pi; the grid rms from
    the oracle's synthesis and Gauss weights must equal the requested 1e-6 s^-1."""
    from aeros_amd.ensemble import rest_states
    T, L = 21, 3
    st = rest_states([(T, L, 2)], noise=1.0e-6)[0]
    vor, div, temp, lnps = st
    assert vor.shape == (2, L, (T + 1) * (T + 2) // 2) and lnps.shape == (2, (T + 1) * (T + 2) // 2)
    assert np.all(div == 0)
    o = Oracle(T)
    nlon, nlat = grid_size(T)
    _, wg = gauss(nlat)
    ms, wsum = 0.0, 0.0
    for m in range(2):
        for k in range(L):
            g = o.synth(vor[m, k])     # (nlat, nlon), Gauss latitudes
            ms += float((g ** 2 * wg[:, None]).sum()) / nlon
            wsum += float(np.sum(wg))
    rms = math.sqrt(ms / wsum)
    assert abs(rms / 1.0e-6 - 1.0) <= 1e-10, rms
    # degrees above 20 carry no noise, l = 0 none either, and T, p_s are the isothermal rest state
    from aeros_amd.seeding import lm_tables
    l, _ = lm_tables(T)
    assert np.all(vor[..., l > 20] == 0) and np.all(vor[..., l == 0] == 0)
    assert np.allclose(temp[..., 0].real, 300.0 * math.sqrt(4 * math.pi))
    assert np.allclose(lnps[:, 0].real, math.log(1.0e5) * math.sqrt(4 * math.pi))


def test_hs_perturb_touches_exactly_the_seeded_coefficients():
    """aeros_hs_perturb (src/physics/aeros_held_suarez.f90:295-336) seeds m = 1..8,
    l = m..m+4 on every level and nothing else; the amplitude scales linearly."""
    from aeros_amd.held_suarez import hs_perturb
    T, L = 21, 4
    s = _Sht(T)
    a = hs_perturb(np.zeros((L, s.nlm), complex), s, 1.0)
    b = hs_perturb(np.zeros((L, s.nlm), complex), s, 2.5)
    want = np.zeros(s.nlm, bool)
    for m in range(1, 9):
        for l in range(m, m + 5):
            want[s.lm(l, m) - 1] = True
    assert np.array_equal(np.abs(a).max(axis=0) > 0, want)
    assert np.allclose(b, 2.5 * a)
    # (cos x, sin y) with different arguments: each part within [-1, 1]
    assert np.all(np.abs(a[:, want].real) <= 1.0) and np.all(np.abs(a[:, want].imag) <= 1.0)


def test_climate_comparison_measures():
    from aeros_amd.held_suarez import compare
    nlat, nlev = 8, 3
    base = {"temp": np.zeros((nlat, nlev)), "u": np.zeros((nlat, nlev))}
    d = {"temp": np.zeros((nlat, nlev)), "u": np.zeros((nlat, nlev))}
    d["temp"][0, 0] = 0.4                 # one point in the north
    d["u"][:] = 0.1                       # a uniform offset
    c = compare(d, base)
    assert math.isclose(c["max_abs_dT_K"], 0.4) and math.isclose(c["max_abs_du_ms"], 0.1)
    assert math.isclose(c["rms_dT_K"], 0.4 / math.sqrt(nlat * nlev))
    # the hemispheric mean halves a difference at one point of one hemisphere
    assert math.isclose(c["sym_max_abs_dT_K"], 0.2) and math.isclose(c["sym_max_abs_du_ms"], 0.1)
