"""Pin the FP64 CPU oracle (oracle/sht_oracle.c) before anything is compared with it.

Checks come from outside the oracle: 50-digit mpmath brute force written by
tests/golden/make_golden.py, numpy/scipy library routines, closed forms, and
the reference's own acceptance tests (cited per test). Each check is chosen so
a plausible mistake (a dropped term, a wrong sign, the Condon-Shortley phase,
a transposed operand) fails it; test_planted_mistakes_are_caught shows that.
"""
import os

import numpy as np
import pytest
import scipy.special

from helpers import lm_tables, rel
from oracle.oracle import Oracle, gauss, grid_size, legendre, lm_index, rfft_forward, rfft_inverse

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def golden(T):
    return np.load(os.path.join(GOLD, f"sht_T{T}.npz"))


# ---------------------------------------------------------------- grid and index


def test_grid_sizes_match_reference_table():
    # tests/test_grid.f90:25-37
    assert grid_size(21) == (64, 32)
    assert grid_size(31) == (96, 48)
    assert grid_size(42) == (128, 64)
    assert grid_size(85) == (256, 128)
    assert grid_size(106) == (320, 160)


def test_grid_size_rule_brute_force():
    for T in range(1, 200):
        nlon, nlat = grid_size(T)
        ok = [n for n in range(3 * T + 1, 4 * T + 40)
              if n % 2 == 0 and all(p in (2, 3, 5) for p in _factors(n))]
        assert nlon == ok[0] and nlat == nlon // 2


def _factors(n):
    out, p = [], 2
    while n > 1:
        while n % p == 0:
            out.append(p)
            n //= p
        p += 1
    return out


def test_lm_index_enumeration():
    # tests/test_spectral.f90:53-68: the index must agree with m-major, l-fastest enumeration
    for T in (5, 21, 42):
        k = 0
        for m in range(T + 1):
            for l in range(m, T + 1):
                k += 1
                assert lm_index(T, l, m) == k
        assert k == (T + 1) * (T + 2) // 2
        assert lm_index(T, T + 1, 0) == -1
        assert lm_index(T, 2, 3) == -1


# ---------------------------------------------------------------- Gauss-Legendre


@pytest.mark.parametrize("n", [8, 32, 45, 48, 64, 160])
def test_gauss_against_numpy(n):
    mu, w = gauss(n)
    x, wx = np.polynomial.legendre.leggauss(n)
    assert np.abs(mu - x[::-1]).max() < 2e-15
    # numpy's own polar weights carry ~1e-11 relative error at n = 160
    assert np.abs(w / wx[::-1] - 1).max() < 5e-11
    assert abs(w.sum() - 2.0) < 1e-14
    assert np.all(np.diff(mu) < 0)          # north -> south
    assert np.array_equal(mu, -mu[::-1])    # exactly antisymmetric
    assert np.array_equal(w, w[::-1])       # exactly symmetric (tests/test_grid.f90:70-76)


@pytest.mark.parametrize("n", [64, 160])
def test_gauss_weights_against_mpmath_legendre(n):
    import mpmath as mp
    mp.mp.dps = 40
    mu, w = gauss(n)
    for i in (0, 1, 2, n // 4, n // 2 - 1):
        z = mp.mpf(mu[i])
        for _ in range(5):
            p = mp.legendre(n, z)
            z -= p / (n * (z * p - mp.legendre(n - 1, z)) / (z * z - 1))
        W = 2 * (1 - z * z) / (n * mp.legendre(n - 1, z)) ** 2
        assert abs(mu[i] - z) < 1e-16
        assert abs(w[i] - W) / W < 5e-13


@pytest.mark.parametrize("T", [5, 21])
def test_gauss_against_mpmath(T):
    z = golden(T)
    mu, w = gauss(int(z["nlat"]))
    assert np.abs(mu - z["mu"]).max() < 2e-16
    assert rel(w, z["w"]) < 3e-15


# ---------------------------------------------------------------- Legendre tables


@pytest.mark.parametrize("T", [5, 21])
def test_legendre_against_mpmath(T):
    z = golden(T)
    P, D = legendre(T, z["mu"])
    assert rel(P, z["P"]) < 1e-14
    assert rel(D, z["D"]) < 1e-14


def test_legendre_against_scipy_without_cs_phase():
    T = 42
    mu, _ = gauss(64)
    P, _ = legendre(T, mu)
    theta = np.arccos(mu)
    ls, ms = lm_tables(T)
    for k in range(len(ls)):
        y = scipy.special.sph_harm_y(int(ls[k]), int(ms[k]), theta, 0.0).real
        # scipy includes (-1)^m; the reference's convention has no CS phase
        assert np.abs(P[k] - (-1.0) ** ms[k] * y).max() < 1e-13


def test_legendre_closed_forms():
    mu, _ = gauss(32)
    P, D = legendre(21, mu)
    s = np.sqrt(1 - mu ** 2)
    assert np.abs(P[0] - 1 / np.sqrt(4 * np.pi)).max() < 1e-16              # Y_00 = 1/sqrt(4 pi)
    assert np.abs(P[1] - np.sqrt(3 / (4 * np.pi)) * mu).max() < 1e-15       # Y_10 = N cos(theta)
    assert np.abs(D[1] + np.sqrt(3 / (4 * np.pi)) * s).max() < 4e-15        # d/dtheta
    assert np.all(P[lm_index(21, 5, 5) - 1] > 0)                            # no CS phase: P_mm > 0


def test_legendre_derivative_by_central_difference():
    T = 31
    theta = np.linspace(0.2, 2.9, 17)
    h = 1e-6
    P0, D0 = legendre(T, np.cos(theta))
    Pp, _ = legendre(T, np.cos(theta + h))
    Pm, _ = legendre(T, np.cos(theta - h))
    fd = (Pp - Pm) / (2 * h)
    assert np.abs(fd - D0).max() < 1e-7 * np.abs(D0).max()


def test_orthonormality_by_quadrature():
    T = 42
    mu, w = gauss(64)
    P, _ = legendre(T, mu)
    ls, ms = lm_tables(T)
    for m in (0, 7, 42):
        sel = ms == m
        G = 2 * np.pi * (P[sel] * w) @ P[sel].T
        assert np.abs(G - np.eye(sel.sum())).max() < 1e-13


# ---------------------------------------------------------------- FFT


@pytest.mark.parametrize("n", [16, 64, 96, 128, 320])
def test_rfft_against_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    mmax = (n - 1) // 3
    X = rfft_forward(x, mmax)
    assert np.abs(X - np.fft.rfft(x)[: mmax + 1]).max() < 1e-13 * n
    Y = rng.standard_normal(mmax + 1) + 1j * rng.standard_normal(mmax + 1)
    Y[0] = Y[0].real
    full = np.zeros(n // 2 + 1, dtype=complex)
    full[: mmax + 1] = Y
    ref = np.fft.irfft(full, n) * n   # = Re X0 + 2 sum_{m>=1} Re(X_m e^{i m phi})
    assert np.abs(rfft_inverse(Y, n) - ref).max() < 1e-13 * n


def test_fft_matches_direct_dft():
    from aeros_amd.seeding import red_spectrum
    for T in (21, 31, 42):
        q = red_spectrum(T, 7 + T)
        assert rel(Oracle(T).synth(q), Oracle(T, fft_direct=True).synth(q)) < 1e-14


# ---------------------------------------------------------------- scalar transforms


@pytest.mark.parametrize("T", [5, 21])
def test_scalar_transforms_against_mpmath(T):
    z = golden(T)
    o = Oracle(T)
    assert rel(o.synth(z["spec_fill"]), z["grid_fill"]) < 2e-15
    assert rel(o.synth(z["spec_red"]), z["grid_red"]) < 3e-15
    assert rel(o.anal(z["anal_grid"]), z["anal_spec"]) < 3e-15


@pytest.mark.parametrize("T", [5, 21])
def test_vector_and_gradient_against_mpmath(T):
    z = golden(T)
    o = Oracle(T)
    vt, vp = o.synth_vec(z["vec_s"], z["vec_t"])
    assert rel(vt, z["vec_vt"]) < 5e-15 and rel(vp, z["vec_vp"]) < 5e-15
    gt, gp = o.synth_grad(z["spec_fill"])
    assert rel(gt, z["grad_gt"]) < 5e-15 and rel(gp, z["grad_gp"]) < 5e-15


def test_y00_and_parseval():
    o = Oracle(21)
    q = np.zeros(o.nlm, complex)
    q[0] = 1.0
    assert np.abs(o.synth(q) - 1 / np.sqrt(4 * np.pi)).max() < 1e-15
    from aeros_amd.seeding import red_spectrum
    q = red_spectrum(21, 3)
    g = o.synth(q)
    # int f^2 dOmega = sum_{m=0} |c|^2 + 2 sum_{m>0} |c|^2 (src/dynamics/aeros_timestep.f90:92-96)
    lhs = o.surface_integral(g * g) / 6.371e6 ** 2
    rhs = np.sum(np.abs(q[o.m_of_lm == 0]) ** 2) + 2 * np.sum(np.abs(q[o.m_of_lm > 0]) ** 2)
    assert abs(lhs - rhs) / rhs < 1e-14


@pytest.mark.parametrize("T", [21, 29, 42])
def test_round_trips(T):
    from aeros_amd.seeding import fill_test_spectrum
    o = Oracle(T)
    q = fill_test_spectrum(T)
    assert rel(o.anal(o.synth(q)), q) < 20 * np.finfo(float).eps   # tests/test_spectral.f90:75-83
    s, t = q.copy(), 0.5 * np.conj(q)
    t[o.m_of_lm == 0] = t[o.m_of_lm == 0].real
    s[0] = t[0] = 0
    s2, t2 = o.anal_vec(*o.synth_vec(s, t))
    assert max(np.abs(s2 - s).max(), np.abs(t2 - t).max()) / max(np.abs(s).max(), np.abs(t).max()) < 20 * np.finfo(float).eps


def test_alias_free_product():
    # nlon >= 3T+1: the product of two T-fields analyses exactly to degree T
    from aeros_amd.seeding import red_spectrum
    T = 21
    o = Oracle(T)
    o2 = Oracle(T, nlon=256, nlat=128)   # a finer grid resolves the product exactly
    a, b = red_spectrum(T, 1), red_spectrum(T, 2)
    ref = Oracle(T, nlon=256, nlat=128).anal(o2.synth(a) * o2.synth(b))
    assert rel(o.anal(o.synth(a) * o.synth(b)), ref) < 1e-14


# ---------------------------------------------------------------- operators


def test_laplacian_exact_and_inverse():
    # tests/test_spectral.f90:110-124
    from aeros_amd.seeding import fill_test_spectrum
    o = Oracle(21)
    q = fill_test_spectrum(21)
    lq = o.laplacian(q)
    l = o.l_of_lm
    assert np.array_equal(lq, q * (-(l * (l + 1)).astype(float) / 6.371e6 ** 2))
    back = o.laplacian(lq, inverse=True)
    q0 = q.copy()
    q0[0] = 0
    assert rel(back, q0) < 1e-12


def test_surface_integral():
    # tests/test_spectral.f90:89-104
    o = Oracle(21)
    a = 6.371e6
    assert abs(o.surface_integral(np.ones((o.nlat, o.nlon))) - 4 * np.pi * a * a) / (4 * np.pi * a * a) < 20 * np.finfo(float).eps
    q = np.zeros(o.nlm, complex)
    q[1] = 1.0
    assert abs(o.surface_integral(o.synth(q))) / (4 * np.pi * a * a) < 20 * np.finfo(float).eps


# ---------------------------------------------------------------- discriminating power


def test_planted_mistakes_are_caught():
    """Mistakes planted in oracle OUTPUTS (on the CPU) must fail the golden checks."""
    z = golden(21)
    o = Oracle(21)
    g = o.synth(z["spec_fill"])
    ls, ms = lm_tables(21)
    # Condon-Shortley phase slipped in
    q_cs = z["spec_fill"] * (-1.0) ** ms
    assert rel(o.synth(q_cs), z["grid_fill"]) > 1e-3
    # missing factor 2 on m > 0 (Parseval weight)
    g_half = (g + o.synth(np.where(ms == 0, z["spec_fill"], 0))) / 2
    assert rel(g_half, z["grid_fill"]) > 1e-3
    # latitude order flipped (south -> north)
    assert rel(g[::-1], z["grid_fill"]) > 1e-3
    # toroidal sign flipped in the vector synthesis
    vt, vp = o.synth_vec(z["vec_s"], -z["vec_t"])
    assert rel(vp, z["vec_vp"]) > 1e-3
    # weights not applied in the analysis: compare the unweighted sum
    bad = o.anal(z["anal_grid"] * 2.0 / o.w[:, None] / o.nlat)
    assert rel(bad, z["anal_spec"]) > 1e-3


# ------------------------------------------------- the CPU-baseline arrangement

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 12, 20, 45, 48, 60, 64, 80, 96, 160, 240])
def test_port_fft_against_numpy(n):
    """The CPU port's Stockham FFT (radix 4, 2, 3, 5 passes) against numpy.fft, both signs."""
    from oracle.oracle import port_cfft
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(port_cfft(x, -1), np.fft.fft(x)) < 1e-14
    assert rel(port_cfft(x, +1), np.fft.ifft(x) * n) < 1e-14
    with pytest.raises(ValueError):
        port_cfft(np.ones(7, complex), -1)


@pytest.mark.parametrize("T", [5, 21, 31, 42, 106])
def test_port_transforms_match_the_oracle(T):
    """The CPU port (hemispheric folding, parity-split tables, half-length real FFT) is
    the oracle's algorithm rearranged: synthesis and analysis agree with orc_synth /
    orc_anal to rounding, including the m = 0 imaginary part that synthesis drops."""
    from oracle.oracle import Port
    o = Oracle(T)
    p = Port(o)
    rng = np.random.default_rng(T)
    q = rng.standard_normal(o.nlm) + 1j * rng.standard_normal(o.nlm)   # m = 0 imaginary parts not zeroed
    g0 = o.synth(q)
    assert rel(p.synth(q), g0) < 1e-14
    assert rel(p.anal(g0), o.anal(g0)) < 1e-14
    # a grid that is not band-limited: the projections still agree
    g = rng.standard_normal((o.nlat, o.nlon))
    assert rel(p.anal(g), o.anal(g)) < 1e-13


def test_port_rejects_grids_it_cannot_fold():
    """An odd latitude count has no pairing of every row; the port refuses it."""
    from oracle.oracle import Port
    with pytest.raises(ValueError):
        Port(Oracle(21, nlon=64, nlat=33))
