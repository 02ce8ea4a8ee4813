"""GPU parity: the HIP transforms against the FP64 oracle and the mpmath golden
vectors, through the C ABI. The bar for the FP64 plain form is the reference's
measure max|x - x_ref| / max|x_ref| <= 1e-13 per field (tests/test_spectral.f90:81)."""
import os

import numpy as np
import pytest

from aeros_amd.seeding import fill_test_spectrum, red_spectra, red_spectrum
from helpers import rel

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a visible MI355X"
    return t


_cache = {}


def sht(T, precision="fp64"):
    from aeros_amd.spectral import AerosSht
    key = T
    if key not in _cache:
        _cache[key] = AerosSht(T)
    s = _cache[key]
    s.set_precision(precision)
    return s


def orc(T):
    from oracle.oracle import Oracle
    key = ("orc", T)
    if key not in _cache:
        _cache[key] = Oracle(T)
    return _cache[key]


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def grid_from(T, seed):
    """A non-band-limited grid field: a synthesized red field times another, plus a constant."""
    o = orc(T)
    a = o.synth(red_spectrum(T, seed))
    b = o.synth(red_spectrum(T, seed + 1))
    return a + 0.3 * a * b + 0.1


# ------------------------------------------------------------------ geometry


@pytest.mark.parametrize("T", [21, 42, 106])
def test_geometry_matches_oracle(torch, T):
    s, o = sht(T), orc(T)
    assert (s.nlon, s.nlat, s.nlm) == (o.nlon, o.nlat, o.nlm)
    assert np.array_equal(s.sinlat, o.mu)
    assert np.abs(s.gauss_w - o.w).max() / o.w.max() < 1e-15
    assert np.array_equal(s.l_of_lm, o.l_of_lm) and np.array_equal(s.m_of_lm, o.m_of_lm)


# ------------------------------------------------------------------ golden vectors


@pytest.mark.parametrize("T", [5, 21])
def test_golden_vectors(torch, T):
    z = np.load(os.path.join(GOLD, f"sht_T{T}.npz"))
    s = sht(T)
    assert rel(s.synthesis(z["spec_fill"]), z["grid_fill"]) < 1e-14
    assert rel(s.synthesis(z["spec_red"]), z["grid_red"]) < 1e-14
    assert rel(s.analysis(z["anal_grid"]), z["anal_spec"]) < 1e-14
    vt, vp = s.synthesis_vec(z["vec_s"], z["vec_t"])
    assert rel(vt, z["vec_vt"]) < 1e-14 and rel(vp, z["vec_vp"]) < 1e-14
    gt, gp = s.gradient(z["spec_fill"])
    assert rel(gt, z["grad_gt"]) < 1e-14 and rel(gp, z["grad_gp"]) < 1e-14


# ------------------------------------------------------------------ per-call (host arrays)


@pytest.mark.parametrize("T", [5, 21, 29, 31, 42, 106])
def test_per_call_matches_oracle(torch, T):
    s, o = sht(T), orc(T)
    q = red_spectrum(T, 100 + T)
    print(f"T{T} grid {s.nlon}x{s.nlat}", flush=True)
    assert rel(s.synthesis(q), o.synth(q)) < TOL64
    g = grid_from(T, 200 + T)
    g_copy = g.copy()
    assert rel(s.analysis(g), o.anal(g)) < TOL64
    assert np.array_equal(g, g_copy)  # input preserved
    sl = red_spectrum(T, 300 + T, zero_l0=True)
    tl = red_spectrum(T, 400 + T, zero_l0=True)
    vt, vp = s.synthesis_vec(sl, tl)
    rvt, rvp = o.synth_vec(sl, tl)
    assert rel(vt, rvt) < TOL64 and rel(vp, rvp) < TOL64
    g2 = grid_from(T, 500 + T)
    s2, t2 = s.analysis_vec(g, g2)
    rs2, rt2 = o.anal_vec(g, g2)
    assert rel(s2, rs2) < TOL64 and rel(t2, rt2) < TOL64
    gt, gp = s.gradient(q)
    rgt, rgp = o.synth_grad(q)
    assert rel(gt, rgt) < TOL64 and rel(gp, rgp) < TOL64


# ------------------------------------------------------------------ batched (device arrays)


@pytest.mark.parametrize("T,nlev", [(21, 16), (31, 20), (42, 20), (106, 3)])
def test_batched_matches_oracle(torch, T, nlev):
    s, o = sht(T), orc(T)
    q = red_spectra(T, nlev, config=1)
    g = s.synthesis(dev(torch, q)).cpu().numpy()
    for k in range(nlev):
        assert rel(g[k], o.synth(q[k])) < TOL64
    gin = np.stack([grid_from(T, 1000 + k) for k in range(nlev)])
    qa = s.analysis(dev(torch, gin)).cpu().numpy()
    for k in range(nlev):
        assert rel(qa[k], o.anal(gin[k])) < TOL64
    sl = red_spectra(T, nlev, config=2, zero_l0=True)
    tl = red_spectra(T, nlev, config=3, zero_l0=True)
    vt, vp = (x.cpu().numpy() for x in s.synthesis_vec(dev(torch, sl), dev(torch, tl)))
    gt, gp = (x.cpu().numpy() for x in s.gradient(dev(torch, q)))
    s2, t2 = (x.cpu().numpy() for x in s.analysis_vec(dev(torch, vt), dev(torch, gin)))
    for k in range(nlev):
        rvt, rvp = o.synth_vec(sl[k], tl[k])
        assert rel(vt[k], rvt) < TOL64 and rel(vp[k], rvp) < TOL64
        rgt, rgp = o.synth_grad(q[k])
        assert rel(gt[k], rgt) < TOL64 and rel(gp[k], rgp) < TOL64
        rs, rt = o.anal_vec(vt[k], gin[k])
        assert rel(s2[k], rs) < TOL64 and rel(t2[k], rt) < TOL64


def test_batched_strided_fields(torch):
    """Field strides larger than the dense size (a sub-view of a bigger state)."""
    from aeros_amd import _lib
    T = 31
    s, o = sht(T), orc(T)
    q = red_spectra(T, 5, config=4)
    qpad = np.zeros((5, s.nlm + 7), dtype=np.complex128)
    qpad[:, : s.nlm] = q
    dq = dev(torch, qpad)
    gpad = torch.zeros((5, s.nlat * s.nlon + 33), dtype=torch.float64, device="cuda")
    s._bind_stream()
    _lib.check(_lib.lib().aer_synth_d(s._h, 5, dq.data_ptr(), 2 * (s.nlm + 7), gpad.data_ptr(),
                                      s.nlat * s.nlon + 33))
    g = gpad.cpu().numpy()
    for k in range(5):
        assert rel(g[k, : s.nlat * s.nlon].reshape(s.nlat, s.nlon), o.synth(q[k])) < TOL64
        assert np.all(g[k, s.nlat * s.nlon:] == 0)  # padding untouched


# ------------------------------------------------------------------ reference acceptance tests


def test_reference_spectral_acceptance(torch):
    """tests/test_spectral.f90 at T21: round trips < 20 eps, quadrature, Laplacian exact."""
    T = 21
    s = sht(T)
    eps = np.finfo(float).eps
    q = fill_test_spectrum(T)
    assert rel(s.analysis(s.synthesis(q)), q) < 20 * eps
    a = 6.371e6
    assert abs(s.surface_integral(np.ones((s.nlat, s.nlon))) - 4 * np.pi * a * a) / (4 * np.pi * a * a) < 20 * eps
    y10 = np.zeros(s.nlm, complex)
    y10[s.lm(1, 0) - 1] = 1
    assert abs(s.surface_integral(s.synthesis(y10))) / (4 * np.pi * a * a) < 20 * eps
    lq = s.laplacian(q)
    l = s.l_of_lm
    assert np.array_equal(lq, q * (-(l * (l + 1)).astype(float) / (a * a)))
    sl, tl = q.copy(), q.copy()
    sl[0] = tl[0] = 0
    s2, t2 = s.analysis_vec(*s.synthesis_vec(sl, tl))
    assert max(np.abs(s2 - sl).max(), np.abs(t2 - tl).max()) / np.abs(sl).max() < 20 * eps


def test_repeat_calls_bit_identical(torch):
    """Deterministic kernels: the pool bit-identity of tests/test_spectral.f90:209-283."""
    T, nlev = 42, 16
    s = sht(T)
    dq = dev(torch, red_spectra(T, nlev, config=5))
    g1 = s.synthesis(dq)
    g2 = s.synthesis(dq)
    assert torch.equal(g1, g2)
    q1 = s.analysis(g1)
    q2 = s.analysis(g1)
    assert torch.equal(q1, q2)
    # a batch of one equals the same field inside a batch
    g3 = s.synthesis(dq[3:4].contiguous())
    assert torch.equal(g3[0], g1[3])


@pytest.mark.parametrize("T", [21, 42, 106])
def test_round_trip_full_size(torch, T):
    """Size-independent property at full sizes: analysis(synthesis(q)) = q."""
    nlev = 20 if T < 100 else 8
    s = sht(T)
    q = red_spectra(T, nlev, config=6)
    q2 = s.analysis(s.synthesis(dev(torch, q))).cpu().numpy()
    assert max(rel(q2[k], q[k]) for k in range(nlev)) < 1e-13


# ------------------------------------------------------------------ mixed precision


def test_fp32_error_is_stated(torch):
    T, nlev = 42, 20
    s, o = sht(T, "fp32"), orc(T)
    try:
        q = red_spectra(T, nlev, config=7)
        g = s.synthesis(dev(torch, q)).cpu().numpy()
        gin = np.stack([grid_from(T, 2000 + k) for k in range(nlev)])
        qa = s.analysis(dev(torch, gin)).cpu().numpy()
        es = max(rel(g[k], o.synth(q[k])) for k in range(nlev))
        ea = max(rel(qa[k], o.anal(gin[k])) for k in range(nlev))
        print(f"fp32: synthesis max rel err {es:.2e}, analysis {ea:.2e}")
        assert es < 1e-6 and ea < 1e-6
    finally:
        s.set_precision("fp64")


@pytest.mark.parametrize("T,nlev", [(21, 8), (42, 20), (106, 3)])
def test_split_error_is_stated(torch, T, nlev):
    """Split-FP32 form (FP32 MFMA on exact 10-bit integer slices, FP64 promotion per
    step, FP64 FFT): every transform kind against the FP64 oracle. The bound is the
    measured error with a margin; the slices carry 50 bits, so the representation
    error is 2^-51 of each row/column scale."""
    s, o = sht(T, "split"), orc(T)
    try:
        q = red_spectra(T, nlev, config=8)
        t = red_spectra(T, nlev, config=9, zero_l0=True)
        gq = s.synthesis(dev(torch, q)).cpu().numpy()
        gin = np.stack([grid_from(T, 3000 + k) for k in range(nlev)])
        qa = s.analysis(dev(torch, gin)).cpu().numpy()
        vt, vp = (x.cpu().numpy() for x in s.synthesis_vec(dev(torch, t), dev(torch, q)))
        ss, tt = (x.cpu().numpy() for x in s.analysis_vec(dev(torch, gin), dev(torch, gin[::-1].copy())))
        gt, gp = (x.cpu().numpy() for x in s.gradient(dev(torch, q)))
        errs = {
            "synth": max(rel(gq[k], o.synth(q[k])) for k in range(nlev)),
            "anal": max(rel(qa[k], o.anal(gin[k])) for k in range(nlev)),
            "synth_vec": max(max(rel(a, b) for a, b in zip((vt[k], vp[k]), o.synth_vec(t[k], q[k]))) for k in range(nlev)),
            "anal_vec": max(max(rel(a, b) for a, b in zip((ss[k], tt[k]), o.anal_vec(gin[k], gin[::-1][k]))) for k in range(nlev)),
            "gradient": max(max(rel(a, b) for a, b in zip((gt[k], gp[k]), o.synth_grad(q[k]))) for k in range(nlev)),
        }
        print(f"split T{T}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        assert max(errs.values()) < 1e-12, errs
    finally:
        s.set_precision("fp64")
