"""GPU parity of the tiled kernel path (fused Legendre+FFT synthesis; analysis FFT of
4 fields x 4 latitude pairs + per-(m, T-m) Legendre blocks) against the FP64 oracle
and against the staged path (separate Legendre and FFT launches), for every
transform kind, both precisions, odd latitude counts and partial tiles."""
import numpy as np
import pytest

from aeros_amd.seeding import red_spectra, red_spectrum
from helpers import rel

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 1e-6


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a visible MI355X"
    return t


_cache = {}


def pair(T, nlat=0):
    from aeros_amd.spectral import AerosSht
    from oracle.oracle import Oracle
    key = (T, nlat)
    if key not in _cache:
        s = AerosSht(T, nlat=nlat)
        _cache[key] = (s, Oracle(T, nlat=nlat) if nlat else Oracle(T))
    return _cache[key]


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def grid_batch(o, T, n, seed):
    out = []
    for k in range(n):
        a = o.synth(red_spectrum(T, seed + 2 * k))
        b = o.synth(red_spectrum(T, seed + 2 * k + 1))
        out.append(a + 0.3 * a * b + 0.1)
    return np.stack(out)


def run_all(torch, s, T, n, o):
    """Every transform kind on n fields; returns a dict of host arrays."""
    q = red_spectra(T, n, config=11)
    sl = red_spectra(T, n, config=12, zero_l0=True)
    tl = red_spectra(T, n, config=13, zero_l0=True)
    g1 = grid_batch(o, T, n, 5000)
    g2 = grid_batch(o, T, n, 7000)
    r = {}
    r["synth"] = s.synthesis(dev(torch, q)).cpu().numpy()
    r["anal"] = s.analysis(dev(torch, g1)).cpu().numpy()
    vt, vp = s.synthesis_vec(dev(torch, sl), dev(torch, tl))
    r["vt"], r["vp"] = vt.cpu().numpy(), vp.cpu().numpy()
    gt, gp = s.gradient(dev(torch, q))
    r["gt"], r["gp"] = gt.cpu().numpy(), gp.cpu().numpy()
    ss, tt = s.analysis_vec(dev(torch, g1), dev(torch, g2))
    r["s"], r["t"] = ss.cpu().numpy(), tt.cpu().numpy()
    r["synth_adj"] = s.synthesis_adjoint(dev(torch, g1)).cpu().numpy()
    r["anal_adj"] = s.analysis_adjoint(dev(torch, q)).cpu().numpy()
    torch.cuda.synchronize()
    return r, (q, sl, tl, g1, g2)


def oracle_all(o, inputs, k):
    q, sl, tl, g1, g2 = inputs
    rvt, rvp = o.synth_vec(sl[k], tl[k])
    rgt, rgp = o.synth_grad(q[k])
    rs, rt = o.anal_vec(g1[k], g2[k])
    return {"synth": o.synth(q[k]), "anal": o.anal(g1[k]), "vt": rvt, "vp": rvp, "gt": rgt, "gp": rgp,
            "s": rs, "t": rt}


# partial tiles: 5 fields (a 4-field group plus one), odd latitude counts, every fixed length
CASES = [(5, 0, 3), (21, 0, 5), (31, 0, 6), (42, 0, 7), (42, 63, 3), (31, 47, 2), (53, 0, 2), (63, 0, 2), (85, 0, 2), (106, 0, 3)]


@pytest.mark.parametrize("T,nlat,n", CASES)
def test_tiled_fp64_matches_oracle_and_staged(torch, T, nlat, n):
    s, o = pair(T, nlat)
    print(f"tiled fp64 T{T} grid {s.nlon}x{s.nlat} nfield {n}", flush=True)
    s.set_precision("fp64")
    s.set_path("staged")
    ref, inputs = run_all(torch, s, T, n, o)
    s.set_path("tiled")
    got, _ = run_all(torch, s, T, n, o)
    for key in ref:
        for k in range(n):
            assert rel(got[key][k], ref[key][k]) < 1e-14, (key, k)
    for k in range(n):
        want = oracle_all(o, inputs, k)
        for key, v in want.items():
            assert rel(got[key][k], v) < TOL64, (key, k)


@pytest.mark.parametrize("T,nlat,n", CASES)
def test_tiled_fp32_error(torch, T, nlat, n):
    s, o = pair(T, nlat)
    print(f"tiled fp32 T{T} grid {s.nlon}x{s.nlat} nfield {n}", flush=True)
    s.set_path("tiled")
    s.set_precision("fp32")
    try:
        got, inputs = run_all(torch, s, T, n, o)
    finally:
        s.set_precision("fp64")
    worst = 0.0
    for k in range(n):
        want = oracle_all(o, inputs, k)
        for key, v in want.items():
            e = rel(got[key][k], v)
            worst = max(worst, e)
            assert e < TOL32, (key, k, e)
    print(f"  fp32 worst rel err {worst:.2e}")


def test_tiled_dynamics_entry_points(torch):
    """The tendency halves (mode factors, the Phi+K term folded into the vector
    analysis, accumulate) on the tiled path equal the staged path."""
    T, L = 42, 6
    s, o = pair(T)
    vor = dev(torch, red_spectra(T, L, 21, zero_l0=True) * 1e-5)
    div = dev(torch, red_spectra(T, L, 22, zero_l0=True) * 1e-6)
    tmp = dev(torch, red_spectra(T, L, 23) * 5.0)
    lnps = dev(torch, red_spectrum(T, 24) * 1e-2)
    gb = grid_batch(o, T, 4 * L + 1, 9000)
    ae, an, ek, dtdt = (dev(torch, gb[i * L:(i + 1) * L]) for i in range(4))
    dl = dev(torch, gb[4 * L])
    outs = {}
    for path in ("staged", "tiled"):
        s.set_path(path)
        g = s.tendency_grid_transforms(vor, div, tmp, lnps)
        t = s.tendency_spectral(ae, an, ek, dtdt, dl)
        torch.cuda.synchronize()
        outs[path] = ({k: v.cpu().numpy() for k, v in g.items()}, {k: v.cpu().numpy() for k, v in t.items()})
    s.set_path("tiled")
    for a, b in zip(outs["staged"], outs["tiled"]):
        for key in a:
            assert rel(b[key], a[key]) < 1e-13, key


def test_tiled_repeat_bit_identical(torch):
    T, n = 42, 16
    s, _ = pair(T)
    s.set_path("tiled")
    dq = dev(torch, red_spectra(T, n, config=5))
    g1 = s.synthesis(dq)
    g2 = s.synthesis(dq)
    assert torch.equal(g1, g2)
    q1 = s.analysis(g1)
    q2 = s.analysis(g1)
    assert torch.equal(q1, q2)


def test_autotune_jobs_keeps_results(torch):
    """aer_autotune_jobs times every candidate configuration, keeps one, and the jobs'
    results under the kept configuration equal the staged path's."""
    import ctypes
    from aeros_amd import _lib as L
    T, n = 31, 6
    s, o = pair(T)
    q = dev(torch, red_spectra(T, n, config=31))
    g = dev(torch, grid_batch(o, T, n, 11000))
    out_g = torch.empty((n, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
    out_q = torch.empty((n, s.nlm), dtype=torch.complex128, device="cuda")
    mk = lambda kind, src, dst: L.AerJob(kind, n, src.data_ptr(), 0, 0, dst.data_ptr(), 0, 0, 0,
                                         0, 0, 0, 0.0, 0.0, 0, 0)
    js = (L.AerJob * 1)(mk(L.AER_JOB_SYNTH_SCALAR, q, out_g))
    ja = (L.AerJob * 1)(mk(L.AER_JOB_ANAL_SCALAR, g, out_q))
    s.set_precision("fp64")
    s.set_path("staged")
    s._bind_stream()
    L.check(L.lib().aer_run_jobs_d(s._h, 1, js))
    L.check(L.lib().aer_run_jobs_d(s._h, 1, ja))
    torch.cuda.synchronize()
    ref_g, ref_q = out_g.clone(), out_q.clone()
    for jl in (js, ja):
        t = (ctypes.c_double * L.AER_TUNE_NCAND)()
        c = ctypes.c_int(-1)
        L.check(L.lib().aer_autotune_jobs(s._h, 1, jl, 3, t, ctypes.byref(c)))
        assert 0 <= c.value < L.AER_TUNE_NCAND
        assert all(t[k] > 0 for k in range(L.AER_TUNE_NCAND))
        print("autotune ms", [round(t[k], 4) for k in range(L.AER_TUNE_NCAND)], "kept", c.value)
    out_g.zero_()
    out_q.zero_()
    L.check(L.lib().aer_run_jobs_d(s._h, 1, js))
    L.check(L.lib().aer_run_jobs_d(s._h, 1, ja))
    torch.cuda.synchronize()
    assert rel(out_g.cpu().numpy(), ref_g.cpu().numpy()) < 1e-14
    assert rel(out_q.cpu().numpy(), ref_q.cpu().numpy()) < 1e-14
    s.set_path("staged")
    assert L.lib().aer_autotune_jobs(s._h, 1, js, 0, None, None) == L.AER_EINVAL


def test_plan_autotune_keeps_results(torch):
    """aer_plan_autotune over a two-truncation plan: same outputs as before tuning."""
    import ctypes
    from aeros_amd import _lib as L
    members = []
    for T, n in ((21, 3), (42, 5)):
        s, o = pair(T)
        s.set_path("staged")
        s.set_precision("fp64")
        q = dev(torch, red_spectra(T, n, config=41 + T))
        g = dev(torch, grid_batch(o, T, n, 12000 + T))
        og = torch.zeros((n, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
        oq = torch.zeros((n, s.nlm), dtype=torch.complex128, device="cuda")
        members.append((s, q, g, og, oq, n))
    cfgs, jobs = [], []
    for s, q, g, og, oq, n in members:
        jobs.append(L.AerJob(L.AER_JOB_SYNTH_SCALAR, n, q.data_ptr(), 0, 0, og.data_ptr(), 0, 0, 0,
                             0, 0, 0, 0.0, 0.0, 0, 0))
        jobs.append(L.AerJob(L.AER_JOB_ANAL_SCALAR, n, g.data_ptr(), 0, 0, oq.data_ptr(), 0, 0, 0,
                             0, 0, 0, 0.0, 0.0, 0, 0))
        cfgs += [s._h, s._h]
    k = len(jobs)
    plan = ctypes.c_void_p()
    L.check(L.lib().aer_plan_create(k, (ctypes.c_void_p * k)(*cfgs), (L.AerJob * k)(*jobs), ctypes.byref(plan)))
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    try:
        L.check(L.lib().aer_plan_run(plan, st))
        torch.cuda.synchronize()
        ref = [(og.clone(), oq.clone()) for _, _, _, og, oq, _ in members]
        t = (ctypes.c_double * (2 * L.AER_TUNE_NCAND))()
        c = (ctypes.c_int * 2)(-1, -1)
        L.check(L.lib().aer_plan_autotune(plan, st, 3, t, c))
        assert all(0 <= c[i] < L.AER_TUNE_NCAND for i in range(2))
        for _, _, _, og, oq, _ in members:
            og.zero_()
            oq.zero_()
        L.check(L.lib().aer_plan_run(plan, st))
        torch.cuda.synchronize()
        for (og0, oq0), (_, _, _, og, oq, _) in zip(ref, members):
            assert rel(og.cpu().numpy(), og0.cpu().numpy()) < 1e-14
            assert rel(oq.cpu().numpy(), oq0.cpu().numpy()) < 1e-14
    finally:
        L.lib().aer_plan_destroy(plan)
        for s, *_ in members:
            s.set_path("staged")
