"""GPU: the fused job-list entry point (aer_run_jobs_d) against the oracle, and
HIP-graph capture of a fused transform set replayed several times, each replay
bit-identical to the eager result."""
import ctypes

import numpy as np
import pytest

from aeros_amd.seeding import red_spectra
from helpers import rel

pytestmark = pytest.mark.gpu
A = 6.371e6


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def job(L, kind, n, ins, outs, modes=(0, 0, 0), sc=(0.0, 0.0), acc=0, strides=(0, 0), prec=0):
    p = [t.data_ptr() for t in ins] + [0] * (3 - len(ins))
    q = [t.data_ptr() for t in outs] + [0] * (2 - len(outs))
    return L.AerJob(kind, n, *p, *q, strides[0], strides[1], *modes, sc[0], sc[1], acc, prec)


def run(L, sht, jobs):
    arr = (L.AerJob * len(jobs))(*jobs)
    sht._bind_stream()
    L.check(L.lib().aer_run_jobs_d(sht._h, len(jobs), arr))


@pytest.mark.parametrize("T", [31, 42])
def test_fused_jobs_match_oracle(torch, T):
    from aeros_amd import _lib as L
    from aeros_amd.spectral import AerosSht
    from oracle.oracle import Oracle
    s, o = AerosSht(T), Oracle(T)
    nlev = 5
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    vor = dev(red_spectra(T, nlev, 31, zero_l0=True) * 1e-5)
    div = dev(red_spectra(T, nlev, 32, zero_l0=True) * 1e-6)
    tem = dev(red_spectra(T, nlev, 33))
    G = lambda: torch.empty((nlev, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
    u, v, tg, gx, gy = G(), G(), G(), G(), G()
    run(L, s, [job(L, L.AER_JOB_SYNTH_VECTOR, nlev, [div, vor], [v, u],
                   (L.AER_MODE_S_FROM_DIV, L.AER_MODE_T_FROM_VOR, 0), (-1.0, 1.0)),
               job(L, L.AER_JOB_SYNTH_SCALAR, nlev, [tem], [tg]),
               job(L, L.AER_JOB_SYNTH_GRAD, nlev, [tem], [gy, gx], sc=(-1 / A, 1 / A))])
    for k in range(nlev):
        ru, rv = o.uv_from_vordiv(vor[k].cpu().numpy(), div[k].cpu().numpy())
        assert rel(u[k].cpu().numpy(), ru) < 1e-13 and rel(v[k].cpu().numpy(), rv) < 1e-13
        assert rel(tg[k].cpu().numpy(), o.synth(tem[k].cpu().numpy())) < 1e-13
        rx, ry = o.gradient(tem[k].cpu().numpy())
        assert rel(gx[k].cpu().numpy(), rx) < 1e-13 and rel(gy[k].cpu().numpy(), ry) < 1e-13
    # analysis: vordiv with the -lap(ek) assembly, then a second call that accumulates
    ek = (tg * 10.0 + u * u).contiguous()
    S = lambda: torch.empty((nlev, s.nlm), dtype=torch.complex128, device="cuda")
    tv, td, tt = S(), S(), S()
    run(L, s, [job(L, L.AER_JOB_ANAL_VECTOR, nlev, [v, u, ek], [td, tv],
                   (L.AER_MODE_DIV_FROM_S, L.AER_MODE_VOR_FROM_T, L.AER_MODE_NEG_LAP), (-1.0, 1.0)),
               job(L, L.AER_JOB_ANAL_SCALAR, nlev, [tg], [tt])])
    tt0 = tt.clone()
    run(L, s, [job(L, L.AER_JOB_ANAL_SCALAR, nlev, [gx], [tt], acc=1)])
    for k in range(nlev):
        uk, vk, ekk = u[k].cpu().numpy(), v[k].cpu().numpy(), ek[k].cpu().numpy()
        rvo, rdi = o.vordiv_from_uv(uk, vk)
        rdi = rdi - o.laplacian(o.anal(ekk))
        assert rel(tv[k].cpu().numpy(), rvo) < 1e-13
        assert rel(td[k].cpu().numpy(), rdi) < 1e-13
        ra = o.anal(tg[k].cpu().numpy())
        assert rel(tt0[k].cpu().numpy(), ra) < 1e-13
        assert rel(tt[k].cpu().numpy(), ra + o.anal(gx[k].cpu().numpy())) < 1e-13


def test_job_validation(torch):
    from aeros_amd import _lib as L
    from aeros_amd.spectral import AerosSht
    s = AerosSht(21)
    x = torch.zeros((1, s.nlm), dtype=torch.complex128, device="cuda")
    g = torch.zeros((1, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
    mixed = [job(L, L.AER_JOB_SYNTH_SCALAR, 1, [x], [g]), job(L, L.AER_JOB_ANAL_SCALAR, 1, [g], [x])]
    with pytest.raises(L.AerError):
        run(L, s, mixed)
    with pytest.raises(L.AerError):
        run(L, s, [job(L, L.AER_JOB_SYNTH_SCALAR, 1, [x], [g], strides=(10, 0))])
    with pytest.raises(L.AerError):
        run(L, s, [job(L, 9, 1, [x], [g])])
    with pytest.raises(L.AerError):
        run(L, s, [job(L, L.AER_JOB_SYNTH_SCALAR, 1, [x], [g], prec=7)])


def _mixed_case(torch, L, s, T, nlev):
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    vor = dev(red_spectra(T, nlev, 51, zero_l0=True) * 1e-5)
    div = dev(red_spectra(T, nlev, 52, zero_l0=True) * 1e-6)
    tem = dev(red_spectra(T, nlev, 53))
    G = lambda: torch.empty((nlev, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
    S = lambda: torch.empty((nlev, s.nlm), dtype=torch.complex128, device="cuda")
    bufs = {"u": G(), "v": G(), "tg": G(), "gx": G(), "gy": G(), "tv": S(), "td": S(), "tt": S()}

    def syn(precs):
        return [job(L, L.AER_JOB_SYNTH_VECTOR, nlev, [div, vor], [bufs["v"], bufs["u"]],
                    (L.AER_MODE_S_FROM_DIV, L.AER_MODE_T_FROM_VOR, 0), (-1.0, 1.0), prec=precs[0]),
                job(L, L.AER_JOB_SYNTH_SCALAR, nlev, [tem], [bufs["tg"]], prec=precs[1]),
                job(L, L.AER_JOB_SYNTH_GRAD, nlev, [tem], [bufs["gy"], bufs["gx"]], sc=(-1 / A, 1 / A), prec=precs[2])]

    def ana(precs):
        ek = (bufs["tg"] * 10.0).contiguous()
        bufs["ek"] = ek
        return [job(L, L.AER_JOB_ANAL_VECTOR, nlev, [bufs["v"], bufs["u"], ek], [bufs["td"], bufs["tv"]],
                    (L.AER_MODE_DIV_FROM_S, L.AER_MODE_VOR_FROM_T, L.AER_MODE_NEG_LAP), (-1.0, 1.0), prec=precs[0]),
                job(L, L.AER_JOB_ANAL_SCALAR, nlev, [bufs["gx"]], [bufs["tt"]], prec=precs[1])]
    return bufs, syn, ana


@pytest.mark.parametrize("T", [31, 42])
def test_mixed_precision_launch_is_bit_identical(torch, T):
    """One launch with FP64 and FP32 jobs gives, bit for bit, what the same jobs give
    in single-precision launches (each job runs its own precision's code path)."""
    from aeros_amd import _lib as L
    from aeros_amd.spectral import AerosSht
    F64, F32 = L.AER_JOB_PREC_FP64, L.AER_JOB_PREC_FP32
    s = AerosSht(T, precision="fp32")
    nlev = 6
    bufs, syn, ana = _mixed_case(torch, L, s, T, nlev)
    outs = {}
    for tag, ps, pa in [("mixed", (F32, F64, F32), (F64, F32)),
                        ("f32", (F32, F32, F32), (F32, F32)),
                        ("f64", (F64, F64, F64), (F64, F64))]:
        run(L, s, syn(ps))
        torch.cuda.synchronize()
        got = {k: bufs[k].clone() for k in ("u", "v", "tg", "gx", "gy")}
        # the analysis always reads the FP64-synthesized grids so that only its own precision differs
        if tag == "mixed":
            run(L, s, syn((F64, F64, F64)))
        run(L, s, ana(pa))
        torch.cuda.synchronize()
        got.update({k: bufs[k].clone() for k in ("tv", "td", "tt")})
        outs[tag] = got
    m, f32, f64 = outs["mixed"], outs["f32"], outs["f64"]
    for k in ("u", "v", "gx", "gy"):
        assert torch.equal(m[k], f32[k]), k
    assert torch.equal(m["tg"], f64["tg"])
    assert torch.equal(m["tv"], f64["tv"]) and torch.equal(m["td"], f64["td"])
    # the FP32 analysis in the mixed launch read FP64-synthesized grids, as the f64 run did;
    # compare it with an FP32 launch on those same grids
    run(L, s, syn((F64, F64, F64)))
    run(L, s, [ana((F64, F32))[1]])
    torch.cuda.synchronize()
    assert torch.equal(m["tt"], bufs["tt"])
    assert rel(m["tt"].cpu().numpy(), f64["tt"].cpu().numpy()) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replay_is_bit_identical(torch, precision):
    from aeros_amd import _lib as L
    from aeros_amd.spectral import AerosSht
    T, nlev = 42, 8
    s = AerosSht(T, precision=precision)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    q = dev(red_spectra(T, nlev, 41))
    gin = torch.empty((nlev, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
    qout = torch.empty((nlev, s.nlm), dtype=torch.complex128, device="cuda")
    syn = [job(L, L.AER_JOB_SYNTH_SCALAR, nlev, [q], [gin])]
    ana = [job(L, L.AER_JOB_ANAL_SCALAR, nlev, [gin], [qout])]

    def step():
        run(L, s, syn)
        run(L, s, ana)

    step()
    torch.cuda.synchronize()
    eager_g, eager_q = gin.clone(), qout.clone()
    gr = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gr):
        step()
    torch.cuda.synchronize()
    s._bind_stream()
    for rep in range(3):
        gin.zero_()
        qout.zero_()
        junk = torch.randn(1000, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        gr.replay()
        torch.cuda.synchronize()
        assert torch.equal(gin, eager_g), f"replay {rep} grid"
        assert torch.equal(qout, eager_q), f"replay {rep} spectra"
        del junk
