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


def job(L, kind, n, ins, outs, modes=(0, 0, 0), sc=(0.0, 0.0), acc=0, strides=(0, 0)):
    p = [t.data_ptr() for t in ins] + [0] * (3 - len(ins))
    q = [t.data_ptr() for t in outs] + [0] * (2 - len(outs))
    return L.AerJob(kind, n, *p, *q, strides[0], strides[1], *modes, sc[0], sc[1], acc)


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
