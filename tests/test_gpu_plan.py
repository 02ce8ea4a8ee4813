"""GPU: compiled plans over several truncations in one launch per stage
(the batching layer for mixed-truncation ensembles), against the oracle, against
the per-handle entry points bit for bit, and replayed from a HIP graph."""
import ctypes

import numpy as np
import pytest

from aeros_amd.seeding import red_spectra
from helpers import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def make_member(torch, L, T, nlev, cfgid, precision="fp64"):
    from aeros_amd.spectral import AerosSht
    s = AerosSht(T, precision=precision)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    m = {"sht": s, "T": T, "nlev": nlev,
         "q": dev(red_spectra(T, nlev, cfgid)),
         "vor": dev(red_spectra(T, nlev, cfgid + 1, zero_l0=True) * 1e-5),
         "div": dev(red_spectra(T, nlev, cfgid + 2, zero_l0=True) * 1e-6)}
    G = lambda: torch.zeros((nlev, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
    S = lambda: torch.zeros((nlev, s.nlm), dtype=torch.complex128, device="cuda")
    m.update(g=G(), u=G(), v=G(), qa=S(), va=S(), da=S())
    P = lambda t: t.data_ptr()
    J = lambda kind, n, ins, outs, md=(0, 0, 0), sc=(0.0, 0.0): L.AerJob(
        kind, n, *(list(map(P, ins)) + [0] * (3 - len(ins))), *(list(map(P, outs)) + [0] * (2 - len(outs))),
        0, 0, *md, *sc, 0)
    m["jobs"] = [
        J(L.AER_JOB_SYNTH_SCALAR, nlev, [m["q"]], [m["g"]]),
        J(L.AER_JOB_SYNTH_VECTOR, nlev, [m["div"], m["vor"]], [m["v"], m["u"]],
          (L.AER_MODE_S_FROM_DIV, L.AER_MODE_T_FROM_VOR, 0), (-1.0, 1.0)),
        J(L.AER_JOB_ANAL_SCALAR, nlev, [m["g"]], [m["qa"]]),
        J(L.AER_JOB_ANAL_VECTOR, nlev, [m["v"], m["u"]], [m["da"], m["va"]],
          (L.AER_MODE_DIV_FROM_S, L.AER_MODE_VOR_FROM_T, 0), (-1.0, 1.0)),
    ]
    return m


def build_plan(L, members):
    cfgs, jobs = [], []
    for m in members:
        for j in m["jobs"]:
            cfgs.append(m["sht"]._h)
            jobs.append(j)
    n = len(jobs)
    plan = ctypes.c_void_p()
    L.check(L.lib().aer_plan_create(n, (ctypes.c_void_p * n)(*cfgs), (L.AerJob * n)(*jobs), ctypes.byref(plan)))
    return plan


def outputs(members):
    return [{k: m[k].clone() for k in ("g", "u", "v", "qa", "va", "da")} for m in members]


def test_mixed_truncation_plan(torch):
    from aeros_amd import _lib as L
    from oracle.oracle import Oracle
    members = [make_member(torch, L, T, n, 60 + 3 * i)
               for i, (T, n) in enumerate([(21, 16), (31, 20), (42, 20), (21, 3), (42, 5)])]
    plan = build_plan(L, members)
    try:
        stream = torch.cuda.current_stream().cuda_stream
        L.check(L.lib().aer_plan_run(plan, ctypes.c_void_p(stream)))
        torch.cuda.synchronize()
        got = outputs(members)
        # against the oracle
        for m, o in zip(members, got):
            orc = Oracle(m["T"])
            for k in range(m["nlev"]):
                q = m["q"][k].cpu().numpy()
                assert rel(o["g"][k].cpu().numpy(), orc.synth(q)) < 1e-13
                ru, rv = orc.uv_from_vordiv(m["vor"][k].cpu().numpy(), m["div"][k].cpu().numpy())
                assert rel(o["u"][k].cpu().numpy(), ru) < 1e-13 and rel(o["v"][k].cpu().numpy(), rv) < 1e-13
                assert rel(o["qa"][k].cpu().numpy(), orc.anal(o["g"][k].cpu().numpy())) < 1e-13
                rvo, rdi = orc.vordiv_from_uv(o["u"][k].cpu().numpy(), o["v"][k].cpu().numpy())
                assert rel(o["va"][k].cpu().numpy(), rvo) < 1e-13 and rel(o["da"][k].cpu().numpy(), rdi) < 1e-13
        # against the per-handle fused entry point, bit for bit
        for m in members:
            for key in ("g", "u", "v", "qa", "va", "da"):
                m[key].zero_()
            s = m["sht"]
            s._bind_stream()
            L.check(L.lib().aer_run_jobs_d(s._h, 2, (L.AerJob * 2)(*m["jobs"][:2])))
            L.check(L.lib().aer_run_jobs_d(s._h, 2, (L.AerJob * 2)(*m["jobs"][2:])))
        torch.cuda.synchronize()
        for m, o in zip(members, got):
            for key in o:
                assert torch.equal(m[key], o[key]), key
        # graph capture of the plan, replayed with eager work in between
        gr = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gr):
            L.check(L.lib().aer_plan_run(plan, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        for rep in range(3):
            for m in members:
                for key in ("g", "u", "v", "qa", "va", "da"):
                    m[key].zero_()
            torch.cuda.synchronize()
            gr.replay()
            torch.cuda.synchronize()
            for m, o in zip(members, got):
                for key in o:
                    assert torch.equal(m[key], o[key]), (rep, key)
    finally:
        L.lib().aer_plan_destroy(plan)


def test_plan_mixes_precisions_bit_identically(torch):
    """Handles of both precisions in one plan: every member's outputs equal, bit for
    bit, its own handle's fused launches."""
    from aeros_amd import _lib as L
    members = [make_member(torch, L, T, n, 80 + 3 * i, prec)
               for i, (T, n, prec) in enumerate([(21, 4, "fp32"), (42, 6, "fp64"), (31, 5, "fp32"), (42, 3, "fp32")])]
    plan = build_plan(L, members)
    try:
        L.check(L.lib().aer_plan_run(plan, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        got = outputs(members)
        for m in members:
            for key in ("g", "u", "v", "qa", "va", "da"):
                m[key].zero_()
            s = m["sht"]
            s._bind_stream()
            L.check(L.lib().aer_run_jobs_d(s._h, 2, (L.AerJob * 2)(*m["jobs"][:2])))
            L.check(L.lib().aer_run_jobs_d(s._h, 2, (L.AerJob * 2)(*m["jobs"][2:])))
        torch.cuda.synchronize()
        for m, o in zip(members, got):
            for key in o:
                assert torch.equal(m[key], o[key]), (m["T"], key)
    finally:
        L.lib().aer_plan_destroy(plan)


def test_plan_validation(torch):
    from aeros_amd import _lib as L
    from aeros_amd.spectral import AerosSht
    a, b = AerosSht(21), AerosSht(31, precision="fp32")
    x = torch.zeros((1, a.nlm), dtype=torch.complex128, device="cuda")
    g = torch.zeros((1, a.nlat, a.nlon), dtype=torch.float64, device="cuda")
    j = L.AerJob(L.AER_JOB_SYNTH_SCALAR, 1, x.data_ptr(), 0, 0, g.data_ptr(), 0, 0, 0, 0, 0, 0, 0.0, 0.0, 0)
    bad = L.AerJob(L.AER_JOB_SYNTH_SCALAR, 1, x.data_ptr(), 0, 0, g.data_ptr(), 0, 0, 0, 0, 0, 0, 0.0, 0.0, 0, 9)
    plan = ctypes.c_void_p()
    cf = (ctypes.c_void_p * 2)(a._h, b._h)
    with pytest.raises(L.AerError):   # unknown job precision
        L.check(L.lib().aer_plan_create(2, cf, (L.AerJob * 2)(j, bad), ctypes.byref(plan)))
    with pytest.raises(L.AerError):   # empty
        L.check(L.lib().aer_plan_create(0, cf, (L.AerJob * 2)(j, j), ctypes.byref(plan)))
    # a field count beyond 2^26 is refused on the host, before any table or launch
    huge = L.AerJob(L.AER_JOB_SYNTH_SCALAR, (1 << 26) + 1, x.data_ptr(), 0, 0, g.data_ptr(), 0, 0, 0, 0, 0, 0,
                    0.0, 0.0, 0)
    with pytest.raises(L.AerError, match="2\\^26"):
        L.check(L.lib().aer_plan_create(2, cf, (L.AerJob * 2)(j, huge), ctypes.byref(plan)))
    with pytest.raises(L.AerError, match="2\\^26"):
        L.check(L.lib().aer_run_jobs_d(a._h, 1, (L.AerJob * 1)(huge)))
