"""Benchmark: spherical-harmonic transform sets per second at T42L20 on one MI355X.

One step = one S_bl transform set of a dynamics step at T42 (128x64 Gaussian
grid, lmax 42) with 20 levels, all fields resident in HBM:

  synthesis half  lnps, grad lnps; per level (u,v) from (zeta,D), zeta, D, T, q,
                  grad T                                   (aer_tendency_grid_transforms_d
                                                            + aer_synth_d for q)
  analysis half   per level (zeta_t, D_t) from (A_e, A_n) with D_t -= lap(Phi+K),
                  T_t, q_t; lnps_t                         (aer_tendency_spectral_d
                                                            + aer_anal_d for q)

= 13L+4 = 264 scalar-equivalent transforms. `value` is the batched
mixed-precision form: FP32 MFMA Legendre GEMM + FP32 FFT for every transform
except the (zeta_t, D_t - lap(Phi+K)) analysis, whose two terms nearly cancel and
which therefore runs FP64 inside the same fused launch (per-job precision). The
all-FP32 form's errors are reported beside it. Also measured on the same GPU: the batched FP64 form and the FP64 plain form (one call per level and
field, the reference's loop structure), plus the per-field max relative error
of the mixed form against FP64. Inputs are seeded synthetic states.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

T, NLEV = 42, 20
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
# kernel kinds of aer_sht_timing: 0-3 the staged path, 2-4 the tiled path
NKIND = 5
KIND_NAMES = {"staged": ["k_leg_synth", "k_fft_synth", "k_fft_anal", "k_leg_anal", None],
              "tiled": [None, None, "k_fft_anal4", "k_leg_anal4", "k_syn_fused"]}
# autotuner candidates (aer_autotune_jobs)
TUNE_NAMES = ["staged", "tiled/16KiB", "tiled/32KiB", "tiled/64KiB"]


def set_counts(L):
    """Transform counts of S_bl: scalar, vector and gradient transforms, by direction."""
    return {"syn_scalar": 4 * L + 1, "syn_vector": L, "syn_grad": L + 1,
            "ana_scalar": 3 * L + 1, "ana_vector": L}


def algorithmic(nlon, nlat, nlm, L):
    """Algorithmic bytes (FP64 I/O) and dense flops per set, per kernel kind
    (0 Legendre synthesis, 1 FFT synthesis, 2 FFT analysis, 3 Legendre analysis,
    4 the fused synthesis = 0 + 1).

    G = 8 nlon nlat, S = 16 nlm; Legendre flops per scalar contraction 4 nlat nlm;
    a vector transform is 4 contractions, a gradient 2; FFT 2.5 nlon log2(nlon) per row."""
    c = set_counts(L)
    G, S = 8 * nlon * nlat, 16 * nlm
    L1 = 4 * nlat * nlm
    F1 = nlat * 2.5 * nlon * np.log2(nlon)
    bytes_ = [S * (c["syn_scalar"] + 2 * c["syn_vector"] + c["syn_grad"]),
              G * (c["syn_scalar"] + 2 * c["syn_vector"] + 2 * c["syn_grad"]),
              G * (c["ana_scalar"] + 2 * c["ana_vector"]),
              S * (c["ana_scalar"] + 2 * c["ana_vector"])]
    flops = [L1 * (c["syn_scalar"] + 4 * c["syn_vector"] + 2 * c["syn_grad"]),
             F1 * (c["syn_scalar"] + 2 * c["syn_vector"] + 2 * c["syn_grad"]),
             F1 * (c["ana_scalar"] + 2 * c["ana_vector"]),
             L1 * (c["ana_scalar"] + 4 * c["ana_vector"])]
    bytes_.append(bytes_[0] + bytes_[1])
    flops.append(flops[0] + flops[1])
    return bytes_, flops


class TransformSet:
    """Device-resident inputs and outputs of one S_bl set, and the calls that run it."""

    def __init__(self, sht, torch, seed_config=1):
        from aeros_amd.seeding import red_spectra, red_spectrum
        self.sht, self.torch = sht, torch
        self.assembly_prec = 0   # AER_JOB_PREC_* of the (zeta_t, D_t) analysis job; 0 = the handle's
        L, nlm = NLEV, sht.nlm
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        self.vor = dev(red_spectra(T, L, seed_config, field0=0, zero_l0=True) * 1e-5)
        self.div = dev(red_spectra(T, L, seed_config, field0=100, zero_l0=True) * 1e-6)
        self.temp = dev(red_spectra(T, L, seed_config, field0=200) * 5.0)
        self.q = dev(red_spectra(T, L, seed_config, field0=300) * 1e-3)
        self.lnps = dev(red_spectrum(T, 1000 * seed_config + 400) * 1e-2)
        # grid inputs of the analysis half: nonlinear products of synthesized fields
        sht.set_precision("fp64")
        a = sht.synthesis(dev(red_spectra(T, L, seed_config, field0=500)))
        b = sht.synthesis(dev(red_spectra(T, L, seed_config, field0=600)))
        c = sht.synthesis(dev(red_spectra(T, L, seed_config, field0=700)))
        torch.cuda.synchronize()
        self.ae = (a * b * 1e-4).contiguous()
        self.an = (b * c * 1e-4 - a * 1e-5).contiguous()
        self.ek = (a * a * 50.0 + c * 100.0).contiguous()
        self.dtdt = (b * c * 1e-5).contiguous()
        self.qdot = (a * c * 1e-8).contiguous()
        self.dlnpsdt = (a[0] * 1e-6 + b[1] * c[2] * 1e-6).contiguous()
        g = lambda: torch.empty((L, sht.nlat, sht.nlon), dtype=torch.float64, device="cuda")
        g2 = lambda: torch.empty((sht.nlat, sht.nlon), dtype=torch.float64, device="cuda")
        s = lambda: torch.empty((L, nlm), dtype=torch.complex128, device="cuda")
        self.gout = {k: g() for k in ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy")}
        self.gout.update({k: g2() for k in ("lnps_g", "dlnpsdx", "dlnpsdy")})
        self.q_g = g()
        self.sout = {k: s() for k in ("vor", "div", "temp")}
        self.sout["lnps"] = torch.empty((nlm,), dtype=torch.complex128, device="cuda")
        self.q_t = s()

    def _jobs(self):
        """The set as two fused job lists: 8 synthesis jobs, 4 analysis jobs."""
        from aeros_amd import _lib as L
        o, so, a = self.gout, self.sout, 1.0 / 6.371e6
        P = lambda t: t.data_ptr()
        J = lambda kind, n, ins, outs, m=(0, 0, 0), sc=(0.0, 0.0), prec=0: L.AerJob(
            kind, n, *(list(map(P, ins)) + [0] * (3 - len(ins))), *(list(map(P, outs)) + [0] * (2 - len(outs))),
            0, 0, m[0], m[1], m[2], sc[0], sc[1], 0, prec)
        syn = [J(L.AER_JOB_SYNTH_SCALAR, 1, [self.lnps], [o["lnps_g"]]),
               J(L.AER_JOB_SYNTH_GRAD, 1, [self.lnps], [o["dlnpsdy"], o["dlnpsdx"]], sc=(-a, a)),
               J(L.AER_JOB_SYNTH_VECTOR, NLEV, [self.div, self.vor], [o["v"], o["u"]],
                 m=(L.AER_MODE_S_FROM_DIV, L.AER_MODE_T_FROM_VOR, 0), sc=(-1.0, 1.0)),
               J(L.AER_JOB_SYNTH_SCALAR, NLEV, [self.vor], [o["vor_g"]]),
               J(L.AER_JOB_SYNTH_SCALAR, NLEV, [self.div], [o["div_g"]]),
               J(L.AER_JOB_SYNTH_SCALAR, NLEV, [self.temp], [o["t_g"]]),
               J(L.AER_JOB_SYNTH_GRAD, NLEV, [self.temp], [o["dtdy"], o["dtdx"]], sc=(-a, a)),
               J(L.AER_JOB_SYNTH_SCALAR, NLEV, [self.q], [self.q_g])]
        ana = [J(L.AER_JOB_ANAL_VECTOR, NLEV, [self.an, self.ae, self.ek], [so["div"], so["vor"]],
                 m=(L.AER_MODE_DIV_FROM_S, L.AER_MODE_VOR_FROM_T, L.AER_MODE_NEG_LAP), sc=(-1.0, 1.0),
                 prec=self.assembly_prec),
               J(L.AER_JOB_ANAL_SCALAR, NLEV, [self.dtdt], [so["temp"]]),
               J(L.AER_JOB_ANAL_SCALAR, 1, [self.dlnpsdt], [so["lnps"]]),
               J(L.AER_JOB_ANAL_SCALAR, NLEV, [self.qdot], [self.q_t])]
        return (L.AerJob * len(syn))(*syn), len(syn), (L.AerJob * len(ana))(*ana), len(ana)

    def set_assembly_prec(self, prec):
        self.assembly_prec = prec
        self._jl = None

    def run_batched(self):
        """One fused launch pair per direction (the whole set = 4 kernels)."""
        from aeros_amd._lib import check, lib
        if getattr(self, "_jl", None) is None:
            self._jl = self._jobs()
        js, ns, ja, na = self._jl
        check(lib().aer_run_jobs_d(self.sht._h, ns, js))
        check(lib().aer_run_jobs_d(self.sht._h, na, ja))

    def autotune(self, reps=20):
        """Kernel configuration per direction by measurement (aer_autotune_jobs): the staged
        path, or the tiled path with an m-band staging budget of 16, 32 or 64 KiB."""
        from aeros_amd._lib import AER_TUNE_NCAND, check, lib
        if getattr(self, "_jl", None) is None:
            self._jl = self._jobs()
        js, ns, ja, na = self._jl
        out = {}
        for name, (jl, n) in (("synthesis", (js, ns)), ("analysis", (ja, na))):
            t = (ctypes.c_double * AER_TUNE_NCAND)()
            c = ctypes.c_int()
            check(lib().aer_autotune_jobs(self.sht._h, n, jl, reps, t, ctypes.byref(c)))
            out[name] = {"ms": [round(t[k], 5) for k in range(AER_TUNE_NCAND)], "kept": TUNE_NAMES[c.value]}
        self.tuned = out
        return out

    def run_calls(self):
        """The same set through the per-purpose entry points (4 calls, 8 kernels)."""
        from aeros_amd._lib import check, lib
        sh, L = self.sht, NLEV
        h, ss, gs = sh._h, 2 * sh.nlm, sh.nlon * sh.nlat
        o, so = self.gout, self.sout
        check(lib().aer_tendency_grid_transforms_d(
            h, L, self.vor.data_ptr(), self.div.data_ptr(), self.temp.data_ptr(), self.lnps.data_ptr(),
            *[o[k].data_ptr() for k in ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy",
                                        "lnps_g", "dlnpsdx", "dlnpsdy")]))
        check(lib().aer_synth_d(h, L, self.q.data_ptr(), ss, self.q_g.data_ptr(), gs))
        check(lib().aer_tendency_spectral_d(
            h, L, self.ae.data_ptr(), self.an.data_ptr(), self.ek.data_ptr(), self.dtdt.data_ptr(),
            self.dlnpsdt.data_ptr(), so["vor"].data_ptr(), so["div"].data_ptr(), so["temp"].data_ptr(),
            so["lnps"].data_ptr()))
        check(lib().aer_anal_d(h, L, self.qdot.data_ptr(), gs, self.q_t.data_ptr(), ss))

    def run_plain(self):
        """The reference's loop structure: one call per level and per transform."""
        from aeros_amd._lib import check, lib
        sh, L = self.sht, NLEV
        h, ss, gs = sh._h, 2 * sh.nlm, sh.nlon * sh.nlat
        o, so = self.gout, self.sout
        sb, gb = 16 * sh.nlm, 8 * gs   # bytes per spectral / grid field
        P = lambda t, k, nb: t.data_ptr() + k * nb
        check(lib().aer_synth_d(h, 1, self.lnps.data_ptr(), ss, o["lnps_g"].data_ptr(), gs))
        check(lib().aer_gradient_d(h, 1, self.lnps.data_ptr(), o["dlnpsdx"].data_ptr(), o["dlnpsdy"].data_ptr()))
        for k in range(L):
            check(lib().aer_uv_from_vordiv_d(h, 1, P(self.vor, k, sb), P(self.div, k, sb),
                                             P(o["u"], k, gb), P(o["v"], k, gb)))
            check(lib().aer_synth_d(h, 1, P(self.vor, k, sb), ss, P(o["vor_g"], k, gb), gs))
            check(lib().aer_synth_d(h, 1, P(self.div, k, sb), ss, P(o["div_g"], k, gb), gs))
            check(lib().aer_synth_d(h, 1, P(self.temp, k, sb), ss, P(o["t_g"], k, gb), gs))
            check(lib().aer_gradient_d(h, 1, P(self.temp, k, sb), P(o["dtdx"], k, gb), P(o["dtdy"], k, gb)))
            check(lib().aer_synth_d(h, 1, P(self.q, k, sb), ss, P(self.q_g, k, gb), gs))
        for k in range(L):
            check(lib().aer_tendency_spectral_d(
                h, 1, P(self.ae, k, gb), P(self.an, k, gb), P(self.ek, k, gb), P(self.dtdt, k, gb),
                self.dlnpsdt.data_ptr(), P(so["vor"], k, sb), P(so["div"], k, sb), P(so["temp"], k, sb),
                so["lnps"].data_ptr()))
            check(lib().aer_anal_d(h, 1, P(self.qdot, k, gb), gs, P(self.q_t, k, sb), ss))

    def outputs(self):
        out = {k: v.clone() for k, v in self.gout.items()}
        out["q_g"] = self.q_g.clone()
        out.update({"tnd_" + k: v.clone() for k, v in self.sout.items()})
        out["tnd_q"] = self.q_t.clone()
        return out


def torch_real(z):
    import torch
    return torch.view_as_real(z)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "transform_set_T42L20_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` in the measured set, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (counters cannot be read inside a timed run; the summary
    names its command). None when the kernel is not in it."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for row in f:
                if row.startswith(f"| `aer::{kernel}<"):
                    cells = [c.strip() for c in row.strip().strip("|").split("|")]
                    return round(float(cells[6]) * 1e6), os.path.relpath(TRAFFIC_PROFILE, ROOT)
    except (OSError, ValueError, IndexError):
        pass
    return None, None


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps   # ms per step


def capture_graph(torch, sht, fn):
    """Record one set into a HIP graph (torch.cuda.CUDAGraph over the library's launches)."""
    fn()                      # first call of this shape allocates the intermediate
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        sht._bind_stream()    # the capture stream
        fn()
    torch.cuda.synchronize()
    sht._bind_stream()        # back to the ambient stream
    return g


def kernel_times(sht, fn, steps):
    """Per-kernel-kind milliseconds per set and launches per set, from HIP events."""
    from aeros_amd._lib import check, lib
    check(lib().aer_sht_timing(sht._h, 1))
    for _ in range(steps):
        fn()
    ms = (ctypes.c_double * NKIND)()
    n = (ctypes.c_int * NKIND)()
    check(lib().aer_sht_timing_read(sht._h, ms, n))
    check(lib().aer_sht_timing(sht._h, 0))
    return [ms[k] / steps for k in range(NKIND)], [n[k] / steps for k in range(NKIND)]


def rel_err(a, ref):
    a = a.cpu().numpy()
    ref = ref.cpu().numpy()
    if a.ndim >= 2 and a.shape[0] == NLEV:
        return max(float(np.abs(a[k] - ref[k]).max() / np.abs(ref[k]).max()) for k in range(NLEV))
    return float(np.abs(a - ref).max() / np.abs(ref).max())


def sref_bytes(nlon, nlat, nlm, L):
    """Algorithmic bytes of one S_ref set (the reference's dry step, 11L+4 transforms)."""
    G, S = 8 * nlon * nlat, 16 * nlm
    Ns, Nv, Ng = 5 * L + 2, 2 * L, L + 1
    return Ns * (G + S) + 2 * Nv * (G + S) + Ng * (S + 2 * G)


def ensemble_config2(torch, precision, steps=50, warmup=5, path="auto"):
    """64 Held-Suarez members of uneven size (24 x T21L16, 24 x T31L20, 16 x T42L20): the
    S_ref transform set of every member in one compiled plan (4 launches), replayed from a
    HIP graph. Returns throughput and the HBM rate of the algorithmic traffic."""
    from aeros_amd import _lib as L
    from aeros_amd.seeding import red_spectra, red_spectrum
    from aeros_amd.spectral import AerosSht
    mix = [(21, 16, 24), (31, 20, 24), (42, 20, 16)]
    shts = {T: AerosSht(T, precision=precision) for T, _, _ in mix}
    for s_ in shts.values():
        s_.set_path("staged" if path == "auto" else path)
    cfgs, jobs, keep, nbytes = [], [], [], 0
    a = 1.0 / 6.371e6
    P = lambda t: t.data_ptr()
    J = lambda kind, n, ins, outs, md=(0, 0, 0), sc=(0.0, 0.0): L.AerJob(
        kind, n, *(list(map(P, ins)) + [0] * (3 - len(ins))), *(list(map(P, outs)) + [0] * (2 - len(outs))),
        0, 0, *md, *sc, 0)
    member = 0
    for T, nlev, count in mix:
        s = shts[T]
        for _ in range(count):
            dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
            vor = dev(red_spectra(T, nlev, 2, member=member, field0=0, zero_l0=True) * 1e-5)
            div = dev(red_spectra(T, nlev, 2, member=member, field0=100, zero_l0=True) * 1e-6)
            tem = dev(red_spectra(T, nlev, 2, member=member, field0=200))
            lnp = dev(red_spectrum(T, 2000 + 300 + 100000 * member) * 1e-2)
            G = lambda n=nlev: torch.empty((n, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
            S = lambda n=nlev: torch.empty((n, s.nlm), dtype=torch.complex128, device="cuda")
            u, v, zg, dg, tg, tx, ty = G(), G(), G(), G(), G(), G(), G()
            lg, lx, ly = G(1), G(1), G(1)
            ae, an, ek, dt = (torch.randn((nlev, s.nlat, s.nlon), dtype=torch.float64, device="cuda") for _ in range(4))
            dl = torch.randn((1, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
            tv, td, tt, tl = S(), S(), S(), S(1)
            keep += [vor, div, tem, lnp, u, v, zg, dg, tg, tx, ty, lg, lx, ly, ae, an, ek, dt, dl, tv, td, tt, tl]
            mj = [J(L.AER_JOB_SYNTH_SCALAR, 1, [lnp], [lg]),
                  J(L.AER_JOB_SYNTH_GRAD, 1, [lnp], [ly, lx], sc=(-a, a)),
                  J(L.AER_JOB_SYNTH_VECTOR, nlev, [div, vor], [v, u],
                    (L.AER_MODE_S_FROM_DIV, L.AER_MODE_T_FROM_VOR, 0), (-1.0, 1.0)),
                  J(L.AER_JOB_SYNTH_SCALAR, nlev, [vor], [zg]),
                  J(L.AER_JOB_SYNTH_SCALAR, nlev, [div], [dg]),
                  J(L.AER_JOB_SYNTH_SCALAR, nlev, [tem], [tg]),
                  J(L.AER_JOB_SYNTH_GRAD, nlev, [tem], [ty, tx], sc=(-a, a)),
                  J(L.AER_JOB_ANAL_VECTOR, nlev, [an, ae, ek], [td, tv],
                    (L.AER_MODE_DIV_FROM_S, L.AER_MODE_VOR_FROM_T, L.AER_MODE_NEG_LAP), (-1.0, 1.0)),
                  J(L.AER_JOB_ANAL_SCALAR, nlev, [dt], [tt]),
                  J(L.AER_JOB_ANAL_SCALAR, 1, [dl], [tl])]
            cfgs += [s._h] * len(mj)
            jobs += mj
            nbytes += sref_bytes(s.nlon, s.nlat, s.nlm, nlev)
            member += 1
    n = len(jobs)
    plan = ctypes.c_void_p()
    L.check(L.lib().aer_plan_create(n, (ctypes.c_void_p * n)(*cfgs), (L.AerJob * n)(*jobs), ctypes.byref(plan)))
    stream = lambda: ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    tuned = None
    if path == "auto":
        t = (ctypes.c_double * (2 * L.AER_TUNE_NCAND))()
        c = (ctypes.c_int * 2)()
        L.check(L.lib().aer_plan_autotune(plan, stream(), 10, t, c))
        tuned = {d: {"ms": [round(t[i * L.AER_TUNE_NCAND + k], 5) for k in range(L.AER_TUNE_NCAND)],
                     "kept": TUNE_NAMES[c[i]]} for i, d in enumerate(("synthesis", "analysis"))}
    run = lambda: L.check(L.lib().aer_plan_run(plan, stream()))
    run()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    torch.cuda.synchronize()
    ms = time_steps(torch, g.replay, steps, warmup)
    L.lib().aer_plan_destroy(plan)
    ntr = sum(count * (11 * nlev + 4) for _, nlev, count in mix)
    return {"workload": "64 members: 24 x T21L16 + 24 x T31L20 + 16 x T42L20, S_ref set (11L+4 transforms) each, "
                        "one compiled plan (3-4 launches) replayed from a HIP graph",
            "precision": precision, "jobs": n, "ms_per_step": round(ms, 4), "autotune": tuned,
            "ensemble_steps_per_s": round(1000.0 / ms, 1),
            "scalar_eq_transforms_per_s": round(ntr * 1000.0 / ms, 0),
            "alg_bytes_per_step": int(nbytes),
            "hbm_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}


def ensemble_config2_step(torch, precision="fp32", steps=50, warmup=5, days=1.0):
    """Whole Held-Suarez steps of the 64-member uneven ensemble (24 x T21L16, 24 x T31L20,
    16 x T42L20) from the seeded rest state: the transforms of all members in two autotuned
    plans, column terms, FP64 leapfrog and mass fixer batched per size, every step after
    the second a HIP-graph replay. Then `days` of integration as a stability check."""
    from aeros_amd.ensemble import EnsembleStepper, rest_states
    mix = [(21, 16, 24), (31, 20, 24), (42, 20, 16)]
    ens = EnsembleStepper(mix, dt=1200.0, precision=precision)
    tuned = ens.autotune()
    ens.set_state(rest_states(mix))
    ens.step()
    ens.step()
    ens.capture()
    ms = time_steps(torch, ens.step, steps, warmup)
    nday = int(round(days * 86400.0 / ens.dt))
    while ens.nstep < nday:
        ens.step()
    torch.cuda.synchronize()
    vmax = max(float(ens.now(i)["vor"].abs().max()) for i in range(len(mix)))
    ntr = sum(n * (11 * L + 4) for _, L, n in mix)
    nbytes = 0
    for T_, L_, n in mix:
        from aeros_amd.spectral import grid_size
        from aeros_amd.seeding import lm_tables
        nlon, nlat = grid_size(T_)
        nbytes += n * sref_bytes(nlon, nlat, lm_tables(T_)[0].size, L_)
    names = lambda c: [TUNE_NAMES[i] for i in c]
    return {"workload": "64 Held-Suarez members (24 x T21L16 + 24 x T31L20 + 16 x T42L20), whole dynamics steps: "
                        "2 fused transform plans + per-size column terms, FP64 leapfrog and mass fixer, "
                        "HIP-graph replay per step",
            "precision": precision + (" (assembly job fp64)" if precision == "fp32" else ""),
            "ms_per_ensemble_step": round(ms, 4), "member_steps_per_s": round(64 * 1000.0 / ms, 0),
            "sref_transforms_per_s": round(ntr * 1000.0 / ms, 0),
            "transform_alg_hbm_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "autotune_kept": {"synthesis": names(tuned["synthesis"][1])[0], "analysis": names(tuned["analysis"][1])[1]},
            "stability_days": days, "max_abs_vor_after": vmax}


def stress_T106(torch, members=32, nlev=30, steps=5, warmup=1):
    """The stress configuration's transform work: the S_ref set (11L+4 transforms) of 32
    T106L30 members (Gaussian 160 x 320) in one compiled plan, split-FP32 MFMA Legendre, and
    the same plan in FP64. Reports sets/s, scalar-equivalent transforms/s, the HBM rate of
    the algorithmic traffic, device memory in use, and the per-field error of the split
    form's outputs against FP64 for the whole ensemble."""
    from aeros_amd import _lib as L
    from aeros_amd.seeding import red_spectra, red_spectrum
    from aeros_amd.spectral import AerosSht
    T = 106
    free0, total = torch.cuda.mem_get_info()
    res = {"workload": f"{members} x T{T}L{nlev} (Gaussian 160 x 320), S_ref set each, one compiled plan",
           "members": members}
    outs = {}
    a = 1.0 / 6.371e6
    P = lambda t: t.data_ptr()
    J = lambda kind, n, ins, outs_, md=(0, 0, 0), sc=(0.0, 0.0): L.AerJob(
        kind, n, *(list(map(P, ins)) + [0] * (3 - len(ins))), *(list(map(P, outs_)) + [0] * (2 - len(outs_))),
        0, 0, *md, *sc, 0)
    s = AerosSht(T)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    ins, nbytes = [], 0
    for mbr in range(members):
        vor = dev(red_spectra(T, nlev, 4, member=mbr, field0=0, zero_l0=True) * 1e-5)
        div = dev(red_spectra(T, nlev, 4, member=mbr, field0=100, zero_l0=True) * 1e-6)
        tem = dev(red_spectra(T, nlev, 4, member=mbr, field0=200))
        lnp = dev(red_spectrum(T, 4000 + 300 + 100000 * mbr) * 1e-2)
        gen = torch.Generator(device="cuda").manual_seed(mbr)
        ae, an, ek, dt = (torch.randn((nlev, s.nlat, s.nlon), dtype=torch.float64, device="cuda", generator=gen)
                          for _ in range(4))
        dl = torch.randn((1, s.nlat, s.nlon), dtype=torch.float64, device="cuda", generator=gen)
        ins.append((vor, div, tem, lnp, ae, an, ek, dt, dl))
        nbytes += sref_bytes(s.nlon, s.nlat, s.nlm, nlev)
    for prec in ("fp64", "split"):
        s.set_precision(prec)
        G = lambda n=nlev: torch.empty((n, s.nlat, s.nlon), dtype=torch.float64, device="cuda")
        S = lambda n=nlev: torch.empty((n, s.nlm), dtype=torch.complex128, device="cuda")
        cfgs, jobs, o_all = [], [], []
        for (vor, div, tem, lnp, ae, an, ek, dt, dl) in ins:
            o = {"u": G(), "v": G(), "vor_g": G(), "div_g": G(), "t_g": G(), "dtdx": G(), "dtdy": G(),
                 "lnps_g": G(1), "dlnpsdx": G(1), "dlnpsdy": G(1), "tnd_vor": S(), "tnd_div": S(), "tnd_temp": S(),
                 "tnd_lnps": S(1)}
            mj = [J(L.AER_JOB_SYNTH_SCALAR, 1, [lnp], [o["lnps_g"]]),
                  J(L.AER_JOB_SYNTH_GRAD, 1, [lnp], [o["dlnpsdy"], o["dlnpsdx"]], sc=(-a, a)),
                  J(L.AER_JOB_SYNTH_VECTOR, nlev, [div, vor], [o["v"], o["u"]],
                    (L.AER_MODE_S_FROM_DIV, L.AER_MODE_T_FROM_VOR, 0), (-1.0, 1.0)),
                  J(L.AER_JOB_SYNTH_SCALAR, nlev, [vor], [o["vor_g"]]),
                  J(L.AER_JOB_SYNTH_SCALAR, nlev, [div], [o["div_g"]]),
                  J(L.AER_JOB_SYNTH_SCALAR, nlev, [tem], [o["t_g"]]),
                  J(L.AER_JOB_SYNTH_GRAD, nlev, [tem], [o["dtdy"], o["dtdx"]], sc=(-a, a)),
                  J(L.AER_JOB_ANAL_VECTOR, nlev, [an, ae, ek], [o["tnd_div"], o["tnd_vor"]],
                    (L.AER_MODE_DIV_FROM_S, L.AER_MODE_VOR_FROM_T, L.AER_MODE_NEG_LAP), (-1.0, 1.0)),
                  J(L.AER_JOB_ANAL_SCALAR, nlev, [dt], [o["tnd_temp"]]),
                  J(L.AER_JOB_ANAL_SCALAR, 1, [dl], [o["tnd_lnps"]])]
            cfgs += [s._h] * len(mj)
            jobs += mj
            o_all.append(o)
        n = len(jobs)
        plan = ctypes.c_void_p()
        L.check(L.lib().aer_plan_create(n, (ctypes.c_void_p * n)(*cfgs), (L.AerJob * n)(*jobs), ctypes.byref(plan)))
        st = lambda: ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        run = lambda: L.check(L.lib().aer_plan_run(plan, st()))
        ms = time_steps(torch, run, steps, warmup)
        free1, _ = torch.cuda.mem_get_info()
        L.lib().aer_plan_destroy(plan)
        ntr = members * (11 * nlev + 4)
        res[prec] = {"ms_per_ensemble_step": round(ms, 3), "sets_per_s": round(members * 1000.0 / ms, 1),
                     "scalar_eq_transforms_per_s": round(ntr * 1000.0 / ms, 0),
                     "hbm_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "device_memory_in_use_GB": round((total - free1) / 1e9, 2)}
        outs[prec] = o_all
    errs = {}
    for k in outs["fp64"][0]:
        errs[k] = max(rel_err_any(a_[k], b_[k]) for a_, b_ in zip(outs["split"], outs["fp64"]))
    res["split_rel_err_vs_fp64_per_field"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
    res["split_max_rel_err"] = max(errs.values())
    res["alg_bytes_per_step"] = int(nbytes)
    s.set_precision("fp64")
    return res


def stress_day_T106(torch, members=32, nlev=30, dt=600.0, hours=24.0, every=16, precisions=("split", "fp64")):
    """The stress configuration end to end: 32 T106L30 Held-Suarez members batched in one
    Stepper (every launch covers all members), one model day forward at dt = 600 s and the
    checkpointed gradient of the summed objective (checkpoints every 16 steps), in the
    split-FP32 form and in FP64. Reports ms per step, S_ref transforms/s and their HBM rate,
    peak device memory, and the split form's per-field error of the final state and of the
    gradient against the FP64 run."""
    from aeros_amd.adjoint import HeldSuarezAdjoint
    from aeros_amd.column import vgrid_init
    from aeros_amd.dynamics import Stepper
    from aeros_amd.seeding import red_spectra
    T = 106
    nsteps = int(round(hours * 3600.0 / dt))
    A, B = vgrid_init(nlev, stretch_a=1.0, stretch_r=2.0, sigma_t=0.0, p_top=1000.0)
    out, finals, grads = {}, {}, {}
    for prec in precisions:
        st = Stepper(T, A, B, dt, tau_diff=6 * 3600.0, ndiff=6, eps_filter=0.06, raw_alpha=0.53,
                     held_suarez=True, mass_fixer=True, precision=prec)
        s = st.sht
        l_of = torch.as_tensor(s.l_of_lm, device="cuda")
        # rest, isothermal 300 K, p_s = 1000 hPa, seeded vorticity noise of grid rms 1e-6 s^-1 on l <= 20
        vor = torch.as_tensor(np.stack([red_spectra(T, nlev, 4, member=m, field0=0, zero_l0=True)
                                        for m in range(members)]), dtype=torch.complex128).cuda()
        vor = torch.where(l_of <= 20, vor, torch.zeros_like(vor)).contiguous()
        s.set_precision("fp64")
        vor *= 1e-6 / float(s.synthesis(vor[0]).pow(2).mean().sqrt())
        s.set_precision(prec)
        temp = torch.zeros_like(vor)
        temp[..., 0] = 300.0 * np.sqrt(4 * np.pi)
        lnps = torch.zeros((members, s.nlm), dtype=torch.complex128, device="cuda")
        lnps[:, 0] = np.log(1.0e5) * np.sqrt(4 * np.pi)
        x0 = [vor, torch.zeros_like(vor), temp, lnps]
        st.dteq = torch.zeros((members, nlev, s.nlat), dtype=torch.float64, device="cuda")
        adj = HeldSuarezAdjoint(st)
        adj.run(x0, 2)                      # warm-up: tables, plans, buffers
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        t0 = time.perf_counter()
        J = adj.run(x0, nsteps)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        finals[prec] = {k: v.clone() for k, v in st.now.items()}
        t0 = time.perf_counter()
        Jg, lam, gq = adj.gradient(x0, nsteps, every=every)
        torch.cuda.synchronize()
        t_grad = time.perf_counter() - t0
        free, total = torch.cuda.mem_get_info()
        grads[prec] = dict(lam, teq=gq)
        ntr = members * (11 * nlev + 4) * nsteps
        nbytes = members * sref_bytes(s.nlon, s.nlat, s.nlm, nlev) * nsteps
        out[prec] = {"J_sum_K": round(J, 9), "forward_ms_per_step": round(1000 * t_fwd / nsteps, 3),
                     "forward_sref_transforms_per_s": round(ntr / t_fwd, 0),
                     "forward_transform_hbm_GBs": round(nbytes / t_fwd / 1e9, 1),
                     "gradient_s": round(t_grad, 3), "gradient_over_forward": round(t_grad / t_fwd, 2),
                     "peak_torch_memory_GB": round(torch.cuda.max_memory_allocated() / 1e9, 2),
                     "device_memory_in_use_GB": round((total - free) / 1e9, 2)}
        del st, adj
        torch.cuda.empty_cache()
    if "split" in finals and "fp64" in finals:
        out["split_final_state_rel_err_vs_fp64"] = {
            k: float(f"{rel_err_any(finals['split'][k], finals['fp64'][k]):.3e}") for k in finals["fp64"]}
        out["split_gradient_rel_err_vs_fp64"] = {
            k: float(f"{rel_err_any(grads['split'][k], grads['fp64'][k]):.3e}") for k in grads["fp64"]}
    out["workload"] = (f"{members} x T{T}L{nlev} Held-Suarez members in one batched stepper, {nsteps} steps of "
                       f"{dt:.0f} s (one day) + checkpointed gradient (every {every})")
    return out


def rel_err_any(a, ref):
    """max|a - ref| / max|ref| over a whole tensor (the reference's measure)."""
    if a.is_complex():
        a, ref = torch_real(a), torch_real(ref)
    a, ref = a.cpu().numpy(), ref.cpu().numpy()
    return float(np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300))


def held_suarez_climate(torch, trunc=31, days=1200.0, members=32):
    """The Held-Suarez protocol of drivers/held_suarez.f90 at T31L20 (1200 days, the last 1000
    averaged) for ensembles of `members` runs stepped together: an FP64 ensemble, the same
    ensemble with FP32 transforms (FP64 assembly job), and an independent FP64 ensemble whose
    difference from the first is the sampling noise of the ensemble-mean climate."""
    from aeros_amd.held_suarez import compare, run_ensemble
    out = {"members": members}
    res = {}
    for name, prec, off in (("fp64", "fp64", 0), ("fp32", "fp32", 0), ("fp64_twin", "fp64", members)):
        sm, d, info = run_ensemble(trunc=trunc, nlev=20, members=members, dt=1800.0, days=days, avg_start=200.0,
                                   precision=prec, seed_offset=off)
        res[name] = d
        out[name] = {"summary": {k: round(v, 3) for k, v in sm.items()}, "wall_s": info["wall_s"],
                     "ms_per_ensemble_step": info["ms_per_ensemble_step"]}
    out["fp32_vs_fp64"] = {k: round(v, 4) for k, v in compare(res["fp32"], res["fp64"]).items()}
    out["fp64_twin_vs_fp64"] = {k: round(v, 4) for k, v in compare(res["fp64_twin"], res["fp64"]).items()}
    out["published_T31L20"] = {"u_max": 35.3, "u_max_lat": -42.7, "vt_max": 18.7, "uv_max": 74.4, "eke_max": 268.0,
                               "t_sfc_equator": 306.8, "t_sfc_pole": 263.0,
                               "source": "reference docs/m1_results.md, T31L20 column (SH jet)"}
    return out


def adjoint_T31(torch, spinup_days=100.0, nsteps=96, every=12, ndir=16, eps=1e-6):
    """Gradient of the Held-Suarez functional (mass-weighted mean T over 30-60 N at the final
    step) through 96 T31L20 steps in FP64, checkpoints every 12, from a spun-up flow:
    timing of the forward run and of the checkpointed gradient, peak device memory, the
    dot-product residual of each transpose transform at T31, and central differences along
    16 seeded directions in (initial zeta, initial T, 2-D T_eq)."""
    from aeros_amd.adjoint import HeldSuarezAdjoint, directional_check, seeded_directions
    from aeros_amd.column import vgrid_init
    from aeros_amd.dynamics import Stepper
    from aeros_amd.held_suarez import initial_state
    from aeros_amd.seeding import red_spectra
    nlev, trunc, dt = 20, 31, 1800.0
    A, B = vgrid_init(nlev, stretch_a=1.0, stretch_r=2.0, sigma_t=0.0, p_top=1000.0)
    st = Stepper(trunc, A, B, dt, tau_diff=6 * 3600.0, ndiff=6, eps_filter=0.06, raw_alpha=0.53,
                 held_suarez=True, mass_fixer=True)
    s = st.sht
    st.set_state(*initial_state(s, nlev))
    for _ in range(int(spinup_days * 86400.0 / dt)):
        st.step()
    x0 = [st.now[k].clone() for k in ("vor", "div", "temp", "lnps")]
    st.dteq = torch.zeros((nlev, s.nlat), dtype=torch.float64, device="cuda")
    adj = HeldSuarezAdjoint(st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    adj.run(x0, nsteps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    J, lam, gq = adj.gradient(x0, nsteps, every=every)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    peak = torch.cuda.max_memory_allocated()
    # dot-product residuals of the transpose transforms at this truncation
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    c1, c2 = dev(red_spectra(trunc, nlev, 3, field0=500)), dev(red_spectra(trunc, nlev, 3, field0=600))
    g1 = s.synthesis(dev(red_spectra(trunc, nlev, 3, field0=700)))
    g2 = s.synthesis(dev(red_spectra(trunc, nlev, 3, field0=800)))
    sd = lambda a, b: float((torch.view_as_real(a) * torch.view_as_real(b)).sum())
    gd = lambda a, b: float((a * b).sum())
    res = {}
    x = s.synthesis(c1)
    res["synthesis"] = abs(gd(x, g1) - sd(c1, s.synthesis_adjoint(g1))) / abs(gd(x, g1))
    x = s.analysis(g1)
    res["analysis"] = abs(sd(x, c1) - gd(g1, s.analysis_adjoint(c1))) / abs(sd(x, c1))
    vt, vp = s.synthesis_vec(c1, c2)
    a_, b_ = s.synthesis_vec_adjoint(g1, g2)
    lhs = gd(vt, g1) + gd(vp, g2)
    res["synthesis_vec"] = abs(lhs - sd(c1, a_) - sd(c2, b_)) / abs(lhs)
    sa, ta = s.analysis_vec(g1, g2)
    at, ap = s.analysis_vec_adjoint(c1, c2)
    lhs = sd(sa, c1) + sd(ta, c2)
    res["analysis_vec"] = abs(lhs - gd(g1, at) - gd(g2, ap)) / abs(lhs)
    gt, gp = s.gradient(c1)
    lhs = gd(gt, g1) + gd(gp, g2)
    res["gradient"] = abs(lhs - sd(c1, s.gradient_adjoint(g1, g2))) / abs(lhs)
    _, _, _, chk = directional_check(adj, x0, nsteps, seeded_directions(st, ndir), eps=eps, every=every)
    rels = sorted(r for _, _, r in chk)
    alg = nsteps * 3 * sref_bytes(s.nlon, s.nlat, s.nlm, nlev)
    return {"workload": f"T{trunc}L{nlev} Held-Suarez, {nsteps} steps of {dt:.0f} s from day {spinup_days:.0f}, "
                        f"FP64, checkpoints every {every}",
            "objective": "mass-weighted mean T over 30-60 N at the final step", "J_K": round(J, 9),
            "forward_ms_per_step": round(1000 * t_fwd / nsteps, 4),
            "gradient_s": round(t_grad, 4),
            "gradient_over_forward": round(t_grad / t_fwd, 2),
            "peak_device_memory_MB": round(peak / 1e6, 1),
            "alg_bytes_whole_run": int(alg), "alg_GBs_whole_run": round(alg / t_grad / 1e9, 1),
            "dot_product_rel_residual": {k: float(f"{v:.2e}") for k, v in res.items()},
            "fd_directions": ndir, "fd_eps": eps,
            "fd_rel_err_max": float(f"{rels[-1]:.3e}"), "fd_rel_err_median": float(f"{rels[len(rels) // 2]:.3e}")}


def dynamics_step(torch, precision, steps=20, warmup=3, path="staged"):
    """Whole Held-Suarez dynamics steps at T42L20 (transform halves, column terms +
    forcing, semi-implicit leapfrog with diffusion and filter, mass fixer), eager
    launches, all state on the device. Returns ms per step."""
    from aeros_amd.dynamics import Stepper
    from aeros_amd.column import vgrid_init
    from aeros_amd.seeding import red_spectra
    A, B = vgrid_init(NLEV, stretch_a=1.0, sigma_t=0.0, p_top=0.0)
    st = Stepper(T, A, B, 1200.0, tau_diff=0.1 * 86400.0, eps_filter=0.06, raw_alpha=0.53,
                 held_suarez=True, mass_fixer=True, precision=precision)
    st.sht.set_path("staged" if path == "auto" else path)
    nlm = st.sht.nlm
    vor = red_spectra(T, NLEV, 3, field0=0, zero_l0=True) * 1e-6
    temp = np.zeros((NLEV, nlm), complex)
    temp[:, 0] = 300.0 * np.sqrt(4 * np.pi)
    lnps = np.zeros(nlm, complex)
    lnps[0] = np.log(1.0e5) * np.sqrt(4 * np.pi)
    st.set_state(vor, np.zeros_like(vor), temp, lnps)
    for _ in range(warmup):
        st.step()
    torch.cuda.synchronize()
    ms = time_steps(torch, st.step, steps, 0)
    return round(ms, 4)


def dynamics_step_graph(torch, precision, steps=200, warmup=10):
    """The same T42L20 Held-Suarez step for one model through EnsembleStepper (one member):
    the step's launches replayed from a HIP graph, no Python between the kernels. Returns
    ms per step."""
    from aeros_amd.ensemble import EnsembleStepper, rest_states
    ens = EnsembleStepper([(T, NLEV, 1)], dt=1200.0, precision=precision)
    ens.autotune()
    ens.set_state(rest_states([(T, NLEV, 1)], seed_config=3))
    ens.step()
    ens.step()
    ens.capture()
    return round(time_steps(torch, ens.step, steps, warmup), 4)


def cpu_baseline(budget_s=12.0):
    """The FP64 CPU restatement timed in the reference's mode-B harness shape
    (drivers/bench_sht.f90:142-187): an OpenMP loop over the 132 units of one set, each
    unit one scalar synthesis + one analysis, one workspace per thread. The transforms are
    the oracle's algorithm in its CPU arrangement (oracle/cpu_port.c: hemispheric folding,
    parity-split tables, a half-length real FFT), checked against the oracle by the tests."""
    import numpy as np
    from aeros_amd.seeding import red_spectra
    from oracle.oracle import Oracle, Port
    p = Port(Oracle(T))
    units = (13 * NLEV + 4) // 2   # 264 scalar-equivalents = 132 (synthesis + analysis) units
    q = red_spectra(T, NLEV, config=1)
    q = np.concatenate([q] * (units // NLEV + 1))[:units]
    threads = min(16, os.cpu_count() or 1)
    t_unit = p.bench_set(q, 1, threads)
    nrep = max(1, min(2000, int(budget_s / max(t_unit * units, 1e-6))))
    t_unit = p.bench_set(q, nrep, threads)
    # one thread as well (the reference quotes per-core figures), on a quarter of the budget
    t1 = p.bench_set(q, 1, 1)
    nrep1 = max(1, min(500, int(0.25 * budget_s / max(t1 * units, 1e-6))))
    t1 = p.bench_set(q, nrep1, 1)
    return {"value": 1.0 / (t_unit * units), "unit": "sets/s", "cores": threads,
            "kind": "port",
            "value_1thread": 1.0 / (t1 * units),
            "sample": f"T42 FP64, the oracle's algorithm in its CPU arrangement (oracle/cpu_port.c: hemispheric "
                      f"folding, parity-split Legendre tables, half-length Stockham real FFT); OpenMP over the "
                      f"{units} (synthesis + analysis) units of one S_bl set, {nrep} sets; "
                      f"{t_unit * 1e6 / 2:.2f} us of wall time per scalar-equivalent transform on {threads} threads; "
                      f"1 thread: {nrep1} sets, {t1 * 1e6 / 2:.1f} us per scalar-equivalent transform"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hs", action="store_true",
                    help="skip the 1200-day Held-Suarez climate ensembles (3 x 32 members, ~90 s)")
    ap.add_argument("--no-stress", action="store_true", help="skip the 32 x T106L30 stress plan")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--only", choices=["set", "ensemble", "ensemble-step", "step", "stress", "stress-day", "hs",
                                       "adjoint"],
                    default=None,
                    help="profiling runs: only the T42L20 set (no ensemble, no FP64 forms), only the "
                         "ensemble, or only whole Held-Suarez dynamics steps")
    ap.add_argument("--path", choices=["auto", "tiled", "staged"], default="auto",
                    help="kernel path of the transforms: the autotuner's choice per direction (default), "
                         "or one path for both (aer_sht_set_path)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device-to-device copy 5 times (calibrates rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the transform set of one member is not split across "
              "GPUs (ensemble members would shard across GPUs with no exchange). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from aeros_amd.spectral import AerosSht
    torch.cuda.set_device(0)
    if args.copy_probe:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": src.numel() * 8, "launches": 5}))
        return
    if args.only == "step":
        print(json.dumps({"held_suarez_step_ms": dynamics_step(torch, args.precision, args.steps, args.warmup,
                                                               args.path),
                          "held_suarez_step_graph_ms": dynamics_step_graph(torch, args.precision, args.steps,
                                                                           args.warmup),
                          "precision": args.precision, "kernel_path": args.path}))
        return
    if args.only == "stress":
        print(json.dumps({"stress_T106L30x32": stress_T106(torch, steps=args.steps, warmup=args.warmup)}))
        return
    if args.only == "hs":
        print(json.dumps({"held_suarez_T31L20_1200d_x32": held_suarez_climate(torch)}))
        return
    if args.only == "stress-day":
        print(json.dumps({"stress_T106L30x32_day": stress_day_T106(torch)}))
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint_T31L20_96": adjoint_T31(torch)}))
        return
    if args.only == "ensemble":
        print(json.dumps({"ensemble_config2": ensemble_config2(torch, args.precision, args.steps, args.warmup,
                                                               args.path)}))
        return
    if args.only == "ensemble-step":
        print(json.dumps({"ensemble_config2_step": ensemble_config2_step(torch, args.precision, args.steps,
                                                                         args.warmup)}))
        return
    sht = AerosSht(T)
    sht.set_path("staged" if args.path == "auto" else args.path)
    ts = TransformSet(sht, torch)
    tune = (lambda: ts.autotune()) if args.path == "auto" else (lambda: None)
    if args.only == "set":
        from aeros_amd import _lib as L
        sht.set_precision(args.precision)
        if args.precision == "fp32":
            ts.set_assembly_prec(L.AER_JOB_PREC_FP64)
        tuned = tune()
        ms = time_steps(torch, ts.run_batched, args.steps, args.warmup)
        print(json.dumps({"set_ms_per_step": ms, "precision": args.precision, "kernel_path": args.path,
                          "autotune": tuned}))
        return
    sht._bind_stream()

    # FP64 reference outputs and the FP64 forms
    sht.set_precision("fp64")
    ts.run_batched()
    ref = ts.outputs()
    tuned64 = tune()
    plain_steps = max(5, args.steps // 10)
    # the per-level tendency call also re-analyses dlnps/dt on every level: 20 extra
    # single-level transforms beyond the set's 264; normalize the plain time to 264
    ms_plain = time_steps(torch, ts.run_plain, plain_steps, 2) * (13 * NLEV + 4) / (14 * NLEV + 3)
    ms_fp64_eager = time_steps(torch, ts.run_batched, args.steps, args.warmup)
    g64 = capture_graph(torch, sht, ts.run_batched)
    ms_fp64 = time_steps(torch, g64.replay, args.steps, args.warmup)
    k64_ms, _ = kernel_times(sht, ts.run_batched, max(10, args.steps // 4))

    from aeros_amd import _lib as L
    sht.set_precision(args.precision)
    errs_all = None
    if args.precision == "fp32":
        # every transform in FP32 ...
        ts.set_assembly_prec(L.AER_JOB_PREC_FP32)
        ts.run_batched()
        errs_all = {k: rel_err(v, ref[k]) for k, v in ts.outputs().items()}
        # ... and the measured form: the cancellation-prone (zeta_t, D_t - lap(Phi+K))
        # analysis job in FP64 inside the same fused launch
        ts.set_assembly_prec(L.AER_JOB_PREC_FP64)
    tuned = tune()
    ts.run_batched()
    errs = {k: rel_err(v, ref[k]) for k, v in ts.outputs().items()}
    ms_eager = time_steps(torch, ts.run_batched, args.steps, args.warmup)
    ms_calls = time_steps(torch, ts.run_calls, args.steps, args.warmup)
    graph = capture_graph(torch, sht, ts.run_batched)
    ms_graph = time_steps(torch, graph.replay, args.steps, args.warmup)
    # the same four launches either way; `value` is the faster way of issuing them
    ms, timing = ((ms_graph, "HIP graph replay of one set per step") if ms_graph <= ms_eager else
                  (ms_eager, "eager launches of one set per step (4 kernels; faster than its HIP-graph "
                             "replay on this runtime, which is reported beside it)"))
    ms_fp64_best = min(ms_fp64, ms_fp64_eager)
    torch.cuda.synchronize()
    graph_errs = {k: rel_err(v, ref[k]) for k, v in ts.outputs().items()}
    assert max(graph_errs.values()) <= max(errs.values()) * 1.0001 + 1e-300, "graph replay differs from eager"
    kms, klaunch = kernel_times(sht, ts.run_batched, max(10, args.steps // 4))

    # split-FP32 form (near-FP64): every job split, FP64 FFT
    sht.set_precision("split")
    ts.set_assembly_prec(L.AER_JOB_PREC_HANDLE)
    ts.run_batched()
    torch.cuda.synchronize()
    errs_split = {k: rel_err(v, ref[k]) for k, v in ts.outputs().items()}
    gsp = capture_graph(torch, sht, ts.run_batched)
    ms_split = time_steps(torch, gsp.replay, args.steps, args.warmup)
    sht.set_precision(args.precision)
    if args.precision == "fp32":
        ts.set_assembly_prec(L.AER_JOB_PREC_FP64)

    bytes_k, flops_k = algorithmic(sht.nlon, sht.nlat, sht.nlm, NLEV)
    set_bytes = sum(bytes_k)
    kernels = {}
    if args.path == "auto":
        sp = "staged" if tuned["synthesis"]["kept"] == "staged" else "tiled"
        ap_ = "staged" if tuned["analysis"]["kept"] == "staged" else "tiled"
        names = KIND_NAMES[sp][:2] + KIND_NAMES[ap_][2:4] + KIND_NAMES[sp][4:]
    else:
        names = KIND_NAMES[args.path]
    for k in range(NKIND):
        if not klaunch[k]:
            continue
        hbm = bytes_k[k] / (kms[k] * 1e-3) / 1e9
        tf = flops_k[k] / (kms[k] * 1e-3) / 1e12
        kernels[names[k]] = {"ms_per_set": round(kms[k], 5), "launches_per_set": klaunch[k],
                             "alg_bytes_per_set": int(bytes_k[k]), "alg_flops_per_set": int(flops_k[k]),
                             "hbm_GBs": round(hbm, 1), "TFLOPs": round(tf, 3),
                             "fp64_ms_per_set": round(k64_ms[k], 5)}
    dom = int(np.argmax(kms))
    pk = PEAK_TFLOPS[args.precision]
    if dom in (0, 3):   # Legendre GEMM: MFMA-bound by its dense flop count
        ach = flops_k[dom] / (kms[dom] * 1e-3) / 1e12
        roof = {"kernel": names[dom], "bound": "mfma", "achieved": round(ach, 3), "peak": pk,
                "unit": "TFLOP/s", "frac": round(ach / pk, 5), "traffic": None}
    else:   # FFT stages and the fused synthesis: HBM-bound by their algorithmic bytes
        ach = bytes_k[dom] / (kms[dom] * 1e-3) / 1e9
        roof = {"kernel": names[dom], "bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBS, 5), "traffic": None}
    roof["alg_per_launch"] = "see kernels[].alg_*_per_set / launches_per_set"
    roof["traffic"], roof["traffic_source"] = measured_traffic(names[dom])

    line = {
        "metric": "SH transform sets/s at T42L20, mixed-prec vs FP64 on same GPU at stated rel err",
        "value": round(1000.0 / ms, 2),
        "unit": "sets/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded red-spectrum states, SplitMix64 + Box-Muller)",
        "config": {"workload": "T42L20 S_bl transform set (13L+4 = 264 scalar-equivalent transforms: "
                               "lnps, grad lnps; per level (u,v), zeta, D, T, q, grad T synthesis; "
                               "(zeta_t, D_t - lap(Phi+K)), T_t, q_t, lnps_t analysis)",
                   "trunc": T, "nlev": NLEV, "grid": f"{sht.nlon}x{sht.nlat}",
                   "precision": ("fp32-mfma; the (zeta_t, D_t) assembly job fp64 in the same launch"
                                 if args.precision == "fp32" else "fp64-mfma"),
                   "io": "fp64 in HBM"},
        "timing": timing,
        "kernel_path": args.path,
        "autotune": tuned,
        "autotune_fp64": tuned64,
        "eager_sets_per_s": round(1000.0 / ms_eager, 2),
        "graph_sets_per_s": round(1000.0 / ms_graph, 2),
        "eager_4_calls_sets_per_s": round(1000.0 / ms_calls, 2),
        "fp64_batched_sets_per_s": round(1000.0 / ms_fp64, 2),
        "fp64_batched_eager_sets_per_s": round(1000.0 / ms_fp64_eager, 2),
        "fp64_plain_sets_per_s": round(1000.0 / ms_plain, 2),
        "speedup_vs_fp64_plain": round(ms_plain / ms, 2),
        "speedup_vs_fp64_batched": round(ms_fp64_best / ms, 3),
        "max_rel_err_vs_fp64": max(errs.values()),
        "rel_err_vs_fp64_per_field": {k: float(f"{v:.3e}") for k, v in errs.items()},
        "rel_err_all_fp32_per_field": (None if errs_all is None else
                                       {k: float(f"{v:.3e}") for k, v in errs_all.items()}),
        "split_fp32_sets_per_s": round(1000.0 / ms_split, 2),
        "max_rel_err_split_vs_fp64": max(errs_split.values()),
        "rel_err_split_per_field": {k: float(f"{v:.3e}") for k, v in errs_split.items()},
        "set_hbm_GBs": round(set_bytes / (ms * 1e-3) / 1e9, 1),
        "roofline": roof,
        "kernels": kernels,
    }
    line["ensemble_config2"] = ensemble_config2(torch, args.precision, path=args.path)
    line["ensemble_config2_step"] = ensemble_config2_step(torch, args.precision)
    line["held_suarez_step_T42L20"] = {
        "what": "one whole dynamics step: transform halves, column terms + HS forcing, semi-implicit "
                "leapfrog + del^6 diffusion + RAW filter (one kernel), FP64 mass fixer; eager launches",
        "ms_per_step_fp64": dynamics_step(torch, "fp64", path=args.path),
        "ms_per_step_fp32_transforms": dynamics_step(torch, "fp32", path=args.path),
        "graph_ms_per_step_fp64": dynamics_step_graph(torch, "fp64"),
        "graph_ms_per_step_fp32_transforms": dynamics_step_graph(torch, "fp32")}
    line["adjoint_T31L20_96"] = adjoint_T31(torch)
    if not args.no_hs:
        line["held_suarez_T31L20_1200d_x32"] = held_suarez_climate(torch)
    if not args.no_stress:
        line["stress_T106L30x32"] = stress_T106(torch)
        line["stress_T106L30x32_day"] = stress_day_T106(torch)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
