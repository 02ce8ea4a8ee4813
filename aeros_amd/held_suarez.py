"""The Held-Suarez benchmark on the GPU: the reference's drivers/held_suarez.f90 protocol
(par/held_suarez.nml) on this package's dynamics step.

  * Twenty evenly spaced sigma levels with a 1000 Pa top, ps_ref 1e5 Pa, semi-implicit
    t_ref 300 K (par/held_suarez.nml &aeros_vert).
  * Leapfrog with RAW filter (eps 0.06, alpha 0.53), del^6 diffusion with a 6 h e-folding
    at l = lmax, mass fixer on (par/held_suarez.nml &aeros).
  * Motionless isothermal start at t_init (aeros_set_isothermal, src/aeros.f90:130-183) plus
    the deterministic temperature seed of aeros_hs_perturb
    (src/physics/aeros_held_suarez.f90:295-336).
  * Daily samples of the zonal-mean statistics from day_avg_start on
    (drivers/held_suarez.f90:149-183), summarized as aeros_diag_summary does.

    python -m aeros_amd.held_suarez --trunc 31 --days 1200 --avg-start 200 --precision fp64
"""
import argparse
import json
import math
import time

import numpy as np

from .column import vgrid_init
from .diagnostics import ZonalStats, summary
from .dynamics import Stepper

DAY = 86400.0


def hs_perturb(temp, sht, amp):
    """aeros_hs_perturb: T_lm += amp (cos(3l + 5m + k), sin(7l + 2m + 3k)) for m = 1..8,
    l = m..m+4 (k the 1-based level index), on a (nlev, nlm) complex array in place."""
    nlev = temp.shape[0]
    for k in range(1, nlev + 1):
        for m in range(1, min(8, sht.mmax) + 1):
            for l in range(m, min(m + 4, sht.lmax) + 1):
                lm = sht.lm(l, m) - 1
                temp[k - 1, lm] += amp * complex(math.cos(3 * l + 5 * m + k), math.sin(7 * l + 2 * m + 3 * k))
    return temp


def initial_state(sht, nlev, t_init=264.0, ps=1.0e5, perturb=1.0):
    """Motionless isothermal state (spectral (0,0) coefficients X sqrt(4 pi)) with the seed."""
    nlm = sht.nlm
    y00 = math.sqrt(4.0 * math.pi)
    vor = np.zeros((nlev, nlm), complex)
    div = np.zeros((nlev, nlm), complex)
    temp = np.zeros((nlev, nlm), complex)
    temp[:, sht.lm(0, 0) - 1] = t_init * y00
    lnps = np.zeros(nlm, complex)
    lnps[sht.lm(0, 0) - 1] = math.log(ps) * y00
    hs_perturb(temp, sht, perturb)
    return vor, div, temp, lnps


def run(trunc=31, nlev=20, dt=1800.0, days=1200.0, avg_start=200.0, sample_days=1.0, precision="fp64",
        t_init=264.0, perturb=1.0, tau_diff_h=6.0, progress=None):
    """Integrate and return (summary dict, finalized statistics dict, run info)."""
    import torch
    A, B = vgrid_init(nlev, stretch_a=1.0, stretch_r=2.0, sigma_t=0.0, p_top=1000.0, ps_ref=1.0e5)
    st = Stepper(trunc, A, B, dt, t_ref=300.0, ps_ref=1.0e5, tau_diff=tau_diff_h * 3600.0, ndiff=6,
                 eps_filter=0.06, raw_alpha=0.53, held_suarez=True, mass_fixer=True, precision=precision)
    s = st.sht
    st.set_state(*initial_state(s, nlev, t_init=t_init, perturb=perturb))
    stats = ZonalStats(s, nlev)
    steps_per_sample = int(round(sample_days * DAY / dt))
    nsample = int(round(days / sample_days))
    t0 = time.time()
    for n in range(1, nsample + 1):
        for _ in range(steps_per_sample):
            st.step()
        day = n * sample_days
        if day >= avg_start:
            now = st.now
            g = s.tendency_grid_transforms(now["vor"], now["div"], now["temp"], now["lnps"])
            stats.accumulate(g, day)
        if progress and n % max(1, nsample // progress) == 0:
            torch.cuda.synchronize()
            g = s.tendency_grid_transforms(st.now["vor"], st.now["div"], st.now["temp"], st.now["lnps"])
            umax = float(torch.sqrt(g["u"] ** 2 + g["v"] ** 2).max())
            print(f"day {day:8.1f}  max|v| {umax:7.2f}  T [{float(g['t_g'].min()):7.2f}, "
                  f"{float(g['t_g'].max()):7.2f}]  {time.time() - t0:7.1f} s", flush=True)
            if not math.isfinite(umax) or umax > 300.0:
                raise RuntimeError(f"the run went unstable at day {day}")
    torch.cuda.synchronize()
    wall = time.time() - t0
    d = stats.finalize(A, B)
    lat = np.degrees(np.arcsin(s.sinlat))
    info = {"trunc": trunc, "nlev": nlev, "dt": dt, "days": days, "avg_start": avg_start,
            "precision": precision, "steps": st.nstep, "wall_s": round(wall, 2),
            "ms_per_step": round(1000.0 * wall / max(st.nstep, 1), 4)}
    return summary(d, lat), d, info


def run_ensemble(trunc=31, nlev=20, members=64, dt=1800.0, days=1200.0, avg_start=200.0, sample_days=1.0,
                 precision="fp64", seed_offset=0, progress=None):
    """The same protocol for an ensemble of `members` runs stepped together
    (aeros_amd.ensemble.EnsembleStepper: one batched launch per stage, HIP-graph replay).
    Member m starts from the isothermal state with the seed amplitude 1 + 1e-3 (m +
    seed_offset), so members decorrelate within the spin-up. Every member's daily samples go
    into one set of sums: the result is the ensemble-mean time-mean climate, whose sampling
    noise is that of a single run over sqrt(members)."""
    import torch

    from .ensemble import EnsembleStepper
    ens = EnsembleStepper([(trunc, nlev, members)], dt=dt, precision=precision)
    s = ens.groups[0].st.sht
    ens.autotune()
    states = [initial_state(s, nlev, perturb=1.0 + 1.0e-3 * (m + seed_offset)) for m in range(members)]
    ens.set_state([tuple(np.stack([st[i] for st in states]) for i in range(4))])
    A, B = vgrid_init(nlev, stretch_a=1.0, stretch_r=2.0, sigma_t=0.0, p_top=1000.0, ps_ref=1.0e5)
    stats = ZonalStats(s, nlev)
    steps_per_sample = int(round(sample_days * DAY / dt))
    nsample = int(round(days / sample_days))
    t0 = time.time()
    for n in range(1, nsample + 1):
        for _ in range(steps_per_sample):
            ens.step()
            if ens.nstep == 2:
                ens.capture()
        day = n * sample_days
        if day >= avg_start or (progress and n % max(1, nsample // progress) == 0):
            now = ens.now(0)
            g = s.tendency_grid_transforms(now["vor"], now["div"], now["temp"], now["lnps"])
            if day >= avg_start:
                for m in range(members):
                    stats.accumulate({k: g[k][m] for k in ("u", "v", "t_g", "lnps_g")}, day)
            if progress and n % max(1, nsample // progress) == 0:
                umax = float(torch.sqrt(g["u"] ** 2 + g["v"] ** 2).max())
                print(f"day {day:8.1f}  max|v| {umax:7.2f}  {time.time() - t0:7.1f} s", flush=True)
                if not math.isfinite(umax) or umax > 300.0:
                    raise RuntimeError(f"the ensemble went unstable at day {day}")
    torch.cuda.synchronize()
    wall = time.time() - t0
    d = stats.finalize(A, B)
    lat = np.degrees(np.arcsin(s.sinlat))
    info = {"trunc": trunc, "nlev": nlev, "members": members, "dt": dt, "days": days, "avg_start": avg_start,
            "precision": precision, "seed_offset": seed_offset, "steps": ens.nstep, "wall_s": round(wall, 2),
            "ms_per_ensemble_step": round(1000.0 * wall / max(ens.nstep, 1), 4)}
    return summary(d, lat), d, info


def compare(da, db):
    """Differences of the time-mean zonal-mean [T] and [u] between two runs: largest and rms,
    as they are and after averaging the two hemispheres (the forcing is symmetric, so the
    hemispheric mean halves the sampling noise of a finite window)."""
    sym = lambda x: 0.5 * (x + x[::-1, :])
    out = {}
    for tag, f in (("", lambda x: x), ("sym_", sym)):
        dT = f(da["temp"]) - f(db["temp"])
        du = f(da["u"]) - f(db["u"])
        out.update({tag + "max_abs_dT_K": float(np.abs(dT).max()), tag + "max_abs_du_ms": float(np.abs(du).max()),
                    tag + "rms_dT_K": float(np.sqrt((dT ** 2).mean())), tag + "rms_du_ms": float(np.sqrt((du ** 2).mean()))})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trunc", type=int, default=31)
    ap.add_argument("--nlev", type=int, default=20)
    ap.add_argument("--dt", type=float, default=1800.0)
    ap.add_argument("--days", type=float, default=1200.0)
    ap.add_argument("--avg-start", type=float, default=200.0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "split", "both"],
                    help="both: an FP64 run, an FP32-transform run and an FP64 twin whose seed differs "
                         "by 1e-12 relative (the chaotic twin measures the sampling noise of the window)")
    ap.add_argument("--out", default=None, help="write the summary (and the FP64-vs-FP32 comparison) as JSON")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="members per run (0: single runs); with --precision both: an FP64 ensemble, the same "
                         "ensemble with FP32 transforms (the FP64 assembly job), and an independent FP64 ensemble")
    args = ap.parse_args()
    import torch
    torch.cuda.set_device(0)
    if args.ensemble > 0:
        runs = ([("fp64", "fp64", 0), ("fp32", "fp32", 0), ("fp64_twin", "fp64", args.ensemble)]
                if args.precision == "both" else [(args.precision, args.precision, 0)])
        res = {}
        for name, p, off in runs:
            print(f"== Held-Suarez T{args.trunc}L{args.nlev} x {args.ensemble} members, {name}", flush=True)
            sm, d, info = run_ensemble(args.trunc, args.nlev, args.ensemble, args.dt, args.days, args.avg_start,
                                       precision=p, seed_offset=off, progress=6)
            print(json.dumps({"summary": sm, "info": info}), flush=True)
            res[name] = (sm, d, info)
        out = {p: {"summary": res[p][0], "info": res[p][2]} for p in res}
        if len(runs) == 3:
            out["fp32_vs_fp64"] = compare(res["fp32"][1], res["fp64"][1])
            out["fp64_twin_vs_fp64"] = compare(res["fp64_twin"][1], res["fp64"][1])
            print(json.dumps({"fp32_vs_fp64": out["fp32_vs_fp64"], "fp64_twin_vs_fp64": out["fp64_twin_vs_fp64"]}),
                  flush=True)
        if args.out:
            with open(args.out, "w") as f:
                json.dump(out, f, indent=1)
        return
    runs = ([("fp64", "fp64", 1.0), ("fp32", "fp32", 1.0), ("fp64_twin", "fp64", 1.0 + 1e-12)]
            if args.precision == "both" else [(args.precision, args.precision, 1.0)])
    res = {}
    for name, p, amp in runs:
        print(f"== Held-Suarez T{args.trunc}L{args.nlev}, {name} ({p} transforms, seed x {amp!r})", flush=True)
        sm, d, info = run(args.trunc, args.nlev, args.dt, args.days, args.avg_start, precision=p, perturb=amp,
                          progress=12)
        print(json.dumps({"summary": sm, "info": info}), flush=True)
        res[name] = (sm, d, info)
    out = {p: {"summary": res[p][0], "info": res[p][2]} for p in res}
    if len(runs) == 3:
        out["fp32_vs_fp64"] = compare(res["fp32"][1], res["fp64"][1])
        out["fp64_twin_vs_fp64"] = compare(res["fp64_twin"][1], res["fp64"][1])
        print(json.dumps({"fp32_vs_fp64": out["fp32_vs_fp64"], "fp64_twin_vs_fp64": out["fp64_twin_vs_fp64"]}),
              flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
