"""A whole Held-Suarez ensemble step on one GPU: members of uneven truncation and level
count, grouped by size, every stage of a step batched across members, the transforms
of all groups fused into compiled plans, and the step replayed from a HIP graph.

The reference integrates one model per process (src/dynamics/aeros_timestep.f90:776-1137);
this module is the ensemble layer its configuration of 64 uneven members needs.

Per step (members of a group share one launch per stage):

  synthesis plan    every group's tendency_grid_transforms jobs      (2 launches)
  column terms      + Held-Suarez forcing, per group                 (1 launch per group)
  analysis plan     every group's tendency_spectral jobs             (2 launches)
  leapfrog          semi-implicit FP64 solve, diffusion, RAW filter  (1 launch per group)
  mass fixer        FP64 synthesis + sum + fix, per group            (2 launches per group)

State lives in three fixed time-level buffers per group that rotate (old, now, new), so
the launches of step n depend only on n mod 3: three HIP graphs, captured once, replay
every later step with no host work beyond the replay.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from .column import column_terms, vgrid_init
from .dynamics import Stepper
from .step import mass

KEYS = ("vor", "div", "temp", "lnps")
GKEYS = ("u", "v", "vor_g", "div_g", "t_g", "dtdx", "dtdy", "lnps_g", "dlnpsdx", "dlnpsdy")
CKEYS = ("ae", "an", "ek", "dtdt", "dlnpsdt")
RADIUS = 6.371e6


def _job(kind, n, ins, outs, modes=(0, 0, 0), scales=(0.0, 0.0), prec=0):
    p = [t.data_ptr() for t in ins] + [0] * (3 - len(ins))
    q = [t.data_ptr() for t in outs] + [0] * (2 - len(outs))
    return L.AerJob(kind, n, *p, *q, 0, 0, *modes, *scales, 0, prec)


class _Plan:
    """A compiled job list over several handles (aer_plan_*)."""

    def __init__(self, cfgs, jobs):
        n = len(jobs)
        h = ctypes.c_void_p()
        L.check(L.lib().aer_plan_create(n, (ctypes.c_void_p * n)(*cfgs), (L.AerJob * n)(*jobs), ctypes.byref(h)))
        self._h = h

    def run(self):
        L.check(L.lib().aer_plan_run(self._h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def autotune(self, reps=10):
        t = (ctypes.c_double * (2 * L.AER_TUNE_NCAND))()
        c = (ctypes.c_int * 2)()
        L.check(L.lib().aer_plan_autotune(self._h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
                                          reps, t, c))
        return [list(t[:L.AER_TUNE_NCAND]), list(t[L.AER_TUNE_NCAND:])], list(c)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().aer_plan_destroy(h)
            except Exception:
                pass
            self._h = None


class _Group:
    """Members of one (truncation, levels): a Stepper for its operators and settings, and
    the fixed buffers of the ensemble step."""

    def __init__(self, trunc, nlev, nmem, dt, precision, **kw):
        A, B = vgrid_init(nlev, stretch_a=1.0, stretch_r=2.0, sigma_t=0.0, p_top=1000.0)
        self.st = Stepper(trunc, A, B, dt, precision=precision, **kw)
        s = self.st.sht
        self.nmem, self.nlev = nmem, nlev
        spec = lambda *lead: torch.zeros(lead + (s.nlm,), dtype=torch.complex128, device="cuda")
        grid = lambda *lead: torch.zeros(lead + (s.nlat, s.nlon), dtype=torch.float64, device="cuda")
        self.X = [{"vor": spec(nmem, nlev), "div": spec(nmem, nlev), "temp": spec(nmem, nlev), "lnps": spec(nmem)}
                  for _ in range(3)]
        self.G = {k: grid(nmem, nlev) if i < 7 else grid(nmem) for i, k in enumerate(GKEYS)}
        self.C = {k: grid(nmem, nlev) if k != "dlnpsdt" else grid(nmem) for k in CKEYS}
        self.R = {"vor": spec(nmem, nlev), "div": spec(nmem, nlev), "temp": spec(nmem, nlev), "lnps": spec(nmem)}
        z = lambda: torch.zeros(nmem, dtype=torch.float64, device="cuda")
        self.target, self.mass, self.lnr, self.lnr_cum = z(), z(), z(), z()

    def synth_jobs(self, now):
        a, G, nf, nm = 1.0 / RADIUS, self.G, self.nmem * self.nlev, self.nmem
        return [_job(L.AER_JOB_SYNTH_SCALAR, nm, [now["lnps"]], [G["lnps_g"]]),
                _job(L.AER_JOB_SYNTH_GRAD, nm, [now["lnps"]], [G["dlnpsdy"], G["dlnpsdx"]], scales=(-a, a)),
                _job(L.AER_JOB_SYNTH_VECTOR, nf, [now["div"], now["vor"]], [G["v"], G["u"]],
                     modes=(L.AER_MODE_S_FROM_DIV, L.AER_MODE_T_FROM_VOR, 0), scales=(-1.0, 1.0)),
                _job(L.AER_JOB_SYNTH_SCALAR, nf, [now["vor"]], [G["vor_g"]]),
                _job(L.AER_JOB_SYNTH_SCALAR, nf, [now["div"]], [G["div_g"]]),
                _job(L.AER_JOB_SYNTH_SCALAR, nf, [now["temp"]], [G["t_g"]]),
                _job(L.AER_JOB_SYNTH_GRAD, nf, [now["temp"]], [G["dtdy"], G["dtdx"]], scales=(-a, a))]

    def anal_jobs(self, assembly_prec):
        C, R, nf, nm = self.C, self.R, self.nmem * self.nlev, self.nmem
        return [_job(L.AER_JOB_ANAL_VECTOR, nf, [C["an"], C["ae"], C["ek"]], [R["div"], R["vor"]],
                     modes=(L.AER_MODE_DIV_FROM_S, L.AER_MODE_VOR_FROM_T, L.AER_MODE_NEG_LAP), scales=(-1.0, 1.0),
                     prec=assembly_prec),
                _job(L.AER_JOB_ANAL_SCALAR, nf, [C["dtdt"]], [R["temp"]]),
                _job(L.AER_JOB_ANAL_SCALAR, nm, [C["dlnpsdt"]], [R["lnps"]])]


class EnsembleStepper:
    """Held-Suarez members of uneven size stepped together.

    groups: list of (trunc, nlev, members). precision: the transforms' compute type
    ("fp32", "split" or "fp64"); with "fp32" the cancellation-prone (zeta_t, D_t - lap(Phi+K))
    analysis runs FP64 inside the same fused launch. The semi-implicit solve, diffusion,
    filter and mass fixer are FP64 in every form. Settings are the Held-Suarez protocol's
    (del^6 diffusion with a 6 h e-folding at lmax, RAW filter 0.06 / 0.53, mass fixer)."""

    def __init__(self, groups, dt=1200.0, precision="fp32", tau_diff=6 * 3600.0, eps_filter=0.06, raw_alpha=0.53):
        self.dt, self.eps, self.raw_alpha = dt, eps_filter, raw_alpha
        kw = dict(tau_diff=tau_diff, ndiff=6, eps_filter=eps_filter, raw_alpha=raw_alpha, held_suarez=True,
                  mass_fixer=True)
        self.groups = [_Group(T, nlev, n, dt, precision, **kw) for T, nlev, n in groups]
        self.assembly_prec = L.AER_JOB_PREC_FP64 if precision == "fp32" else 0
        self.synth = [self._synth_plan(p) for p in range(3)]
        self.anal = self._plan([j for g in self.groups for j in g.anal_jobs(self.assembly_prec)],
                               [g.st.sht._h for g in self.groups for _ in range(3)])
        self.nstep = 0
        self.graphs = None
        self.tuned = None

    @staticmethod
    def _plan(jobs, cfgs):
        return _Plan(cfgs, jobs)

    def _synth_plan(self, p):
        """The synthesis plan of rotation phase p (reads time level p). A plan takes each
        handle's kernel configuration when it is built."""
        jobs = [j for g in self.groups for j in g.synth_jobs(g.X[p])]
        return self._plan(jobs, [g.st.sht._h for g in self.groups for _ in range(7)])

    @property
    def members(self):
        return sum(g.nmem for g in self.groups)

    def set_state(self, states):
        """states: one (vor, div, temp, lnps) tuple per group, arrays with a leading member
        dimension ((nmem, nlev, nlm) and (nmem, nlm))."""
        if len(states) != len(self.groups):
            raise ValueError(f"expected {len(self.groups)} group states, got {len(states)}")
        for g, st in zip(self.groups, states):
            if len(st) != 4:
                raise ValueError("a group state is (vor, div, temp, lnps)")
            for k, a in zip(KEYS, st):
                t = torch.as_tensor(a, dtype=torch.complex128)
                if tuple(t.shape) != tuple(g.X[0][k].shape):
                    raise ValueError(f"{k}: expected shape {tuple(g.X[0][k].shape)}, got {tuple(t.shape)}")
                g.X[0][k].copy_(t)
        self.nstep = 0   # captured graphs stay valid: the buffers do not move

    def now(self, gi):
        """Current state of group gi (views of the rotating buffers)."""
        return self.groups[gi].X[self.nstep % 3]

    def _launches(self, n):
        """Every launch of step n >= 0 (no host synchronization, no allocation after the
        first call at each shape)."""
        p = n % 3
        self.synth[p].run()
        for g in self.groups:
            column_terms(g.st.sht, g.st.vg, g.G, held_suarez=True, out=g.C)
        self.anal.run()
        for g in self.groups:
            old, now, new = g.X[(n + 2) % 3], g.X[p], g.X[(n + 1) % 3]
            eps = self.eps if n > 0 else 0.0
            g.st.sht._bind_stream()
            L.check(L.lib().aer_leapfrog_batch_d(
                g.st.si._h, g.nmem, g.st.si.flags(True, True, eps), eps, self.raw_alpha,
                *[old[k].data_ptr() for k in KEYS], *[now[k].data_ptr() for k in KEYS],
                *[g.R[k].data_ptr() for k in KEYS], *[new[k].data_ptr() for k in KEYS]))
            g.st.sht._bind_stream()
            L.check(L.lib().aer_mass_batch_d(g.st.sht._h, g.nmem, new["lnps"].data_ptr(), 1, g.target.data_ptr(),
                                             g.mass.data_ptr(), g.lnr.data_ptr()))
            g.lnr_cum.add_(g.lnr)

    def step(self):
        """One step: eager for the start-up steps, then a graph replay."""
        n = self.nstep
        if n == 0:
            for g in self.groups:
                for k in KEYS:
                    g.X[2][k].copy_(g.X[0][k])   # X^(-1) := X^0
                m, _ = mass(g.st.sht, g.X[0]["lnps"])
                g.target.copy_(m)
                g.lnr_cum.zero_()
                g.st.si.set_step(self.dt)
        elif n == 1:
            for g in self.groups:
                g.st.si.set_step(2.0 * self.dt)
        if n >= 1 and self.graphs is not None:
            self.graphs[n % 3].replay()
        else:
            self._launches(n)
        self.nstep = n + 1

    def capture(self):
        """Capture the three rotation phases as HIP graphs. The leapfrog's step length is a
        launch argument, so capture after the second step has set it to 2 dt."""
        if self.nstep < 2:
            raise RuntimeError("capture after the first two steps")
        torch.cuda.synchronize()
        graphs = []
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for p in range(3):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=s):
                    self._launches(p + 3)   # any n >= 1 with n mod 3 == p
                graphs.append(g)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graphs = graphs

    def autotune(self, reps=10):
        """Autotune the synthesis and analysis plans' kernel configurations (timed on scratch
        work: the state buffers of the plans are overwritten, so call before set_state)."""
        out = {"synthesis": self.synth[0].autotune(reps), "analysis": self.anal.autotune(reps)}
        # the kept configuration now sits on the handles: rebuild the other phases' plans
        self.synth[1:] = [self._synth_plan(p) for p in (1, 2)]
        self.graphs = None
        self.tuned = out
        return out


def rest_states(groups, seed_config=2, noise=1.0e-6, lmax_noise=20, t0=300.0, ps=1.0e5):
    """The ensemble's initial states: rest, isothermal t0, p_s = ps, plus seeded vorticity
    noise of grid rms `noise` (s^-1) on degrees 1..lmax_noise; member m of the whole
    ensemble uses seeds 1000 config + level + 100000 m."""
    from .seeding import lm_tables, red_spectra
    out, m0 = [], 0
    for T, nlev, n in groups:
        l, _ = lm_tables(T)
        nlm = l.size
        vor = np.stack([red_spectra(T, nlev, seed_config, member=m0 + i, zero_l0=True) for i in range(n)])
        vor[..., l > lmax_noise] = 0.0
        # grid rms of an orthonormal expansion: sqrt(sum_m w_m |c|^2 / 4 pi), w_0 = 1, w_m>0 = 2
        wm = np.where(np.arange(nlm) <= T, 1.0, 2.0)
        rms = np.sqrt((wm * np.abs(vor) ** 2).sum(-1).mean() / (4.0 * math.pi))
        vor *= noise / rms
        temp = np.zeros((n, nlev, nlm), complex)
        temp[..., 0] = t0 * math.sqrt(4.0 * math.pi)
        lnps = np.zeros((n, nlm), complex)
        lnps[:, 0] = math.log(ps) * math.sqrt(4.0 * math.pi)
        out.append((vor, np.zeros_like(vor), temp, lnps))
        m0 += n
    return out
