"""Gradients of a Held-Suarez functional through the semi-implicit leapfrog, on the GPU.

The reference has no adjoint; this module builds one from hand-written transposes:

  * the transform halves: the transpose transforms (aer_synth_adj_d, aer_synth_vec_adj_d,
    aer_synth_grad_adj_d, aer_anal_adj_d, aer_anal_vec_adj_d) with the per-degree
    factors of the forward jobs applied on the spectral side;
  * the column terms and forcing: aer_column_terms_adj_d (a reverse sweep per column);
  * the leapfrog, semi-implicit solve, diffusion and RAW filter: aer_leapfrog_adj_d;
  * the mass fixer: its linearization, ln p_s -> ln p_s - c00 e00 (grad M / M).

Inner products are Euclidean: grid fields over their values, spectral fields over the
real and imaginary parts of every coefficient. The forward state is recomputed from
step checkpoints (Stepper.checkpoint) one segment at a time, as the reference's restart
state would allow: memory is one segment of states, not the whole run.

The objective is the mass-weighted mean of the temperature over 30-60 N at the final
step: weights w_j dp_k with dp_k the layer thickness at p_s = 1e5 Pa. The controls are
the initial spectral state and an additive T_eq field dteq [nlev, nlat] of the
Held-Suarez forcing. The mass fixer's target is the mass of the initial ln p_s, and its
sensitivity is carried back to that initial field.
"""
import math

import torch

from .column import column_terms, column_terms_adjoint
from .step import mass

KEYS = ("vor", "div", "temp", "lnps")
RADIUS = 6.371e6


def dot(a, b):
    """Euclidean inner product of two dicts (or tensors) of CUDA tensors; complex tensors
    count their real and imaginary parts."""
    if isinstance(a, dict):
        return sum(dot(a[k], b[k]) for k in a)
    ra = torch.view_as_real(a) if a.is_complex() else a
    rb = torch.view_as_real(b) if b.is_complex() else b
    return float((ra * rb).sum())


def _lfac(sht, fn):
    l = torch.as_tensor(sht.l_of_lm, dtype=torch.float64, device="cuda")
    return fn(l)


class HeldSuarezAdjoint:
    """Forward run with checkpoints, and its reverse sweep, for one Stepper configuration."""

    def __init__(self, stepper):
        if stepper.sht.precision not in ("fp64", "split"):
            raise ValueError("the adjoint transposes the FP64 or split-FP32 transforms (near-exact) only")
        self.st = stepper
        self.target_bar = 0.0
        s = stepper.sht
        ll1 = _lfac(s, lambda l: l * (l + 1.0))
        inv = torch.where(ll1 > 0, 1.0 / torch.where(ll1 > 0, ll1, 1.0), torch.zeros_like(ll1))
        # per-degree factors of the forward jobs (src/dynamics/aeros_vordiv.f90:100-164,
        # src/dynamics/aeros_tendency.f90:458-460)
        self.f_s_from_div = -RADIUS * inv
        self.f_t_from_vor = RADIUS * inv
        self.f_sraw_of_tdiv = -ll1 / RADIUS        # tdiv = this * S (S = anal_vec's 1/(l(l+1)) form)
        self.f_traw_of_tvor = ll1 / RADIUS
        self.f_neg_lap = ll1 / (RADIUS * RADIUS)
        prec = s.precision
        s.set_precision("fp64")   # c00 as the mass fixer reads it (an FP64 analysis of 1)
        self.c00 = float(s.analysis(torch.ones((s.nlat, s.nlon), dtype=torch.float64, device="cuda"))[0].real)
        s.set_precision(prec)
        w = torch.as_tensor(s.gauss_w, dtype=torch.float64, device="cuda")
        self.wrow = (w * (2.0 * math.pi / s.nlon))[:, None]

    # ------------------------------------------------------------ objective
    def objective_weights(self):
        s, vg = self.st.sht, self.st.vg
        import numpy as np
        lat = np.degrees(np.arcsin(s.sinlat))
        band = torch.as_tensor((lat >= 30.0) & (lat <= 60.0), dtype=torch.float64, device="cuda")
        dp = torch.as_tensor(np.diff(vg.a_half) + np.diff(vg.b_half) * 1.0e5, dtype=torch.float64, device="cuda")
        w = torch.as_tensor(s.gauss_w, dtype=torch.float64, device="cuda")
        wk = dp[:, None] * (w * band)[None, :]                  # [nlev, nlat]
        wk = wk / (wk.sum() * s.nlon)
        w = wk[:, :, None].expand(self.st.nlev, s.nlat, s.nlon)
        shape = tuple(self.st.now["temp"].shape[:-1]) + (s.nlat, s.nlon)
        return w.expand(shape).contiguous()

    def objective(self, now):
        """J (for an ensemble, the sum of the members' values)."""
        t = self.st.sht.synthesis(now["temp"])
        return float((t * self.objective_weights()).sum())

    def objective_adjoint(self):
        """Adjoint of the final state: d J / d X^n (temperature only)."""
        s = self.st.sht
        lam = {k: torch.zeros_like(v) for k, v in self.st.now.items()}
        lam["temp"] = s.synthesis_adjoint(self.objective_weights())
        return lam

    # ------------------------------------------------------------ forward pieces
    def step_record(self):
        """One Stepper step (the same calls in the same order) that also returns what its
        adjoint needs: X^n as it entered, the step settings and ln p_s before the fixer."""
        st = self.st
        s, now = st.sht, st.now
        rec = {"nstep": st.nstep, "now": {k: v.clone() for k, v in now.items()}}
        if st.nstep == 0:
            h = st.dt
            st.old = {k: v.clone() for k, v in now.items()}
            if st.mass_fixer and st.mass_target is None:
                st.mass_target = st._mass_target(now["lnps"])
        else:
            h = 2.0 * st.dt
        if h != st.h:
            st.si.set_step(h)
            st.h = h
        g = s.tendency_grid_transforms(now["vor"], now["div"], now["temp"], now["lnps"])
        c = column_terms(s, st.vg, g, phis=st.phis, held_suarez=st.hs, dteq=st.dteq)
        tnd = s.tendency_spectral(c["ae"], c["an"], c["ek"], c["dtdt"], c["dlnpsdt"])
        eps = st.eps if st.nstep > 0 else 0.0
        new = st.si.leapfrog(st.old, now, tnd, semi_implicit=st.semi, diffuse=st.diffuse,
                             eps_filter=eps, raw_alpha=st.raw_alpha)
        st.old, st.now = now, new
        rec.update(h=h, eps=eps, lnps_prefix=new["lnps"].clone())
        if st.mass_fixer:
            _, lnr = mass(s, st.now["lnps"], fix=True, target=st.mass_target)
            st.lnr_cum = lnr if st.lnr_cum is None else st.lnr_cum + lnr
        st.nstep += 1
        st.last_grid = g
        return rec

    # ------------------------------------------------------------ adjoint pieces
    def tg_adjoint(self, gbar):
        """Transpose of AerosSht.tendency_grid_transforms."""
        s = self.st.sht
        a = RADIUS
        sl, tl = s.synthesis_vec_adjoint(-gbar["v"], gbar["u"])    # v = -V_theta, u = V_phi
        lam = {"div": self.f_s_from_div * sl, "vor": self.f_t_from_vor * tl}
        lam["vor"] = lam["vor"] + s.synthesis_adjoint(gbar["vor_g"])
        lam["div"] = lam["div"] + s.synthesis_adjoint(gbar["div_g"])
        lam["temp"] = s.synthesis_adjoint(gbar["t_g"]) + s.gradient_adjoint(-gbar["dtdy"] / a, gbar["dtdx"] / a)
        lam["lnps"] = (s.synthesis_adjoint(gbar["lnps_g"]) +
                       s.gradient_adjoint(-gbar["dlnpsdy"] / a, gbar["dlnpsdx"] / a))
        return lam

    def ts_adjoint(self, tbar):
        """Transpose of AerosSht.tendency_spectral."""
        s = self.st.sht
        vt, vp = s.analysis_vec_adjoint(self.f_sraw_of_tdiv * tbar["div"], self.f_traw_of_tvor * tbar["vor"])
        return {"an": (-vt).contiguous(), "ae": vp, "ek": s.analysis_adjoint(self.f_neg_lap * tbar["div"]),
                "dtdt": s.analysis_adjoint(tbar["temp"]), "dlnpsdt": s.analysis_adjoint(tbar["lnps"])}

    def mass_fix_adjoint(self, lam_lnps, lnps_prefix):
        """Transpose of ln p_s -> ln p_s + c00 ln(target / M(ln p_s)) e00 at ln p_s = lnps_prefix."""
        s = self.st.sht
        eg = torch.exp(s.synthesis(lnps_prefix)) * self.wrow       # ([nmem,] nlat, nlon)
        q = eg.sum(dim=(-2, -1))
        coef = self.c00 * lam_lnps[..., 0].real
        self.target_bar = self.target_bar + coef   # d(ln target), added to the initial ln p_s at the end
        return lam_lnps - (coef / q)[..., None] * s.synthesis_adjoint(eg)

    def mass_target_adjoint(self, lam_lnps, lnps0):
        """Adds the target's share: ln target = ln M(ln p_s at step 0)."""
        s = self.st.sht
        eg = torch.exp(s.synthesis(lnps0)) * self.wrow
        return lam_lnps + (self.target_bar / eg.sum(dim=(-2, -1)))[..., None] * s.synthesis_adjoint(eg)

    def step_adjoint(self, rec, lam_old, lam_now):
        """(adjoint of X^n filtered, adjoint of X^(n+1) after the fixer) -> (adjoints of
        X^(n-1), X^n) for the step recorded in rec; also returns the dteq sensitivity."""
        st = self.st
        s = st.sht
        lam_new = dict(lam_now)
        if st.mass_fixer:
            lam_new["lnps"] = self.mass_fix_adjoint(lam_now["lnps"], rec["lnps_prefix"])
        if rec["h"] != st.h:
            st.si.set_step(rec["h"])
            st.h = rec["h"]
        c = lambda d: {k: d[k].contiguous() for k in KEYS}
        go, gc, gt = st.si.leapfrog_adjoint(c(lam_new), c(lam_old), semi_implicit=st.semi, diffuse=st.diffuse,
                                            eps_filter=rec["eps"], raw_alpha=st.raw_alpha)
        cbar = self.ts_adjoint(gt)
        now = rec["now"]
        g = s.tendency_grid_transforms(now["vor"], now["div"], now["temp"], now["lnps"])
        res = column_terms_adjoint(s, st.vg, g, cbar, phis=st.phis, held_suarez=st.hs, dteq=st.dteq,
                                   want_teq=st.hs)
        gbar, teq = res if st.hs else (res, None)
        lt = self.tg_adjoint(gbar)
        lam_prev_now = {k: gc[k] + lt[k] for k in KEYS}
        lam_prev_old = go
        if rec["nstep"] == 0:   # the start-up step copies X^n into X^(n-1)
            lam_prev_now = {k: lam_prev_now[k] + go[k] for k in KEYS}
            lam_prev_old = {k: torch.zeros_like(v) for k, v in go.items()}
        return lam_prev_old, lam_prev_now, teq

    # ------------------------------------------------------------ driver
    def run(self, state, nsteps):
        """Forward only: J after nsteps from state (vor, div, temp, lnps)."""
        self.st.set_state(*state)
        self.st.mass_target = None
        self.st.lnr_cum = None
        for _ in range(nsteps):
            self.st.step()
        return self.objective(self.st.now)

    def gradient(self, state, nsteps, every=12):
        """J and its gradient w.r.t. the initial spectral state and dteq, with the forward
        state recomputed from checkpoints taken every `every` steps."""
        st = self.st
        st.set_state(*state)
        st.mass_target = None
        st.lnr_cum = None
        lnps0 = st.now["lnps"].clone()
        self.target_bar = torch.zeros(lnps0.shape[:-1], dtype=torch.float64, device="cuda")
        ckpts = {0: st.checkpoint()}
        for n in range(nsteps):
            st.step()
            if (n + 1) % every == 0 and n + 1 < nsteps:
                ckpts[n + 1] = st.checkpoint()
        J = self.objective(st.now)
        lam_old = {k: torch.zeros_like(v) for k, v in st.now.items()}
        lam_now = self.objective_adjoint()
        g_teq = None
        for start in sorted(ckpts, reverse=True):
            end = min(start + every, nsteps)
            st.restore(ckpts[start])
            recs = [self.step_record() for _ in range(start, end)]
            for rec in reversed(recs):
                lam_old, lam_now, teq = self.step_adjoint(rec, lam_old, lam_now)
                if teq is not None:
                    g_teq = teq if g_teq is None else g_teq + teq
        if st.mass_fixer:
            lam_now["lnps"] = self.mass_target_adjoint(lam_now["lnps"], lnps0)
        if g_teq is None:
            g_teq = torch.zeros(self.teq_shape(), dtype=torch.float64, device="cuda")
        return J, lam_now, g_teq

    def teq_shape(self):
        """Shape of a per-member dteq field: (nlev, nlat), or (nmem, nlev, nlat)."""
        return tuple(self.st.now["temp"].shape[:-1]) + (self.st.sht.nlat,)


def seeded_directions(st, n, config=3, zeta_rms=1.0e-5, temp_rms=10.0, teq_rms=10.0):
    """n reproducible directions in (initial zeta, initial T, dteq): seeded red spectra
    (seed 1000 config + field + 100000 direction) scaled to the given grid rms, and a
    seeded normal dteq field; divergence and ln p_s directions are zero."""
    import numpy as np

    from .seeding import normals, red_spectra
    s = st.sht
    out = []
    for i in range(n):
        dev = lambda a: torch.as_tensor(a, dtype=torch.complex128).cuda().contiguous()
        z = dev(red_spectra(s.trunc, st.nlev, config, member=i, field0=0, zero_l0=True))
        t = dev(red_spectra(s.trunc, st.nlev, config, member=i, field0=100, zero_l0=True))
        z *= zeta_rms / float(s.synthesis(z).pow(2).mean().sqrt())
        t *= temp_rms / float(s.synthesis(t).pow(2).mean().sqrt())
        q = normals(1000 * config + 200 + 100000 * i, st.nlev * s.nlat).real.reshape(st.nlev, s.nlat)
        d = {"vor": z, "div": torch.zeros_like(z), "temp": t,
             "lnps": torch.zeros(s.nlm, dtype=torch.complex128, device="cuda")}
        out.append((d, torch.as_tensor(np.ascontiguousarray(q) * teq_rms, dtype=torch.float64).cuda()))
    return out


def directional_check(adj, state, nsteps, directions, eps=1.0e-6, every=12):
    """Central differences (J(x + eps d) - J(x - eps d)) / (2 eps) of the forward code along
    each (state direction, dteq direction) against the adjoint's <grad, d>. Returns J, the
    gradient and one (fd, adjoint, relative difference) per direction."""
    st = adj.st
    st.set_state(*state)
    q0 = st.dteq if st.dteq is not None else torch.zeros(adj.teq_shape(), dtype=torch.float64, device="cuda")
    st.dteq = q0
    J, lam, gq = adj.gradient(state, nsteps, every=every)
    res = []
    for d, dq in directions:
        vals = []
        for sgn in (1.0, -1.0):
            st.dteq = (q0 + sgn * eps * dq).contiguous()
            vals.append(adj.run([x + sgn * eps * d[k] for x, k in zip(state, KEYS)], nsteps))
        st.dteq = q0
        fd = (vals[0] - vals[1]) / (2.0 * eps)
        ad = dot(lam, d) + dot(gq, dq)
        res.append((fd, ad, abs(fd - ad) / max(abs(ad), 1e-300)))
    return J, lam, gq, res
