"""One semi-implicit leapfrog dynamics step on the GPU, in the order of the
reference's aeros_timestep_step (src/dynamics/aeros_timestep.f90:776-1137) for
the dry core with Held-Suarez forcing:

  1. spectral -> grid transform half   (aeros_tendency_grid, transforms)
     column terms + Held-Suarez        (column_terms, aeros_hs_apply)
     grid -> spectral transform half   (aeros_tendency_spectral)
  2. vorticity leapfrog + semi-implicit solve of D, T, ln p_s
  3. implicit horizontal diffusion
  4. Robert-Asselin-Williams filter    (from the second step on)
  5. shift the time levels
  7. mass fixer                        (optional)

The start-up step is forward over dt; every later step is a leapfrog over 2 dt
(:825-840). Everything stays on the device; the only host work is the launches.
"""
import torch

from .column import VGrid, column_terms
from .spectral import AerosSht
from .step import SemiImplicit, mass

HS_KEYS = ("vor", "div", "temp", "lnps")


class Stepper:
    """State (X^(n-1), X^n) of one model on one GPU and the step that advances it."""

    def __init__(self, trunc, a_half, b_half, dt, t_ref=300.0, ps_ref=1.0e5, si_alpha=0.5,
                 tau_diff=0.0, tau_diff_div=0.0, ndiff=6, eps_filter=0.0, raw_alpha=1.0,
                 held_suarez=True, mass_fixer=False, semi_implicit=True, precision="fp64"):
        self.sht = AerosSht(trunc, precision=precision)
        self.vg = VGrid(a_half, b_half)
        self.nlev = self.vg.nlev
        self.dt = dt
        self.si = SemiImplicit(self.sht, self.vg, dt, t_ref=t_ref, ps_ref=ps_ref, alpha=si_alpha)
        self.h = dt
        self.diffuse = tau_diff > 0.0
        if self.diffuse:
            self.si.set_diffusion(tau_diff, tau_diff_div, ndiff)
        self.eps, self.raw_alpha = eps_filter, raw_alpha
        self.hs, self.mass_fixer, self.semi = held_suarez, mass_fixer, semi_implicit
        self.phis = None
        self.nstep = 0
        self.mass_target = None
        self.now = self.old = None

    def set_state(self, vor, div, temp, lnps):
        """X^n from spectral (nlev, nlm) / (nlm,) arrays (numpy or CUDA tensors)."""
        dev = lambda a: torch.as_tensor(a, dtype=torch.complex128).cuda().contiguous()
        self.now = {"vor": dev(vor), "div": dev(div), "temp": dev(temp), "lnps": dev(lnps)}
        self.old = None
        self.nstep = 0

    def step(self):
        s, now = self.sht, self.now
        if self.nstep == 0:
            h = self.dt
            self.old = {k: v.clone() for k, v in now.items()}
            if self.mass_fixer and self.mass_target is None:
                self.mass_target = mass(s, now["lnps"])[0].item()
        else:
            h = 2.0 * self.dt
        if h != self.h:
            self.si.set_step(h)
            self.h = h
        g = s.tendency_grid_transforms(now["vor"], now["div"], now["temp"], now["lnps"])
        c = column_terms(s, self.vg, g, phis=self.phis, held_suarez=self.hs)
        tnd = s.tendency_spectral(c["ae"], c["an"], c["ek"], c["dtdt"], c["dlnpsdt"])
        eps = self.eps if self.nstep > 0 else 0.0
        new = self.si.leapfrog(self.old, now, tnd, semi_implicit=self.semi, diffuse=self.diffuse,
                               eps_filter=eps, raw_alpha=self.raw_alpha)
        self.old, self.now = now, new
        if self.mass_fixer:
            mass(s, self.now["lnps"], fix=True, target=self.mass_target)
        self.nstep += 1
        self.last_grid = g
        return self.now
