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
CKPT_VERSION = 1


class Stepper:
    """State (X^(n-1), X^n) of one model on one GPU and the step that advances it."""

    def __init__(self, trunc, a_half, b_half, dt, t_ref=300.0, ps_ref=1.0e5, si_alpha=0.5,
                 tau_diff=0.0, tau_diff_div=0.0, ndiff=6, eps_filter=0.0, raw_alpha=1.0,
                 held_suarez=True, mass_fixer=False, semi_implicit=True, precision="fp64", dteq=None):
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
        self.dteq = dteq          # optional additive Held-Suarez T_eq [nlev, nlat] (CUDA float64)
        self.nstep = 0
        self.mass_target = None
        self.lnr_cum = None       # running sum of the mass fixer's ln r (device scalar)
        self.now = self.old = None

    def set_state(self, vor, div, temp, lnps):
        """X^n from spectral (nlev, nlm) / (nlm,) arrays (numpy or CUDA tensors). An
        ensemble of members of this size passes (nmem, nlev, nlm) / (nmem, nlm): every
        launch of a step then covers all members."""
        dev = lambda a: torch.as_tensor(a, dtype=torch.complex128).cuda().contiguous()
        now = {"vor": dev(vor), "div": dev(div), "temp": dev(temp), "lnps": dev(lnps)}
        self.si._members(now)
        self.now = now
        self.old = None
        self.nstep = 0
        self.mass_target = None
        self.lnr_cum = None

    @property
    def members(self):
        """Ensemble size, or 0 for a single model."""
        return self.now["vor"].shape[0] if self.now is not None and self.now["vor"].dim() == 3 else 0

    def _mass_target(self, lnps):
        m = mass(self.sht, lnps)[0]
        return m.item() if m.dim() == 0 else m.clone()

    def step(self):
        s, now = self.sht, self.now
        if self.nstep == 0:
            h = self.dt
            self.old = {k: v.clone() for k, v in now.items()}
            if self.mass_fixer and self.mass_target is None:
                self.mass_target = self._mass_target(now["lnps"])
        else:
            h = 2.0 * self.dt
        if h != self.h:
            self.si.set_step(h)
            self.h = h
        g = s.tendency_grid_transforms(now["vor"], now["div"], now["temp"], now["lnps"])
        c = column_terms(s, self.vg, g, phis=self.phis, held_suarez=self.hs, dteq=self.dteq)
        tnd = s.tendency_spectral(c["ae"], c["an"], c["ek"], c["dtdt"], c["dlnpsdt"])
        eps = self.eps if self.nstep > 0 else 0.0
        new = self.si.leapfrog(self.old, now, tnd, semi_implicit=self.semi, diffuse=self.diffuse,
                               eps_filter=eps, raw_alpha=self.raw_alpha)
        self.old, self.now = now, new
        if self.mass_fixer:
            _, lnr = mass(s, self.now["lnps"], fix=True, target=self.mass_target)
            self.lnr_cum = lnr if self.lnr_cum is None else self.lnr_cum + lnr
        self.nstep += 1
        self.last_grid = g
        return self.now

    # ------------------------------------------------------------ checkpoint / restart
    def checkpoint(self):
        """The integrator state a resumed run needs to continue bit-identically, as in
        aeros_timestep_write_restart (src/dynamics/aeros_timestep.f90:1303-1337) for the dry
        core: both leapfrog levels X^(n-1) and X^n, nstep, the mass-fixer target and its
        accumulated ln r. Device copies; grid fields are recomputed every step and are not
        part of it."""
        if self.now is None:
            raise RuntimeError("no state to checkpoint")
        return {"version": CKPT_VERSION, "nstep": self.nstep,
                "old": None if self.old is None else {k: v.clone() for k, v in self.old.items()},
                "now": {k: v.clone() for k, v in self.now.items()},
                "mass_target": self.mass_target,
                "lnr_cum": None if self.lnr_cum is None else self.lnr_cum.clone(),
                "shape": (self.nlev, self.sht.nlm, self.sht.nlat, self.sht.nlon), "members": self.members}

    def restore(self, ck):
        """Resume from checkpoint(); the shapes must match this stepper's."""
        if ck.get("version") != CKPT_VERSION:
            raise ValueError("checkpoint version mismatch")
        if tuple(ck["shape"]) != (self.nlev, self.sht.nlm, self.sht.nlat, self.sht.nlon):
            raise ValueError(f"checkpoint of shape {tuple(ck['shape'])} does not fit this stepper")
        if self.now is not None and int(ck.get("members", 0)) != self.members:
            raise ValueError("checkpoint of another ensemble size")
        self.now = {k: v.clone() for k, v in ck["now"].items()}
        self.old = None if ck["old"] is None else {k: v.clone() for k, v in ck["old"].items()}
        self.nstep = int(ck["nstep"])
        self.mass_target = ck["mass_target"]
        self.lnr_cum = None if ck["lnr_cum"] is None else ck["lnr_cum"].clone()


def save_checkpoint(ck, path):
    """A checkpoint to a .npz file: spectral levels as complex128 arrays, scalars and the
    shape (nlev, nlm, nlat, nlon) as metadata. Loaded without pickle."""
    import numpy as np
    host = lambda x: np.array(np.nan) if x is None else (x.cpu().numpy() if torch.is_tensor(x) else np.array(x))
    arrs = {"version": np.array(ck["version"]), "nstep": np.array(ck["nstep"]),
            "shape": np.array(ck["shape"], dtype=np.int64), "members": np.array(int(ck.get("members", 0))),
            "mass_target": host(ck["mass_target"]), "lnr_cum": host(ck["lnr_cum"])}
    for lev in ("old", "now"):
        if ck[lev] is not None:
            for k, v in ck[lev].items():
                arrs[f"{lev}_{k}"] = v.cpu().numpy()
    np.savez(path, **arrs)


def load_checkpoint(path, device="cuda"):
    import numpy as np
    z = np.load(path, allow_pickle=False)
    levs = {}
    for lev in ("old", "now"):
        keys = [k for k in HS_KEYS if f"{lev}_{k}" in z.files]
        levs[lev] = {k: torch.from_numpy(z[f"{lev}_{k}"]).to(device) for k in keys} if keys else None
    members = int(z["members"]) if "members" in z.files else 0
    mt, lc = z["mass_target"], z["lnr_cum"]
    if members:   # per-member device arrays
        mass_target = None if mt.ndim == 0 else torch.from_numpy(mt).to(device)
        lnr_cum = None if lc.ndim == 0 else torch.from_numpy(lc).to(device)
    else:
        mass_target = None if np.isnan(float(mt)) else float(mt)
        lnr_cum = None if np.isnan(float(lc)) else torch.tensor(float(lc), dtype=torch.float64, device=device)
    return {"version": int(z["version"]), "nstep": int(z["nstep"]), "old": levs["old"], "now": levs["now"],
            "mass_target": mass_target, "lnr_cum": lnr_cum, "members": members,
            "shape": tuple(int(x) for x in z["shape"])}
