"""Checkpoint / restart of the GPU dynamics step: a split run (K/2 steps, checkpoint to a
file, a fresh stepper, K/2 more) equals the straight run bit for bit at both leapfrog
levels and in the scalars -- the split-run bar of the reference's tests/test_restart.f90:1-22."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a visible MI355X"
    return t


def _stepper(T=21, nlev=8):
    from aeros_amd.column import vgrid_init
    from aeros_amd.dynamics import Stepper
    A, B = vgrid_init(nlev, stretch_a=1.0, sigma_t=0.0, p_top=1000.0)
    return Stepper(T, A, B, 1800.0, tau_diff=6 * 3600.0, eps_filter=0.06, raw_alpha=0.53,
                   held_suarez=True, mass_fixer=True)


@pytest.mark.parametrize("K", [2, 13])
def test_split_run_is_bit_identical(torch, tmp_path, K):
    from aeros_amd.dynamics import load_checkpoint, save_checkpoint
    from aeros_amd.held_suarez import initial_state
    a = _stepper()
    ic = initial_state(a.sht, a.nlev)
    a.set_state(*ic)
    for _ in range(K):
        a.step()
    b = _stepper()
    b.set_state(*ic)
    for _ in range(K // 2):
        b.step()
    path = tmp_path / "ck.npz"
    save_checkpoint(b.checkpoint(), path)
    del b
    c = _stepper()
    c.restore(load_checkpoint(path))
    for _ in range(K - K // 2):
        c.step()
    assert c.nstep == a.nstep == K
    assert c.mass_target == a.mass_target
    assert float(c.lnr_cum) == float(a.lnr_cum)
    for lev in ("old", "now"):
        for k in ("vor", "div", "temp", "lnps"):
            assert torch.equal(getattr(c, lev)[k], getattr(a, lev)[k]), (lev, k)


def test_checkpoint_shape_is_checked(torch):
    a = _stepper(T=21, nlev=8)
    from aeros_amd.held_suarez import initial_state
    a.set_state(*initial_state(a.sht, a.nlev))
    a.step()
    ck = a.checkpoint()
    b = _stepper(T=31, nlev=8)
    with pytest.raises(ValueError):
        b.restore(ck)
