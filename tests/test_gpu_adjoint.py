"""GPU: the adjoint (transpose) transforms satisfy <A x, y> = <x, A^T y> in the real
inner products (spectral: sum of re*re + im*im over all coefficients; grid:
unweighted sum), for every transform kind, truncation and precision, and the
transposes of the two scalar transforms are tied to each other through the
quadrature weights."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def sdot(a, b):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    return float(np.sum(a.real * b.real + a.imag * b.imag))


def gdot(a, b):
    return float(np.sum(a.cpu().numpy() * b.cpu().numpy()))


def check(lhs, rhs, scale, tol):
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs, abs(lhs - rhs) / scale)


@pytest.mark.parametrize("T,prec,tol", [(21, "fp64", 1e-13), (29, "fp64", 1e-13), (42, "fp64", 1e-13),
                                        (106, "fp64", 1e-13), (42, "split", 1e-13), (42, "fp32", 1e-5)])
def test_dot_product_identities(torch, T, prec, tol):
    from aeros_amd.spectral import AerosSht
    s = AerosSht(T, precision=prec)
    rng = np.random.default_rng(T)
    nf = 3
    spec = lambda: torch.from_numpy(rng.standard_normal((nf, s.nlm)) + 1j * rng.standard_normal((nf, s.nlm))).cuda()
    grid = lambda: torch.from_numpy(rng.standard_normal((nf, s.nlat, s.nlon))).cuda()
    c, c2, g, g2 = spec(), spec(), grid(), grid()
    # scale of each identity: sum of |terms|, so the bound is relative to the work done
    sc = lambda x, y: float(np.sum(np.abs(x.cpu().numpy()) * np.abs(y.cpu().numpy())))
    # scalar synthesis and analysis
    Sc = s.synthesis(c)
    check(gdot(Sc, g), sdot(c, s.synthesis_adjoint(g)), sc(Sc, g), tol)
    Ag = s.analysis(g)
    check(sdot(Ag, c), gdot(g, s.analysis_adjoint(c)), sc(Ag, c), tol)
    # vector synthesis and analysis
    vt, vp = s.synthesis_vec(c, c2)
    st, tt = s.synthesis_vec_adjoint(g, g2)
    check(gdot(vt, g) + gdot(vp, g2), sdot(c, st) + sdot(c2, tt), sc(vt, g) + sc(vp, g2), tol)
    sa, ta = s.analysis_vec(g, g2)
    at, ap = s.analysis_vec_adjoint(c, c2)
    check(sdot(sa, c) + sdot(ta, c2), gdot(g, at) + gdot(g2, ap), sc(sa, c) + sc(ta, c2), tol)
    # gradient
    gt, gp = s.gradient(c)
    check(gdot(gt, g) + gdot(gp, g2), sdot(c, s.gradient_adjoint(g, g2)), sc(gt, g) + sc(gp, g2), tol)


def test_adjoints_through_the_quadrature(torch):
    """analysis^T c = W synthesis(c / w_m) and synthesis^T g = w_m analysis(g / W) (FP64)."""
    from aeros_amd.spectral import AerosSht
    T = 31
    s = AerosSht(T)
    rng = np.random.default_rng(1)
    c = rng.standard_normal(s.nlm) + 1j * rng.standard_normal(s.nlm)
    m = np.array([mm for mm in range(T + 1) for _ in range(mm, T + 1)])
    wm = np.where(m == 0, 1.0, 2.0)
    W = (s.gauss_w * 2 * np.pi / s.nlon)[:, None]
    got = s.analysis_adjoint(torch.from_numpy(c).cuda()).cpu().numpy()
    ref = W * s.synthesis(torch.from_numpy(c / wm).cuda()).cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    g = rng.standard_normal((s.nlat, s.nlon))
    got = s.synthesis_adjoint(torch.from_numpy(g).cuda()).cpu().numpy()
    ref = wm * s.analysis(torch.from_numpy(g / W).cuda()).cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
