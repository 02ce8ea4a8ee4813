"""Write the 50-digit golden vectors that pin the FP64 oracle.

Brute force in mpmath, from textbook formulas that share no code path with
the oracle's recurrences or FFT:

* Gauss-Legendre nodes: Newton on the explicit Legendre polynomial sum, 50 digits.
* Normalized associated Legendre functions, no Condon-Shortley phase:
  P_lm(x) = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) (1-x^2)^{m/2} d^m/dx^m P_l(x),
  with P_l(x) = 2^-l sum_k (-1)^k C(l,k) C(2l-2k,l) x^(l-2k) differentiated term
  by term; dP_lm/dtheta from the product rule on the same closed form.
* Transforms as direct sums over every (l, m) and every longitude, no FFT.

Inputs are the reference's own deterministic test spectrum
(tests/test_spectral.f90:147-172) and a seeded red spectrum. Run once in the
build container (python tests/golden/make_golden.py); the .npz files it writes
are committed.
"""
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from aeros_amd.seeding import fill_test_spectrum, lm_tables, red_spectrum  # noqa: E402

mp.mp.dps = 50


def legendre_poly(n, x):
    s = mp.mpf(0)
    for k in range(n // 2 + 1):
        s += (-1) ** k * mp.binomial(n, k) * mp.binomial(2 * n - 2 * k, n) * x ** (n - 2 * k)
    return s / mp.mpf(2) ** n


def dlegendre_poly(n, x):
    s = mp.mpf(0)
    for k in range(n // 2 + 1):
        p = n - 2 * k
        if p >= 1:
            s += (-1) ** k * mp.binomial(n, k) * mp.binomial(2 * n - 2 * k, n) * p * x ** (p - 1)
    return s / mp.mpf(2) ** n


def gauss_nodes(n):
    """Descending nodes (north -> south) and weights, 50 digits."""
    mus, ws = [], []
    for i in range(n):
        x = mp.cos(mp.pi * (i + mp.mpf(3) / 4) / (n + mp.mpf(1) / 2))
        for _ in range(100):
            dx = legendre_poly(n, x) / dlegendre_poly(n, x)
            x -= dx
            if abs(dx) < mp.mpf(10) ** -45:
                break
        d = dlegendre_poly(n, x)
        mus.append(x)
        ws.append(2 / ((1 - x * x) * d * d))
    return mus, ws


def coeff_deriv(l, m):
    """Coefficients of d^m/dx^m P_l(x) as {power: coefficient}."""
    out = {}
    for k in range(l // 2 + 1):
        p = l - 2 * k
        if p < m:
            continue
        c = mp.mpf((-1) ** k) * mp.binomial(l, k) * mp.binomial(2 * l - 2 * k, l) / mp.mpf(2) ** l
        c *= mp.factorial(p) / mp.factorial(p - m)
        out[p - m] = out.get(p - m, 0) + c
    return out


def poly_eval(cs, x):
    return sum((c * x ** p for p, c in cs.items()), mp.mpf(0))


def poly_deriv(cs):
    return {p - 1: c * p for p, c in cs.items() if p >= 1}


def plm_table(T, mus):
    """P[lm][j], D[lm][j] (D = dP/dtheta) in mpmath."""
    P, D = [], []
    for m in range(T + 1):
        for l in range(m, T + 1):
            norm = mp.sqrt((2 * l + 1) / (4 * mp.pi) * mp.factorial(l - m) / mp.factorial(l + m))
            Q = coeff_deriv(l, m)
            dQ = poly_deriv(Q)
            prow, drow = [], []
            for x in mus:
                s = mp.sqrt(1 - x * x)
                q = poly_eval(Q, x)
                dq = poly_eval(dQ, x)
                prow.append(norm * s ** m * q)
                # d/dtheta [s^m Q(x)] = m x s^(m-1) Q - s^(m+1) Q'
                dval = -s ** (m + 1) * dq
                if m > 0:
                    dval += m * x * s ** (m - 1) * q
                drow.append(norm * dval)
            P.append(prow)
            D.append(drow)
    return P, D


def to_mpc(c):
    return [mp.mpc(float(z.real), float(z.imag)) for z in c]


def synth(T, nlon, mus, P, c, which="scalar", D=None, t=None):
    """Direct-sum synthesis. which: scalar -> f; vector -> (vt, vp) from (S=c, T=t)."""
    nlat = len(mus)
    ls, ms = lm_tables(T)
    phis = [2 * mp.pi * i / nlon for i in range(nlon)]
    outs = [[[mp.mpf(0)] * nlon for _ in range(nlat)] for _ in range(1 if which == "scalar" else 2)]
    for j, x in enumerate(mus):
        s = mp.sqrt(1 - x * x)
        F = [[mp.mpc(0)] * (T + 1) for _ in range(len(outs))]
        for k in range(len(ls)):
            l, m = int(ls[k]), int(ms[k])
            if which == "scalar":
                F[0][m] += c[k] * P[k][j]
            else:
                q = m * P[k][j] / s
                tt = t[k] if t is not None else mp.mpc(0)
                F[0][m] += c[k] * D[k][j] + 1j * q * tt
                F[1][m] += 1j * q * c[k] - tt * D[k][j]
        for o in range(len(outs)):
            for i, ph in enumerate(phis):
                v = mp.re(F[o][0])
                for m in range(1, T + 1):
                    v += 2 * mp.re(F[o][m] * mp.expj(m * ph))
                outs[o][j][i] = v
    return [np.array([[float(v) for v in row] for row in out]) for out in outs]


def anal(T, nlon, mus, ws, P, grid):
    nlat = len(mus)
    ls, ms = lm_tables(T)
    G = []
    for j in range(nlat):
        row = [mp.mpf(float(v)) for v in grid[j]]
        Gj = []
        for m in range(T + 1):
            acc = mp.mpc(0)
            for i in range(nlon):
                acc += row[i] * mp.expj(-2 * mp.pi * m * i / nlon)
            Gj.append(acc * 2 * mp.pi / nlon)
        G.append(Gj)
    out = []
    for k in range(len(ls)):
        m = int(ms[k])
        acc = mp.mpc(0)
        for j in range(nlat):
            acc += ws[j] * P[k][j] * G[j][m]
        out.append(complex(acc))
    return np.array(out)


def make(T, nlon, nlat, path):
    print(f"T{T}: {nlon}x{nlat}", flush=True)
    mus, ws = gauss_nodes(nlat)
    P, D = plm_table(T, mus)
    fill = fill_test_spectrum(T)
    red = red_spectrum(T, seed=424242)
    s_fill = fill.copy()
    t_fill = np.conj(fill) * 0.5
    l, m = lm_tables(T)
    t_fill[m == 0] = t_fill[m == 0].real
    s_fill[l == 0] = 0
    t_fill[l == 0] = 0
    (g_fill,) = synth(T, nlon, mus, P, to_mpc(fill))
    (g_red,) = synth(T, nlon, mus, P, to_mpc(red))
    vt, vp = synth(T, nlon, mus, P, to_mpc(s_fill), "vector", D, to_mpc(t_fill))
    gt, gp = synth(T, nlon, mus, P, to_mpc(fill), "vector", D, None)
    # analysis of a grid that is NOT band-limited: the red grid plus a product term
    g_in = g_red + 0.3 * g_red * g_fill
    a_in = anal(T, nlon, mus, ws, P, g_in)
    np.savez_compressed(
        path,
        trunc=T, nlon=nlon, nlat=nlat,
        mu=np.array([float(x) for x in mus]), w=np.array([float(x) for x in ws]),
        P=np.array([[float(v) for v in r] for r in P]), D=np.array([[float(v) for v in r] for r in D]),
        spec_fill=fill, spec_red=red, grid_fill=g_fill, grid_red=g_red,
        vec_s=s_fill, vec_t=t_fill, vec_vt=vt, vec_vp=vp,
        grad_gt=gt, grad_gp=gp, anal_grid=g_in, anal_spec=a_in)


if __name__ == "__main__":
    make(5, 16, 8, os.path.join(HERE, "sht_T5.npz"))
    make(21, 64, 32, os.path.join(HERE, "sht_T21.npz"))
