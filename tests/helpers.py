"""Shared test utilities."""
import numpy as np


def rel(a, ref):
    """The reference's error measure: max|x - x_ref| / max|x_ref| (tests/test_spectral.f90:81)."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    d = np.abs(ref).max()
    return float(np.abs(a - ref).max() / (d if d > 0 else 1.0))


def lm_tables(T):
    ls, ms = [], []
    for m in range(T + 1):
        for l in range(m, T + 1):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)
