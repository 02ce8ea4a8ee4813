"""Seeded synthetic states for the transform path.

A counter-based SplitMix64 stream with Box-Muller normals, so every field of
every member is reproducible from (config, field id, member) alone:
seed = 1000*config + field_id + 100000*member.

Spectral fields follow the red spectrum used for the benchmark inputs:
c_lm = (N(0,1) + i N(0,1)) * (l+1)^-1.5 with the m = 0 imaginary parts zeroed
(a real field has real m = 0 coefficients, reference tests/test_spectral.f90:166-167),
and l = 0 optionally zeroed for vector potentials (reference
src/dynamics/aeros_vordiv.f90:44-51 discards it).
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs of the stream started at `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def normals(seed: int, n: int) -> np.ndarray:
    """n complex numbers whose real and imaginary parts are independent N(0,1)."""
    z = splitmix64(seed, 2 * n)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2.0 * np.pi * u2) + 1j * (r * np.sin(2.0 * np.pi * u2))


def lm_tables(trunc: int):
    """0-based (l_of_lm, m_of_lm), m-major with l fastest."""
    ls, ms = [], []
    for m in range(trunc + 1):
        for l in range(m, trunc + 1):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)


def red_spectrum(trunc: int, seed: int, zero_l0: bool = False, slope: float = -1.5) -> np.ndarray:
    """One seeded spectral field (nlm,) complex128."""
    l, m = lm_tables(trunc)
    c = normals(seed, l.size) * (l + 1.0) ** slope
    c[m == 0] = c[m == 0].real
    if zero_l0:
        c[l == 0] = 0.0
    return c


def red_spectra(trunc: int, nfield: int, config: int = 1, member: int = 0,
                field0: int = 0, zero_l0: bool = False) -> np.ndarray:
    """(nfield, nlm) complex128; field k uses seed 1000*config + field0 + k + 100000*member."""
    return np.stack([red_spectrum(trunc, 1000 * config + field0 + k + 100000 * member, zero_l0)
                     for k in range(nfield)])


def fill_test_spectrum(trunc: int) -> np.ndarray:
    """The reference's deterministic test spectrum (tests/test_spectral.f90:147-172):
    amplitude 1/(l+1)^2, value cos(3l+m) + i sin(l+2m), m = 0 real."""
    l, m = lm_tables(trunc)
    amp = 1.0 / ((l + 1.0) ** 2)
    c = amp * np.cos(3.0 * l + m) + 1j * amp * np.sin(l + 2.0 * m)
    c[m == 0] = c[m == 0].real
    return c
