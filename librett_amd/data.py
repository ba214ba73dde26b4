"""Seeded synthetic tensors for benchmarks and tests.

Values are uniform in [-1, 1) and come from a counter-based splitmix64 hash
(Steele, Lea and Flood, OOPSLA 2014): draw k of stream ``seed`` is
``mix(seed*G + (k+1)*G)`` with ``G = 0x9E3779B97F4A7C15``. float32 takes the
top 24 bits, float64 the top 53 bits, complex128 two consecutive float64
draws. The stream is a pure function of (seed, k), so any slice can be
generated independently.
"""
import numpy as np

G = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

DTYPES = {4: np.float32, 8: np.float64, 16: np.complex128}


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def draws(seed, n, start=0, chunk=1 << 24):
    out = np.empty(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * G
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            k = np.arange(start + s + 1, start + e + 1, dtype=np.uint64)
            out[s:e] = _mix(base + k * G)
    return out


def uniform(nelem, elem, seed):
    """Flat array of ``nelem`` elements of ``elem`` bytes (4, 8 or 16)."""
    if elem == 4:
        u = draws(seed, nelem)
        return ((u >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0) * 2.0 - 1.0).astype(np.float32)
    if elem in (8, 16):
        n = nelem * (elem // 8)
        u = draws(seed, n)
        f = (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0
        return f.view(np.complex128) if elem == 16 else f
    raise ValueError(elem)
