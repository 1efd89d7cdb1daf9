"""Synthetic inputs shared by the tests, smoke() and bench.py.

SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): value i is mixed from the state
seed + (i + 1) * 0x9E3779B97F4A7C15 (mod 2^64); u = (z >> 11) * 2^-53 and the
sample is a + (b - a) * u, evaluated unfused.  The C library exports the same
stream as custenFillUniform for large fills.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(n: int, seed: int, start: int = 0) -> np.ndarray:
    """Raw 64-bit outputs [start, start + n) of the stream for `seed`."""
    i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(n: int, seed: int, a: float, b: float, start: int = 0) -> np.ndarray:
    u = (splitmix64(n, seed, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return np.float64(a) + np.float64(b - a) * u


def field(nx: int, ny: int, seed: int, a: float = -1.0, b: float = 1.0) -> np.ndarray:
    """ny x nx row-major FP64 field."""
    return uniform(nx * ny, seed, a, b).reshape(ny, nx)
