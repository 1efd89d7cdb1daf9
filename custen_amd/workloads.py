"""Synthetic workloads of the benchmark configurations (shared by bench.py
and the tests): sizes, seeds and weights as the measurement plan states them.

  batched grids : 512 square grids, N uniform over {128,...,2048} drawn with
                  SplitMix64 seed 4, grid g filled from seed 1000 + g,
                  13-point biharmonic (5x5), periodic, 100 steps.
  served CH     : Poisson arrivals (seed 5), N uniform over {256,...,2048},
                  c0 ~ U[-0.1, 0.1) from seed 5000 + r, 200 steps each.
"""
from __future__ import annotations

import numpy as np

from . import synth

BATCH_SIZES = (128, 256, 512, 1024, 2048)
CH_SIZES = (256, 512, 1024, 2048)


def biharmonic_weights(n: int) -> np.ndarray:
    """13-point biharmonic as 5x5 weights / h^4, h = 2 pi / n."""
    h = 2 * np.pi / n
    pat = np.array([[0, 0, 1, 0, 0], [0, 2, -8, 2, 0], [1, -8, 20, -8, 1],
                    [0, 2, -8, 2, 0], [0, 0, 1, 0, 0]], dtype=np.float64)
    return (pat / h ** 4).ravel()


def laplacian9_weights(n: int) -> np.ndarray:
    """9-point Laplacian [1,4,1;4,-20,4;1,4,1] / (6 h^2)."""
    h = 2 * np.pi / n
    return np.array([1, 4, 1, 4, -20, 4, 1, 4, 1], dtype=np.float64) / (6 * h * h)


def batch_sizes(count: int = 512, seed: int = 4) -> list[int]:
    z = synth.splitmix64(count, seed)
    return [BATCH_SIZES[int(v % np.uint64(len(BATCH_SIZES)))] for v in z]


def ch_requests(count: int, rate_per_s: float, seed: int = 5, sizes=CH_SIZES):
    """(arrival_time_s, n, ic_seed) for `count` Poisson arrivals."""
    u = synth.uniform(2 * count, seed, 0.0, 1.0)
    gaps = -np.log1p(-u[:count]) / rate_per_s
    t = np.cumsum(gaps)
    t -= t[0]
    ns = [sizes[int(np.floor(v * len(sizes))) % len(sizes)] for v in u[count:]]
    return [(float(t[r]), ns[r], 5000 + r) for r in range(count)]
