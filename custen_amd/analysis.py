"""Coarsening analysis of Cahn-Hilliard snapshots, the counterpart of the
reference's cuPentCahnADI/plotting.py:27-86.

For each snapshot c (an n x n periodic field on [0, lx)^2) two length-scale
measures are computed:

  s(t)  = 1 / (1 - <c^2>)                      (plotting.py:64-66)
  1/k_1 = sum(|k|^-1 P(k)) / sum(P(k))         (plotting.py:68-74)

where P(k) = |FFT2(c)|^2 over the wavenumbers with kx != 0 and ky != 0 (the
reference drops the first row and column of the spectrum, [1:n, 1:n]), and
|k| = sqrt(kx^2 + ky^2) with integer wavenumbers from fftfreq(n, 1/n).

<c^2> is the domain average.  method="periodic" takes the exact average of the
periodic samples (the mean); method="simpson" follows the reference, which
integrates with Simpson's rule over linspace(0, 2 pi, n) and divides by
(2 pi)^2 (plotting.py:64-65).

Snapshots are the .npy files custenCHSnapshot writes; `snapshot_name` gives the
reference's naming (cahn_hilliard_<time with 10 decimals>).

  python -m custen_amd.analysis DIR [--method periodic|simpson] [--lx L]
prints one JSON object per snapshot, in time order.
"""
from __future__ import annotations

import argparse
import json
import os
import re

import numpy as np


def snapshot_name(time: float, ext: str = ".npy") -> str:
    """File name of the snapshot at `time` (cuPentCahnADI.cu:107-110 naming)."""
    return f"cahn_hilliard_{time:0.10f}{ext}"


def mean_square(c: np.ndarray, method: str = "periodic", lx: float = 2 * np.pi) -> float:
    c = np.asarray(c, dtype=np.float64)
    if method == "periodic":
        return float(np.mean(np.square(c)))
    if method == "simpson":
        from scipy.integrate import simpson

        n = c.shape[0]
        x = np.linspace(0, lx, n)
        y = np.linspace(0, lx, c.shape[1])
        return float(simpson(simpson(np.square(c), x=y), x=x) / (lx * lx))
    raise ValueError(f"unknown method {method!r}")


def inverse_k1(c: np.ndarray) -> float:
    c = np.asarray(c, dtype=np.float64)
    n = c.shape[0]
    if c.shape != (n, n):
        raise ValueError("square field expected")
    k = np.fft.fftfreq(n, 1.0 / n)
    kx = k[1:][None, :]
    ky = k[1:][:, None]
    inv_mod = 1.0 / np.sqrt(kx * kx + ky * ky)
    power = np.square(np.abs(np.fft.fft2(c)))[1:n, 1:n]
    numer = float(np.sum(power))
    return float(np.sum(inv_mod * power)) / numer if numer > 0 else float("nan")


def coarsening(c: np.ndarray, method: str = "periodic", lx: float = 2 * np.pi) -> dict:
    avg = mean_square(c, method, lx)
    return {"s": 1.0 / (1.0 - avg), "inv_k1": inverse_k1(c), "mean_square": avg}


def load_series(directory: str):
    """[(time, field)] of the cahn_hilliard_<time>.npy files in `directory`,
    sorted by time (plotting.py:35-40)."""
    out = []
    for name in os.listdir(directory):
        m = re.match(r"cahn_hilliard_(\d+\.\d+)\.npy$", name)
        if m:
            out.append((float(m.group(1)), os.path.join(directory, name)))
    out.sort()
    return [(t, np.load(p, allow_pickle=False)) for t, p in out]


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("directory")
    ap.add_argument("--method", default="periodic", choices=["periodic", "simpson"])
    ap.add_argument("--lx", type=float, default=2 * np.pi)
    args = ap.parse_args(argv)
    for t, c in load_series(args.directory):
        row = {"t": t, "t^(1/3)": t ** (1.0 / 3.0)}
        row.update(coarsening(c, args.method, args.lx))
        print(json.dumps(row))


if __name__ == "__main__":
    main()
