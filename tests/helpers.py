"""Shared test helpers: exact-arithmetic emulation of the oracle's rounding
contract (used to pin the oracle on tiny inputs) and comparison utilities."""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def round_binary(v: Fraction, mant: int, emin: int) -> float:
    """Round an exact rational to the nearest binary floating-point number with
    `mant` significand bits (ties to even), minimum normal exponent `emin`."""
    if v == 0:
        return 0.0
    sign = -1 if v < 0 else 1
    a = abs(v)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, emin)
    scale = Fraction(2) ** (mant - 1 - e)
    q = a * scale
    n = q.numerator // q.denominator
    rem = q - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return sign * float(Fraction(n) / scale)


def fma64(a: float, b: float, c: float) -> float:
    """Correctly rounded fused multiply-add in binary64 (exact rationals)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def fma32(a, b, c) -> np.float32:
    return np.float32(round_binary(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)),
                                   24, -126))


def brute_stencil(u, w, H, L, V, T, periodic, fma=fma64, zero=0.0):
    """Definition of the stencil apply, evaluated with exact-rational fma:
    out(y,x) = fma chain over j<V (outer), i<H (inner) from 0, periodic wrap or
    interior-only.  Returns (out, written-mask)."""
    ny, nx = u.shape
    R, B = H - 1 - L, V - 1 - T
    out = np.zeros_like(u)
    mask = np.zeros(u.shape, dtype=bool)
    for y in range(ny):
        for x in range(nx):
            if not periodic and not (T <= y < ny - B and L <= x < nx - R):
                continue
            s = u.dtype.type(zero)
            for j in range(V):
                for i in range(H):
                    yy, xx = y - T + j, x - L + i
                    if periodic:
                        yy %= ny
                        xx %= nx
                    s = fma(w[j * H + i], u[yy, xx], s)
            out[y, x] = s
            mask[y, x] = True
    return out, mask


def bits_equal(a, b) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    it = np.uint64 if a.dtype == np.float64 else np.uint32
    return bool(np.array_equal(a.view(it), b.view(it)))


def mismatches(a, b) -> int:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    it = np.uint64 if a.dtype == np.float64 else np.uint32
    return int(np.count_nonzero(a.view(it) != b.view(it)))


def scaled_error(y, yref, u, w, H, L, V, T, periodic):
    """max |y - yref| / max(|yref|, S) with S = sum_k |w_k| |u_k| (the rounding
    scale of the sum), over the points the reference writes."""
    ny, nx = u.shape
    absu = np.abs(u.astype(np.float64))
    S = np.zeros((ny, nx))
    for j in range(V):
        for i in range(H):
            S += abs(float(w[j * H + i])) * np.roll(np.roll(absu, T - j, axis=0), L - i, axis=1)
    den = np.maximum(np.abs(yref.astype(np.float64)), S)
    den[den == 0] = 1.0
    err = np.abs(y.astype(np.float64) - yref.astype(np.float64)) / den
    if not periodic:
        R, B = H - 1 - L, V - 1 - T
        err = err[T:ny - B, L:nx - R]
    return float(err.max()) if err.size else 0.0
