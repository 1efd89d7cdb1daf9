"""Coarsening analysis (custen_amd/analysis.py, the counterpart of
cuPentCahnADI/plotting.py) against closed forms."""
import os

import numpy as np
import pytest

from custen_amd import analysis


def grid(n):
    x = np.arange(n) * (2 * np.pi / n)
    return np.meshgrid(x, x)  # X along columns, Y along rows


def test_single_mode_closed_forms():
    X, Y = grid(64)
    c = 0.8 * np.cos(3 * X) * np.cos(4 * Y)
    r = analysis.coarsening(c)
    assert r["mean_square"] == pytest.approx(0.8 ** 2 / 4, rel=1e-14)
    assert r["s"] == pytest.approx(1 / (1 - 0.16), rel=1e-14)
    assert r["inv_k1"] == pytest.approx(1 / 5, rel=1e-13)


def test_two_modes_weighted_by_power():
    X, Y = grid(128)
    c = np.cos(3 * X) * np.cos(4 * Y) + 0.5 * np.cos(6 * X) * np.sin(8 * Y)
    # powers 1 : 0.25 at |k| = 5 and 10
    assert analysis.inverse_k1(c) == pytest.approx((0.2 + 0.25 * 0.1) / 1.25, rel=1e-13)


def test_axis_modes_are_excluded_like_the_reference():
    """The reference keeps only kx != 0 and ky != 0 ([1:n, 1:n]): a mode along
    one axis adds nothing."""
    X, Y = grid(64)
    mixed = np.cos(2 * X) * np.cos(2 * Y)
    assert analysis.inverse_k1(mixed + 0.7 * np.cos(5 * X)) == pytest.approx(
        analysis.inverse_k1(mixed), rel=1e-12)


def test_simpson_close_to_periodic_mean():
    X, Y = grid(256)
    c = 0.5 * np.sin(X) * np.cos(2 * Y) + 0.1
    a = analysis.mean_square(c, "periodic")
    b = analysis.mean_square(c, "simpson")
    assert abs(a - b) < 2e-2 * a


def test_series_in_time_order(tmp_path):
    X, Y = grid(32)
    for t in (2.5, 0.1, 10.0):
        np.save(tmp_path / analysis.snapshot_name(t), np.cos(X) * t)
    series = analysis.load_series(str(tmp_path))
    assert [t for t, _ in series] == [0.1, 2.5, 10.0]
    assert np.array_equal(series[2][1], np.cos(X) * 10.0)
    assert os.path.basename(analysis.snapshot_name(0.1)) == "cahn_hilliard_0.1000000000.npy"
