"""Pins the WENO5 advection oracle (cso_weno_f64) independently of its code:
closed forms (constant fields, one-sided stencils that cannot see a step),
exact symmetries (mirror images under u -> -u, periodic translation), the
fifth-order convergence of the reference example's fields
(examples/src/2d_xyWENOADV_p.cu: phi = cos x sin y, u = sin y, v = -sin x, whose
advection term is known in closed form), and the size of the single-precision
squares' effect."""
import numpy as np
import pytest

import oracle
from custen_amd import synth
from tests.helpers import bits_equal


def example_fields(n):
    """The reference example's inputs on an n x n grid of [0, 2 pi)^2 and the
    exact advection term u phi_x + v phi_y."""
    h = 2 * np.pi / n
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)  # X varies along columns (x), Y along rows (y)
    phi = np.cos(X) * np.sin(Y)
    u = np.sin(Y)
    v = -np.sin(X)
    exact = u * (-np.sin(X) * np.sin(Y)) + v * (np.cos(X) * np.cos(Y))
    return phi, u, v, exact, h


def random_case(nx, ny, seed):
    phi = synth.field(nx, ny, seed=seed)
    u = synth.field(nx, ny, seed=seed + 1)
    v = synth.field(nx, ny, seed=seed + 2)
    return phi, u, v


@pytest.mark.parametrize("single", [True, False])
def test_constant_field_gives_zero(single):
    phi = np.full((12, 17), 3.25)
    _, u, v = random_case(17, 12, 40)
    out = oracle.weno(phi, u, v, 0.1, 0.2, single)
    assert np.all(out == 0.0)


@pytest.mark.parametrize("single", [True, False])
def test_convergence_fifth_order(single):
    errs = []
    for n in (32, 64, 128):
        phi, u, v, exact, h = example_fields(n)
        errs.append(np.abs(oracle.weno(phi, u, v, h, h, single) - exact).max())
    assert errs[0] < 1e-4
    for a, b in zip(errs, errs[1:]):
        assert a / b > 24.0  # 2^5 = 32 for a fifth-order scheme


def test_example_kat():
    """The reference example's configuration, checked at truncation level."""
    phi, u, v, exact, h = example_fields(256)
    out = oracle.weno(phi, u, v, h, h, True)
    assert np.abs(out - exact).max() < 3e-9


@pytest.mark.parametrize("single", [True, False])
def test_upwind_side_cannot_see_a_step(single):
    """A periodic step in x at column s (and back down at column 0): with u > 0
    the flux at column x uses the differences ending at x-2 .. x+2, so columns
    whose differences avoid both jumps give exactly 0, s-3 included; with
    u <= 0 it uses x-1 .. x+3, so s+2 gives 0.  Swapping the two sides would
    put the jump into both of those stencils."""
    nx, ny, s = 24, 5, 12
    phi = np.zeros((ny, nx))
    phi[:, s:] = 1.0  # jumps at s (up) and at 0 (down, periodic)
    zero_v = np.zeros((ny, nx))
    pos = oracle.weno(phi, np.ones((ny, nx)), zero_v, 0.5, 0.5, single)
    neg = oracle.weno(phi, -np.ones((ny, nx)), zero_v, 0.5, 0.5, single)
    # u > 0: columns s-3 .. 3 + ... (those not reaching s or the wrap at 0)
    for x in (s - 3, s - 4, s - 5, s - 6, s - 7, s - 8, s - 9):
        if x - 2 >= 1:  # differences x-2 .. x+2 avoid both jumps (at 0 and s)
            assert np.all(pos[:, x] == 0.0), x
    assert np.all(pos[:, s - 2] != 0.0)  # reaches difference s
    # u <= 0: columns s+2 .. where x-1 >= s+1 and x+3 <= nx-1
    for x in range(s + 2, nx - 3):
        assert np.all(neg[:, x] == 0.0), x
    assert np.all(neg[:, s + 1] != 0.0)
    # u = 0 takes the u <= 0 side and multiplies by zero
    zero_u = oracle.weno(phi, zero_v, zero_v, 0.5, 0.5, single)
    assert np.all(zero_u == 0.0)


@pytest.mark.parametrize("single", [True, False])
def test_step_in_y_upwinds_on_v(single):
    nx, ny, s = 5, 24, 12
    phi = np.zeros((ny, nx))
    phi[s:, :] = 1.0
    zero_u = np.zeros((ny, nx))
    pos = oracle.weno(phi, zero_u, np.ones((ny, nx)), 0.5, 0.5, single)
    neg = oracle.weno(phi, zero_u, -np.ones((ny, nx)), 0.5, 0.5, single)
    for y in range(3, s - 2):
        assert np.all(pos[y] == 0.0), y
    assert np.all(pos[s - 2] != 0.0)
    for y in range(s + 2, ny - 3):
        assert np.all(neg[y] == 0.0), y
    assert np.all(neg[s + 1] != 0.0)


@pytest.mark.parametrize("single", [True, False])
def test_mirror_symmetry_is_bitwise(single):
    """phi'(x) = phi(-x) with u' = -u(-x): every difference flips sign and the
    upwind side flips, so out'(x) = out(-x) exactly (negation commutes with
    rounding, squares ignore the sign)."""
    phi, u, v = random_case(23, 19, 50)
    u[u == 0.0] = 0.5  # the mirror of u = 0 (which takes the u <= 0 side) is not symmetric
    v[v == 0.0] = 0.5
    out = oracle.weno(phi, u, v, 0.3, 0.7, single)

    def mirror_x(a):
        return np.roll(a[:, ::-1], 1, axis=1)  # index x -> -x (mod nx)

    def mirror_y(a):
        return np.roll(a[::-1, :], 1, axis=0)

    mx = oracle.weno(mirror_x(phi), -mirror_x(u), mirror_x(v), 0.3, 0.7, single)
    assert bits_equal(mx, mirror_x(out))
    my = oracle.weno(mirror_y(phi), mirror_y(u), -mirror_y(v), 0.3, 0.7, single)
    assert bits_equal(my, mirror_y(out))


def test_translation_equivariance_is_bitwise():
    phi, u, v = random_case(31, 26, 60)
    out = oracle.weno(phi, u, v, 0.25, 0.5)
    sx, sy = 7, 11
    r = lambda a: np.roll(np.roll(a, sy, axis=0), sx, axis=1)  # noqa: E731
    assert bits_equal(oracle.weno(r(phi), r(u), r(v), 0.25, 0.5), r(out))


def test_single_squares_are_used_and_small():
    """Single-precision squares change the weights by ~2^-24 relative: the
    results differ from the double-square ones somewhere, by a tiny amount."""
    phi, u, v = random_case(40, 30, 70)
    a = oracle.weno(phi, u, v, 0.1, 0.1, True)
    b = oracle.weno(phi, u, v, 0.1, 0.1, False)
    assert not bits_equal(a, b)
    scale = np.abs(b).max()
    assert np.abs(a - b).max() < 1e-6 * scale


def test_threads_identical():
    phi, u, v = random_case(64, 48, 80)
    a = oracle.weno(phi, u, v, 0.1, 0.2, True, nthreads=1)
    b = oracle.weno(phi, u, v, 0.1, 0.2, True, nthreads=4)
    assert bits_equal(a, b)
