"""Pins the CPU oracle (oracle/custen_oracle.c) before it is trusted as the
parity reference.  Each check is independent of the oracle's own code:
exact-rational brute force on tiny grids, closed forms (impulses, constant
fields, circulant inverses by DFT), invariants (translation equivariance,
zero-sum weights, mass conservation) and the truncation-level known answers of
the reference's example programs (examples/src/*.cu), whose `answer` arrays
are the only expected values the reference ships."""
import numpy as np
import pytest

import oracle
from custen_amd import synth
from tests.helpers import bits_equal, brute_stencil, fma32

SHAPES = [  # (H, L, V, T): symmetric, asymmetric, x-only, y-only
    (5, 2, 5, 2),
    (3, 1, 3, 1),
    (4, 1, 2, 0),
    (2, 0, 3, 2),
    (5, 2, 1, 0),
    (4, 3, 1, 0),
    (1, 0, 5, 2),
    (1, 0, 4, 1),
]


@pytest.mark.parametrize("H,L,V,T", SHAPES)
@pytest.mark.parametrize("periodic", [True, False])
def test_brute_force_f64(H, L, V, T, periodic):
    u = synth.field(9, 7, seed=11)
    w = synth.uniform(H * V, seed=12, a=-3.0, b=3.0)
    sentinel = 1234.5
    got = oracle.apply(u, w, H, L, V, T, periodic, out=np.full_like(u, sentinel))
    ref, mask = brute_stencil(u, w, H, L, V, T, periodic)
    assert bits_equal(got[mask], ref[mask])
    assert np.all(got[~mask] == sentinel)


@pytest.mark.parametrize("H,L,V,T", SHAPES[:4])
@pytest.mark.parametrize("periodic", [True, False])
def test_brute_force_f32(H, L, V, T, periodic):
    u = synth.field(8, 6, seed=13).astype(np.float32)
    w = synth.uniform(H * V, seed=14, a=-3.0, b=3.0).astype(np.float32)
    got = oracle.apply(u, w, H, L, V, T, periodic, out=np.full_like(u, 7.0))
    ref, mask = brute_stencil(u, w, H, L, V, T, periodic, fma=fma32)
    assert bits_equal(got[mask], ref[mask])
    assert np.all(got[~mask] == 7.0)


def test_order_matters():
    """The emulated fma chain detects a wrong summation order: evaluating the
    same terms i-outer gives different bits on this input."""
    u = synth.field(9, 7, seed=21)
    w = synth.uniform(25, seed=22, a=-1e3, b=1e3)
    got = oracle.apply(u, w, 5, 2, 5, 2, True)
    wt = w.reshape(5, 5).T.copy().ravel()  # transposed weights = wrong operand
    bad = oracle.apply(u.T.copy(), wt, 5, 2, 5, 2, True).T
    assert not bits_equal(got, bad)


@pytest.mark.parametrize("H,L,V,T", SHAPES)
def test_impulse_shift(H, L, V, T):
    """A unit weight at (j, i) gives an exact periodic shift by (T-j, L-i)."""
    u = synth.field(11, 10, seed=3)
    for j in range(V):
        for i in range(H):
            w = np.zeros(H * V)
            w[j * H + i] = 1.0
            got = oracle.apply(u, w, H, L, V, T, True)
            assert bits_equal(got, np.roll(np.roll(u, T - j, axis=0), L - i, axis=1))


def test_translation_equivariance():
    u = synth.field(64, 48, seed=5)
    w = synth.uniform(25, seed=6, a=-1, b=1)
    a = oracle.apply(np.roll(u, (3, -7), axis=(0, 1)), w, 5, 2, 5, 2, True)
    b = np.roll(oracle.apply(u, w, 5, 2, 5, 2, True), (3, -7), axis=(0, 1))
    assert bits_equal(a, b)


def test_threads_identical():
    u = synth.field(200, 130, seed=8)
    w = synth.uniform(25, seed=9, a=-1, b=1)
    assert bits_equal(oracle.apply(u, w, 5, 2, 5, 2, True, nthreads=1),
                      oracle.apply(u, w, 5, 2, 5, 2, True, nthreads=4))


def test_x_np_zero_right_edge():
    """x-np writes 0.0 on the right R columns (2d_x_np_kernel.cu:164-176) and
    leaves the left L columns untouched."""
    u = synth.field(16, 4, seed=2)
    w = synth.uniform(5, seed=3, a=-1, b=1)
    out = oracle.apply(u, w, 5, 1, 1, 0, False, zero_right_edge=True,
                       out=np.full_like(u, 9.0))
    assert np.all(out[:, :1] == 9.0)
    assert np.all(out[:, -3:] == 0.0)
    ref, mask = brute_stencil(u, w, 5, 1, 1, 0, False)
    assert bits_equal(out[mask], ref[mask])


def _d2_weights_8th(h):
    # 9-point 8th-order d2/dx2 (examples/src/2d_x_p.cu:99-114)
    c = np.array([-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])
    return c / h ** 2


def test_example_kat_x_p():
    """examples/src/2d_x_p.cu: sin(x) with the 9-point d2/dx2 gives -sin(x)."""
    n = 1024
    h = 2 * np.pi / n
    x = np.arange(n) * h
    u = np.tile(np.sin(x), (8, 1))
    out = oracle.apply(u, _d2_weights_8th(h), 9, 4, 1, 0, True)
    assert np.max(np.abs(out + u)) < 1e-9


def test_example_kat_y_np():
    """examples/src/2d_y_np.cu: sin(y), 9-point d2/dy2, interior only."""
    n = 512
    h = 2 * np.pi / n
    y = np.arange(n) * h
    u = np.tile(np.sin(y)[:, None], (1, 8))
    out = oracle.apply(u, _d2_weights_8th(h), 1, 0, 9, 4, False)
    assert np.max(np.abs(out[4:-4] + u[4:-4])) < 1e-8


def test_example_kat_xy_p():
    """examples/src/2d_xy_p.cu:82-120: sin(x)cos(y) with the 3x3 cross
    derivative [1,0,-1;0,0,0;-1,0,1]/(4 dx dy) gives -cos(x) sin(y)."""
    n = 256
    h = 2 * np.pi / n
    x = np.arange(n) * h
    u = np.sin(x)[None, :] * np.cos(x)[:, None]
    w = np.array([1, 0, -1, 0, 0, 0, -1, 0, 1]) / (4 * h * h)
    out = oracle.apply(u, w, 3, 1, 3, 1, True)
    ans = -np.cos(x)[None, :] * np.sin(x)[:, None]
    assert np.max(np.abs(out - ans)) < 1e-3  # second order, h = 2.45e-2


def test_fun_x_central_difference_kat():
    """examples/src/2d_x_np_fun.cu:49-54 with coe = 1/dx^2: -sin(x)."""
    n = 1024
    h = 2 * np.pi / n
    u = np.tile(np.sin(np.arange(n) * h), (4, 1))
    out = oracle.apply_fun(oracle.FUN_X_CENTRAL_DIFF, 0, u, np.array([1 / h ** 2]), 3, 1, 1, 0,
                           False)
    assert np.max(np.abs(out[:, 1:-1] + u[:, 1:-1])) < 1e-5


def test_fun_weighted_sums_equal_weight_stencils():
    """The xy weighted-sum point function equals the weight stencil, and the
    cube-minus function equals the weight stencil applied to u^3 - u (formed
    here by numpy, which does not fuse)."""
    u = synth.field(40, 30, seed=31)
    c = synth.uniform(9, seed=32, a=-1, b=1)
    for periodic in (True, False):
        a = oracle.apply_fun(oracle.FUN_XY_WSUM, 2, u, c, 3, 1, 3, 1, periodic)
        b = oracle.apply(u, c, 3, 1, 3, 1, periodic)
        assert bits_equal(a, b)
        q = u * u * u - u
        a = oracle.apply_fun(oracle.FUN_XY_CUBE_MINUS, 2, u, c, 3, 1, 3, 1, periodic)
        b = oracle.apply(q, c, 3, 1, 3, 1, periodic)
        assert bits_equal(a, b)
    c9 = synth.uniform(9, seed=33, a=-1, b=1)
    a = oracle.apply_fun(oracle.FUN_Y_WSUM, 1, u, c9, 1, 0, 9, 4, True)
    b = oracle.apply(u, c9, 1, 0, 9, 4, True)
    assert bits_equal(a, b)
    # x weighted sum over numCoe points starting L left of the centre (the
    # native caller's devArg1X function), centred and one-sided
    c5 = synth.uniform(5, seed=34, a=-1, b=1)
    for L in (2, 1, 4):
        for periodic in (True, False):
            a = oracle.apply_fun(oracle.FUN_X_WSUM, 0, u, c5, 5, L, 1, 0, periodic)
            b = oracle.apply(u, c5, 5, L, 1, 0, periodic)
            assert bits_equal(a, b), (L, periodic)


# ---------------------------------------------------------------------------
# Cahn-Hilliard pieces
# ---------------------------------------------------------------------------

def test_pent_solve_matches_dense():
    m = 40
    rng = np.random.default_rng(0)
    diags = [rng.uniform(-1, 1, m) for _ in range(5)]
    diags[2] = diags[2] + 8.0  # diagonally dominant
    A = np.zeros((m, m))
    for k, off in enumerate((-2, -1, 0, 1, 2)):
        for i in range(m):
            j = i + off
            if 0 <= j < m:
                A[i, j] = diags[k][i]
    b = rng.uniform(-1, 1, m)
    x = oracle.pent_factor_solve(*diags, b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [16, 64, 256])
def test_cyclic_solve_matches_dft(n):
    """The periodic pentadiagonal matrix [s, -4s, 1+6s, -4s, s] is circulant with
    eigenvalues 1 + s (2 - 2 cos(2 pi k / n))^2: the solve is an exact DFT
    division (the closed form fixes every coefficient and the wrap-around)."""
    ch = oracle.CH(n)
    s = ch.st.sigma_lin
    r = synth.uniform(n, seed=40 + n, a=-1, b=1)
    x = r.copy()
    ch.cyclic_solve(x)
    th = 2 * np.pi * np.arange(n) / n
    lam = 1 + s * (2 - 2 * np.cos(th)) ** 2
    ref = np.real(np.fft.ifft(np.fft.fft(r) / lam))
    kappa = 1 + 16 * s
    assert np.max(np.abs(x - ref)) <= 1e-14 * kappa * np.max(np.abs(r)) * n


def test_ch_weights_and_parameters():
    n = 256
    ch = oracle.CH(n)
    st = ch.st
    dx = 2 * np.pi / n
    dt = 0.1 * dx
    assert st.dx == dx and st.dt == dt
    assert np.isclose(st.sigma_lin, 2 * dt * 0.01 / (3 * dx ** 4), rtol=1e-15)
    assert np.isclose(st.sigma_nl, dt / 3 * 2 / dx ** 2, rtol=1e-15)
    wl = np.array(st.wlin[:])
    wn = np.array(st.wnl[:])
    assert abs(wl.sum()) <= 1e-12 * np.abs(wl).sum() and abs(wn.sum()) <= 1e-12 * np.abs(wn).sum()
    # the 5x5 weights are -sigma * (discrete biharmonic * h^4)
    pat = np.array([[0, 0, 1, 0, 0], [0, 2, -8, 2, 0], [1, -8, 20, -8, 1],
                    [0, 2, -8, 2, 0], [0, 0, 1, 0, 0]], dtype=float)
    assert np.allclose(wl.reshape(5, 5), -st.sigma_lin * pat, rtol=1e-15)


def test_ch_omega_schur_complement():
    """omega is the inverse of the 2x2 Schur complement of the cyclic matrix
    with respect to its leading (n-2) block (BatchHyper.cu:421-514)."""
    n = 32
    ch = oracle.CH(n)
    st = ch.st
    a, b, c, d, e = st.a, st.b, st.c, st.d, st.e
    A = np.zeros((n, n))
    for i in range(n):
        for off, v in zip((-2, -1, 0, 1, 2), (a, b, c, d, e)):
            A[i, (i + off) % n] = v
    E, F = A[:-2, :-2], A[:-2, -2:]
    G, Hm = A[-2:, :-2], A[-2:, -2:]
    schur = Hm - G @ np.linalg.solve(E, F)
    om = np.array(st.omega[:]).reshape(2, 2)
    assert np.allclose(om, np.linalg.inv(schur), rtol=1e-9)


def test_ch_mass_conservation():
    n = 64
    ch = oracle.CH(n)
    c = synth.uniform(n * n, seed=5000, a=-0.1, b=0.1)
    cold = c.copy()
    m0 = c.sum()
    ch.step(c, cold, steps=20)
    assert abs(c.sum() - m0) <= 1e-12 * n * n
    assert np.all(np.isfinite(c))


def test_ch_step_linear_limit():
    """For a field small enough that u^3 is below rounding, one step from
    cOld = c is the linear implicit update: c1 = A^-1 (c + L c) in each
    direction, computed here with dense numpy algebra."""
    n = 16
    ch = oracle.CH(n)
    st = ch.st
    c0 = synth.uniform(n * n, seed=77, a=-1e-9, b=1e-9)
    c = c0.copy()
    cold = c0.copy()
    ch.step(c, cold)
    U = c0.reshape(n, n)
    wl = np.array(st.wlin[:]).reshape(5, 5)
    wn = np.array(st.wnl[:]).reshape(3, 3)
    lin = sum(wl[j, i] * np.roll(np.roll(U, 2 - j, 0), 2 - i, 1) for j in range(5) for i in range(5))
    nl = sum(wn[j, i] * np.roll(np.roll(-U, 1 - j, 0), 1 - i, 1) for j in range(3) for i in range(3))
    rhs = lin + nl
    A = np.zeros((n, n))
    for i in range(n):
        for off, v in zip((-2, -1, 0, 1, 2), (st.a, st.b, st.c, st.d, st.e)):
            A[i, (i + off) % n] = v
    X = np.linalg.solve(A, rhs.T).T  # x-direction solves (rows)
    X = np.linalg.solve(A, X)        # y-direction solves (columns)
    ref = U + X
    assert np.max(np.abs(c.reshape(n, n) - ref)) <= 1e-6 * np.max(np.abs(U))


def test_splitmix_implementations_agree():
    from custen_amd import api
    a = oracle.fill_uniform(4096, 123, -0.1, 0.1)
    b = synth.uniform(4096, 123, -0.1, 0.1)
    assert bits_equal(a, b)
    assert np.all((a >= -0.1) & (a < 0.1))
    try:
        L = api.lib()
    except ImportError:
        pytest.skip("native library not built")
    c = np.empty(4096)
    api.custenFillUniform(c, 4096, 123, -0.1, 0.1)
    assert bits_equal(a, c)


def _pent_systems(m, batch, seed):
    """Interleaved (m, batch) diagonals of diagonally dominant systems."""
    ds, dl, du, dw = (synth.field(batch, m, seed=seed + k) for k in range(4))
    d = synth.field(batch, m, seed=seed + 4, a=6.0, b=9.0)
    ds[:2] = 0.0
    dl[:1] = 0.0
    du[-1:] = 0.0
    dw[-2:] = 0.0
    b = synth.field(batch, m, seed=seed + 5)
    return ds, dl, d, du, dw, b


@pytest.mark.parametrize("m,batch", [(4, 1), (5, 3), (12, 7), (33, 2)])
def test_pent_batch_matches_dense_and_single(m, batch):
    """Batched interleaved factor + solve: each system's solution matches a dense
    LU solve to rounding level and the single-system oracle bit for bit."""
    ds, dl, d, du, dw, b = _pent_systems(m, batch, 300 + m)
    fac = oracle.pent_factor_batch(ds, dl, d, du, dw)
    x = oracle.pent_solve_batch(*fac, b)
    for k in range(batch):
        A = (np.diag(d[:, k]) + np.diag(dl[1:, k], -1) + np.diag(ds[2:, k], -2) +
             np.diag(du[:-1, k], 1) + np.diag(dw[:-2, k], 2))
        ref = np.linalg.solve(A, b[:, k])
        assert np.abs(x[:, k] - ref).max() < 1e-13 * np.abs(ref).max()
        single = oracle.pent_factor_solve(ds[:, k], dl[:, k], d[:, k], du[:, k], dw[:, k],
                                          b[:, k])
        assert bits_equal(x[:, k], single)
