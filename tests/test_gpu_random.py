"""Seeded random parity sweeps over the surfaces the fixed-case tests cover
one geometry at a time: FP32 storage, the one-column-per-lane strips,
host-resident streaming, the point functions and WENO5.  Sizes are drawn so
that widths are odd and even, below and above 128-column multiples, and
heights fill and do not fill whole chunks; every result is compared bit for
bit with the oracle."""
import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import bits_equal, mismatches
from tests.test_gpu_stencil import SENTINEL, create_weights, dev, oracle_weights

pytestmark = pytest.mark.gpu

VARIANTS = ["2DXp", "2DXnp", "2DYp", "2DYnp", "2DXYp", "2DXYnp"]


def _shape(rng, variant):
    if variant.startswith("2DXY"):
        H = int(rng.choice([3, 5, 7]))
        V = int(rng.choice([3, 5]))
        return H, int(rng.integers(0, H)), V, int(rng.integers(0, V))
    if variant.startswith("2DX"):
        H = int(rng.choice([3, 5, 9, 11]))
        return H, int(rng.integers(0, H)), 1, 0
    V = int(rng.choice([3, 5, 9, 13]))
    return 1, 0, V, int(rng.integers(0, V))


def _cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        variant = VARIANTS[int(rng.integers(len(VARIANTS)))]
        H, L, V, T = _shape(rng, variant)
        nx = int(rng.integers(16, 700))
        ny = int(rng.integers(max(V, 4), 260))
        out.append((variant, nx, ny, H, L, V, T, int(rng.integers(1 << 30))))
    return out


@pytest.mark.parametrize("variant,nx,ny,H,L,V,T,seed", _cases(16, 11))
def test_random_f32(gpu, variant, nx, ny, H, L, V, T, seed):
    torch = gpu
    dir_ = 2 if variant.startswith("2DXY") else (0 if variant.startswith("2DX") else 1)
    periodic = 0 if variant.endswith("np") else 1
    u = synth.field(nx, ny, seed % 65536).astype(np.float32)
    w64 = synth.uniform(H * V, seed % 65536 + 1, -1.0, 1.0)
    init = np.full_like(u, np.float32(SENTINEL))
    du, dout, dw = dev(torch, u), dev(torch, init), dev(torch, w64)
    h = api.cuSten_t()
    api.custenCreate2DF32(h, 0, 1, nx, ny, dir_, periodic, dout, du, dw, H, L, H - 1 - L, V, T,
                          V - 1 - T)
    if seed % 2:
        api.custenSetBandConfig(h, 4 if seed % 4 == 1 else 8, 1 + seed % 3)
    api.custenCompute2DF32(h, api.DEVICE)
    torch.cuda.synchronize()
    api.custenDestroy2DF32(h)
    ref = oracle.apply(u, w64.astype(np.float32), H, L, V, T, bool(periodic),
                       zero_right_edge=(dir_ == 0 and not periodic), out=init.copy())
    assert mismatches(dout.cpu().numpy(), ref) == 0


@pytest.mark.parametrize("variant,nx,ny,H,L,V,T,seed", _cases(12, 12))
def test_random_one_column_per_lane(gpu, variant, nx, ny, H, L, V, T, seed):
    """custenSetTileConfigEx(..., columnsPerLane = 1): 64-column strips."""
    torch = gpu
    u = synth.field(nx, ny, seed % 65536)
    w = synth.uniform(H * V, seed % 65536 + 1, -2.0, 2.0)
    init = np.full_like(u, SENTINEL)
    du, dout, dw = dev(torch, u), dev(torch, init), dev(torch, w)
    h = api.cuSten_t()
    create_weights(variant, h, 1, nx, ny, dout, du, dw, H, L, V, T)
    api.custenSetTileConfig(h, 8 if seed % 2 else 16, 1 + seed % 4, 1)
    getattr(api, f"cuStenCompute{variant}")(h, api.DEVICE)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)
    ref = oracle_weights(variant, u, w, H, L, V, T, init)
    assert mismatches(dout.cpu().numpy(), ref) == 0


@pytest.mark.parametrize("variant,nx,ny,H,L,V,T,seed", _cases(12, 13))
def test_random_host_streaming(gpu, variant, nx, ny, H, L, V, T, seed):
    """Host-resident data in numTiles tiles (a divisor of ny), pinned or
    pageable memory, with or without the band walker on the tiles."""
    torch = gpu
    divisors = [d for d in (1, 2, 3, 4, 5, 6, 8) if ny % d == 0 and ny // d >= max(V, 1)]
    nt = divisors[seed % len(divisors)]
    u = synth.field(nx, ny, seed % 65536)
    w = synth.uniform(H * V, seed % 65536 + 1, -1.0, 1.0)
    init = np.full_like(u, SENTINEL)
    if seed % 2:
        hin, hout = torch.from_numpy(u.copy()).pin_memory(), torch.from_numpy(init.copy()).pin_memory()
    else:
        hin, hout = torch.from_numpy(u.copy()), torch.from_numpy(init.copy())
    h = api.cuSten_t()
    create_weights(variant, h, nt, nx, ny, hout, hin, w, H, L, V, T)
    if seed % 3 == 0:
        api.custenSetBandConfig(h, 4, 2)
    getattr(api, f"cuStenCompute{variant}")(h, api.HOST)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)
    ref = oracle_weights(variant, u, w, H, L, V, T, init)
    assert mismatches(hout.numpy(), ref) == 0, (nt, mismatches(hout.numpy(), ref))


def _fun_cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        periodic = bool(rng.integers(2))
        H = int(rng.choice([3, 5]))
        nx = int(rng.integers(16, 700))
        ny = int(rng.integers(8, 200))
        out.append((periodic, H, nx, ny, int(rng.integers(1 << 30))))
    return out


@pytest.mark.parametrize("periodic,H,nx,ny,seed", _fun_cases(10, 14))
def test_random_cube_minus(gpu, periodic, H, nx, ny, seed):
    """The cube-minus point function (inlined), strip kernel or band walker."""
    torch = gpu
    L = T = H // 2
    variant = "2DXYpFun" if periodic else "2DXYnpFun"
    u = synth.field(nx, ny, seed % 65536)
    c = synth.uniform(H * H, seed % 65536 + 1, -1.0, 1.0)
    init = np.full_like(u, SENTINEL)
    du, dout, dc = dev(torch, u), dev(torch, init), dev(torch, c)
    fn = api.custenGetBuiltinFunction("xy_cube_minus")
    h = api.cuSten_t()
    getattr(api, f"cuStenCreate{variant}")(h, 0, 1, nx, ny, 32, 8, dout, du, dc, H, L, L, H, T, T,
                                           fn)
    if seed % 2:
        api.custenSetBandConfig(h, 8, 1 + seed % 3)
    getattr(api, f"cuStenCompute{variant}")(h, api.DEVICE)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)
    ref = oracle.apply_fun(oracle.FUN_XY_CUBE_MINUS, 2, u, c, H, L, H, T, periodic,
                           out=init.copy())
    assert mismatches(dout.cpu().numpy(), ref) == 0


@pytest.mark.parametrize("nx,ny,seed", [(int(a), int(b), int(c)) for a, b, c in
                                        np.random.default_rng(15).integers(
                                            [8, 8, 0], [600, 300, 1 << 30], size=(8, 3))])
def test_random_weno(gpu, nx, ny, seed):
    torch = gpu
    phi = synth.field(nx, ny, seed % 65536)
    u = synth.field(nx, ny, seed % 65536 + 1)
    v = synth.field(nx, ny, seed % 65536 + 2)
    dx, dy = 2 * np.pi / nx, 2 * np.pi / ny
    dphi, du, dv = dev(torch, phi), dev(torch, u), dev(torch, v)
    dout = torch.full_like(dphi, 7.0)
    h = api.cuSten_t()
    api.cuStenCreate2DXYWENOADVp(h, 0, 1, nx, ny, 32, 32, dx, dy, du, dv, dout, dphi)
    single = bool(seed % 2)
    if not single:
        api.custenSetWenoSquares(h, False)
    api.cuStenCompute2DXYWENOADVp(h, api.DEVICE)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYWENOADVp(h)
    ref = oracle.weno(phi, u, v, dx, dy, single, nthreads=8)
    got = dout.cpu().numpy()
    assert bits_equal(got, ref), mismatches(got, ref)


@pytest.mark.parametrize("n,steps,mode", [(int(a), int(b), int(c)) for a, b, c in
                                          np.random.default_rng(16).integers(
                                              [8, 1, 0], [300, 5, 2], size=(8, 3))])
def test_random_ch_sizes(gpu, n, steps, mode):
    """Cahn-Hilliard ADI at random grid sizes (not multiples of the 64-point
    tiles or blocks), fused and plain modes, against the oracle."""
    c0 = synth.uniform(n * n, 7000 + n, -0.1, 0.1).reshape(n, n)
    s = api.CHSolver(api.ch_params(n), c0)
    s.run(steps, mode=mode)
    c, co = s.field()
    s.close()
    ref = oracle.CH(n)
    rc, rco = c0.ravel().copy(), c0.ravel().copy()
    ref.step(rc, rco, steps=steps)
    assert mismatches(c, rc.reshape(n, n)) == 0
    assert mismatches(co, rco.reshape(n, n)) == 0


@pytest.mark.parametrize("sizes,seed", [([int(x) for x in np.random.default_rng(17 + k).integers(
    8, 400, size=int(np.random.default_rng(17 + k).integers(2, 9)))], 17 + k) for k in range(4)])
def test_random_batch(gpu, sizes, seed):
    """Batched grids of random uneven (non-square-multiple) sizes, fused with
    the strip kernel and with the band walker, against the oracle."""
    torch = gpu
    w = synth.uniform(25, seed, -0.05, 0.05)
    fields = [synth.field(n, n, seed * 100 + g) for g, n in enumerate(sizes)]
    for band in (False, True):
        ins = [torch.from_numpy(f.copy()).cuda() for f in fields]
        outs = [torch.zeros_like(t) for t in ins]
        b = api.Batch(outs, ins, torch.from_numpy(w).cuda())
        b.set_tile_config(4 if band else 8, 2, band)
        b.run(3)
        torch.cuda.synchronize()
        cur = outs if b.parity == 1 else ins
        for g, f in enumerate(fields):
            ref = f
            for _ in range(3):
                ref = oracle.apply(ref, w, 5, 2, 5, 2, True)
            assert bits_equal(cur[g].cpu().numpy(), ref), (band, g, sizes[g])
        b.close()
