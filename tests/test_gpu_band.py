"""GPU parity of the band walker (custenSetBandConfig) against the oracle.

The walker covers the interior 128-column strips of FP64 weight stencils; edge
strips, other dtypes and shapes run the strip kernel.  Every case here is
compared bit for bit with the oracle, with ragged widths and heights so that
bands end inside a chunk, np write masks fall inside walker strips, and
chunks of stores that are not one-store-per-row occur next to full ones."""
import zlib

import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import bits_equal, mismatches
from tests.test_gpu_stencil import SENTINEL, create_weights, dev, oracle_weights

pytestmark = pytest.mark.gpu

SHAPES = [
    # variant, H, L, V, T
    ("2DXp", 3, 1, 1, 0), ("2DXp", 5, 2, 1, 0), ("2DXnp", 7, 3, 1, 0), ("2DXp", 9, 4, 1, 0),
    ("2DXnp", 11, 5, 1, 0), ("2DXp", 13, 6, 1, 0),
    ("2DYp", 1, 0, 3, 1), ("2DYnp", 1, 0, 5, 2), ("2DYp", 1, 0, 7, 3), ("2DYp", 1, 0, 9, 4),
    ("2DYp", 1, 0, 11, 5), ("2DYnp", 1, 0, 13, 6),
    ("2DXYp", 3, 1, 3, 1), ("2DXYnp", 3, 1, 3, 1), ("2DXYp", 5, 2, 5, 2), ("2DXYnp", 5, 2, 5, 2),
    ("2DXYp", 5, 0, 5, 4), ("2DXYp", 3, 1, 5, 2), ("2DXYp", 5, 2, 3, 1), ("2DXYp", 7, 3, 7, 3),
    ("2DXYnp", 9, 4, 9, 4),
    # one-sided windows: all of the halo on one side of the strip or band
    ("2DXYp", 5, 4, 5, 0), ("2DXYnp", 3, 2, 5, 0), ("2DXp", 7, 6, 1, 0), ("2DXnp", 9, 0, 1, 0),
    ("2DYp", 1, 0, 7, 0), ("2DYnp", 1, 0, 9, 8),
    # 1 x 1: a scaled copy through the walker
    ("2DXYp", 1, 0, 1, 0), ("2DXYnp", 1, 0, 1, 0),
]


def _run(torch, variant, nx, ny, H, L, V, T, rpw, chunks, seed_extra=0):
    seed = zlib.crc32(repr((variant, nx, ny, H, L, V, T, seed_extra)).encode()) & 0xFFFF
    u = synth.field(nx, ny, seed)
    w = synth.uniform(H * V, seed + 1, -2.0, 2.0)
    init = np.full_like(u, SENTINEL)
    du, dout, dw = dev(torch, u), dev(torch, init), dev(torch, w)
    h = api.cuSten_t()
    create_weights(variant, h, 1, nx, ny, dout, du, dw, H, L, V, T)
    api.custenSetBandConfig(h, rpw, chunks)
    assert api.custenGetBandMode(h) == 1
    getattr(api, f"cuStenCompute{variant}")(h, api.DEVICE)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)
    return dout.cpu().numpy(), oracle_weights(variant, u, w, H, L, V, T, init)


@pytest.mark.parametrize("variant,H,L,V,T", SHAPES)
@pytest.mark.parametrize("rpw,chunks", [(4, 1), (4, 3), (8, 2)])
@pytest.mark.parametrize("nx,ny", [(3 * 128 + 44, 101), (3 * 128, 96)])
def test_band_shapes(gpu, variant, H, L, V, T, rpw, chunks, nx, ny):
    # 428 x 101: 3 interior strips + a ragged one, 101 rows end inside a
    # chunk (masked stores); 384 x 96: whole strips and chunks, which for the
    # periodic variants is the unmasked (every point written) instantiation
    got, ref = _run(gpu, variant, nx, ny, H, L, V, T, rpw, chunks)
    assert mismatches(got, ref) == 0, f"{mismatches(got, ref)} points differ"


@pytest.mark.parametrize("nx,ny", [(128, 64), (256, 17), (1024, 256), (640, 333)])
@pytest.mark.parametrize("variant", ["2DXYp", "2DXYnp"])
def test_band_sizes(gpu, nx, ny, variant):
    """Grid sizes where a band is one chunk, where every strip is interior,
    and where the walker's one-wrap bound (rows >= V + rows per wave) decides
    between the walker and the strip kernel (ny = 17 with 8 rows per wave)."""
    for rpw, chunks in ((4, 2), (8, 1), (4, 16)):
        got, ref = _run(gpu, variant, nx, ny, 5, 2, 5, 2, rpw, chunks)
        assert mismatches(got, ref) == 0, (rpw, chunks)


@pytest.mark.parametrize("variant", ["2DXYp", "2DXYnp", "2DYp", "2DXnp"])
@pytest.mark.parametrize("nt,nx,ny", [(1, 520, 150), (3, 520, 150), (3, 512, 192)])
def test_band_host_streaming(gpu, variant, nt, nx, ny):
    """Host-resident data streams through halo-extended staging tiles (rows
    addressed linearly and clamped); the walker runs on each tile (512 x 192
    in 3 tiles: whole chunks, the unmasked instantiation when periodic)."""
    torch = gpu
    H, L, V, T = (5, 2, 5, 2)
    if variant.startswith("2DX") and not variant.startswith("2DXY"):
        V, T = 1, 0
    if variant.startswith("2DY"):
        H, L = 1, 0
    u = synth.field(nx, ny, 301)
    w = synth.uniform(H * V, 302, -1.0, 1.0)
    init = np.full_like(u, SENTINEL)
    hin = torch.from_numpy(u).pin_memory()
    hout = torch.from_numpy(init).pin_memory()
    h = api.cuSten_t()
    create_weights(variant, h, nt, nx, ny, hout, hin, w, H, L, V, T)
    api.custenSetBandConfig(h, 4, 2)
    getattr(api, f"cuStenCompute{variant}")(h, api.HOST)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)
    ref = oracle_weights(variant, u, w, H, L, V, T, init)
    assert mismatches(hout.numpy(), ref) == 0


@pytest.mark.parametrize("variant,H,L,V,T", [("2DXYnp", 7, 3, 7, 3), ("2DXYnp", 7, 1, 7, 5),
                                              ("2DXYnp", 9, 4, 9, 4), ("2DXYnp", 9, 8, 9, 0)])
@pytest.mark.parametrize("nx,ny", [(428, 101), (640, 333), (1000, 77), (256, 40)])
@pytest.mark.parametrize("rpw,chunks", [(4, 2), (8, 1), (4, 5)])
def test_band_window_split(gpu, variant, H, L, V, T, nx, ny, rpw, chunks):
    """Non-periodic stencils of more than 36 points: the written window runs
    unmasked, the edge strips (one launch, by list) and the ragged rows that
    are left run masked, the rows outside the window are not walked.  Sizes
    give ragged rows for every rows-per-wave (or none), one or two edge
    strips, and a right edge inside the ragged column strip."""
    got, ref = _run(gpu, variant, nx, ny, H, L, V, T, rpw, chunks)
    assert mismatches(got, ref) == 0, f"{mismatches(got, ref)} points differ"


@pytest.mark.parametrize("nt,nx,ny", [(3, 520, 150), (4, 640, 204), (5, 384, 105)])
def test_band_window_split_host_streaming(gpu, nt, nx, ny):
    """The window split on host-streamed tiles: each tile is a region whose
    rows start at a global offset, so the row window starts inside the first
    tile only and ends inside the last."""
    torch = gpu
    H, L, V, T = 7, 3, 7, 3
    u = synth.field(nx, ny, 305)
    w = synth.uniform(H * V, 306, -1.0, 1.0)
    init = np.full_like(u, SENTINEL)
    hin = torch.from_numpy(u).pin_memory()
    hout = torch.from_numpy(init).pin_memory()
    h = api.cuSten_t()
    create_weights("2DXYnp", h, nt, nx, ny, hout, hin, w, H, L, V, T)
    api.custenSetBandConfig(h, 4, 2)
    api.cuStenCompute2DXYnp(h, api.HOST)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYnp(h)
    ref = oracle_weights("2DXYnp", u, w, H, L, V, T, init)
    assert mismatches(hout.numpy(), ref) == 0


def test_band_repeated_and_swapped(gpu):
    """Several Computes with Swap in between: the ring and counters start
    clean for every wave of every launch."""
    torch = gpu
    nx, ny = 768, 200
    u = synth.field(nx, ny, 303)
    w = synth.uniform(25, 304, -0.3, 0.3)
    a, b, dw = dev(torch, u), dev(torch, np.zeros_like(u)), dev(torch, w)
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, 1, nx, ny, 32, 32, b, a, dw, 5, 2, 2, 5, 2, 2)
    api.custenSetBandConfig(h, 8, 3)
    ref = u
    for step in range(4):
        api.cuStenCompute2DXYp(h, api.DEVICE)
        api.cuStenSwap2DXYp(h, b if step % 2 == 0 else a)
        ref = oracle.apply(ref, w, 5, 2, 5, 2, True)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYp(h)
    assert bits_equal(a.cpu().numpy(), ref)


def test_band_weighted_sum_function(gpu):
    """The built-in weighted-sum point function runs on the stencil kernels,
    so it takes the walker too."""
    torch = gpu
    nx, ny = 400, 130
    u = synth.field(nx, ny, 305)
    c = synth.uniform(25, 306, -1.0, 1.0)
    init = np.full_like(u, SENTINEL)
    du, dout, dc = dev(torch, u), dev(torch, init), dev(torch, c)
    fn = api.custenGetBuiltinFunction("xy_weighted_sum")
    h = api.cuSten_t()
    api.cuStenCreate2DXYpFun(h, 0, 1, nx, ny, 32, 8, dout, du, dc, 5, 2, 2, 5, 2, 2, fn)
    api.custenSetBandConfig(h, 4, 4)
    api.cuStenCompute2DXYpFun(h, api.DEVICE)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYpFun(h)
    ref = oracle.apply_fun(oracle.FUN_XY_WSUM, 2, u, c, 5, 2, 5, 2, True, out=init.copy())
    assert mismatches(dout.cpu().numpy(), ref) == 0


def test_band_config_switch(gpu):
    """custenSetTileConfig switches the walker off; both give the same bits."""
    torch = gpu
    nx, ny = 1024, 512
    u = synth.field(nx, ny, 307)
    w = synth.uniform(25, 308, -1.0, 1.0)
    du, dw = dev(torch, u), dev(torch, w)
    outs = []
    for band in (True, False):
        dout = dev(torch, np.zeros_like(u))
        h = api.cuSten_t()
        api.cuStenCreate2DXYp(h, 0, 1, nx, ny, 32, 32, dout, du, dw, 5, 2, 2, 5, 2, 2)
        api.custenSetBandConfig(h, 4, 4)
        if not band:
            api.custenSetTileConfig(h, 4, 2)
        assert api.custenGetBandMode(h) == (1 if band else 0)
        api.cuStenCompute2DXYp(h, api.DEVICE)
        torch.cuda.synchronize()
        api.cuStenDestroy2DXYp(h)
        outs.append(dout.cpu().numpy())
    assert bits_equal(outs[0], outs[1])
    assert bits_equal(outs[0], oracle.apply(u, w, 5, 2, 5, 2, True))


F32_SHAPES = [
    # dir, H, L, V, T
    (2, 5, 2, 5, 2), (2, 3, 1, 3, 1), (2, 7, 3, 7, 3), (0, 5, 2, 1, 0), (0, 11, 5, 1, 0),
    (1, 1, 0, 5, 2), (1, 1, 0, 13, 6), (2, 3, 1, 5, 2),
]


@pytest.mark.parametrize("dir_,H,L,V,T", F32_SHAPES)
@pytest.mark.parametrize("periodic", [1, 0])
@pytest.mark.parametrize("nx,ny", [(3 * 128 + 44, 101), (3 * 128, 96)])
def test_band_f32(gpu, dir_, H, L, V, T, periodic, nx, ny):
    """FP32 storage on the walker (each 128-column row body is two 4-byte
    loads per lane): bit for bit against the FP32 oracle."""
    torch = gpu
    u = synth.field(nx, ny, 311 + H * 7 + V).astype(np.float32)
    w64 = synth.uniform(H * V, 312, -1, 1)
    init = np.full_like(u, np.float32(SENTINEL))
    du, dout, dw = dev(torch, u), dev(torch, init), dev(torch, w64)
    for rpw, chunks in ((4, 2), (8, 3)):
        dout.copy_(dev(torch, init))
        h = api.cuSten_t()
        api.custenCreate2DF32(h, 0, 1, nx, ny, dir_, periodic, dout, du, dw, H, L, H - 1 - L, V,
                              T, V - 1 - T)
        api.custenSetBandConfig(h, rpw, chunks)
        api.custenCompute2DF32(h, api.DEVICE)
        torch.cuda.synchronize()
        api.custenDestroy2DF32(h)
        ref = oracle.apply(u, w64.astype(np.float32), H, L, V, T, bool(periodic),
                           zero_right_edge=(dir_ == 0 and not periodic), out=init.copy())
        assert mismatches(dout.cpu().numpy(), ref) == 0, (rpw, chunks)


@pytest.mark.parametrize("variant", ["2DXYpFun", "2DXYnpFun"])
@pytest.mark.parametrize("H,L,V,T", [(3, 1, 3, 1), (5, 2, 5, 2), (3, 0, 3, 2)])
@pytest.mark.parametrize("nx,ny", [(3 * 128 + 44, 101), (3 * 128, 96)])
def test_band_cube_minus_function(gpu, variant, H, L, V, T, nx, ny):
    """The cube-minus point function (u^3 - u under a weighted sum, as the
    Cahn-Hilliard non-linear term) on the walker: the transform is applied as
    values are read from the ring, and the result equals the oracle's
    restatement of the function bit for bit."""
    torch = gpu
    u = synth.field(nx, ny, 321 + H)
    c = synth.uniform(H * V, 322, -1.0, 1.0)
    init = np.full_like(u, SENTINEL)
    fn = api.custenGetBuiltinFunction("xy_cube_minus")
    for rpw, chunks in ((4, 2), (8, 3)):
        du, dout, dc = dev(torch, u), dev(torch, init), dev(torch, c)
        h = api.cuSten_t()
        getattr(api, f"cuStenCreate{variant}")(h, 0, 1, nx, ny, 32, 8, dout, du, dc, H, L,
                                               H - 1 - L, V, T, V - 1 - T, fn)
        api.custenSetBandConfig(h, rpw, chunks)
        getattr(api, f"cuStenCompute{variant}")(h, api.DEVICE)
        torch.cuda.synchronize()
        getattr(api, f"cuStenDestroy{variant}")(h)
        ref = oracle.apply_fun(oracle.FUN_XY_CUBE_MINUS, 2, u, c, H, L, V, T,
                               variant == "2DXYpFun", out=init.copy())
        assert mismatches(dout.cpu().numpy(), ref) == 0, (rpw, chunks)


def _random_cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        variant, H, L, V, T = SHAPES[int(rng.integers(len(SHAPES)))]
        # mostly even widths (pair-aligned rows: the walker runs), some odd
        # (unaligned rows: every strip falls back to the strip kernel)
        nx = int(rng.integers(64, 750)) * 2 if rng.random() < 0.75 else int(rng.integers(128, 1500))
        ny = int(rng.integers(V + 8, 400))
        rpw = int(rng.choice([4, 8]))
        chunks = int(rng.integers(1, 6))
        out.append((variant, nx, ny, H, L, V, T, rpw, chunks))
    return out


@pytest.mark.parametrize("variant,nx,ny,H,L,V,T,rpw,chunks", _random_cases(24, 2024))
def test_band_random_geometry(gpu, variant, nx, ny, H, L, V, T, rpw, chunks):
    """Seeded random grid sizes, shapes and band configurations (widths that
    are and are not multiples of 128, heights that do and do not fill whole
    chunks and bands), bit for bit against the oracle."""
    got, ref = _run(gpu, variant, nx, ny, H, L, V, T, rpw, chunks, seed_extra=nx * ny)
    assert mismatches(got, ref) == 0, f"{mismatches(got, ref)} points differ"
