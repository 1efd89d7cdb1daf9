"""GPU parity of every cuSten variant against the CPU oracle, through the C ABI.

Bar: bit-exact equality with the oracle on every written point (the kernels
and the oracle evaluate the same fma chain in the same order), untouched
points keep their prior value, plus the scaled error metric
max|y - y_ref| / max(|y_ref|, sum|w||u|) <= 1e-14 reported for the record."""
import zlib

import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import bits_equal, mismatches, scaled_error

pytestmark = pytest.mark.gpu

SENTINEL = 12345.0


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def create_weights(variant, h, nt, nx, ny, out, inp, w, H, L, V, T, bx=32, by=8):
    R, B = H - 1 - L, V - 1 - T
    fn = getattr(api, f"cuStenCreate{variant}")
    if variant.startswith("2DXY"):
        fn(h, 0, nt, nx, ny, bx, by, out, inp, w, H, L, R, V, T, B)
    elif variant.startswith("2DX"):
        fn(h, 0, nt, nx, ny, bx, by, out, inp, w, H, L, R)
    else:
        fn(h, 0, nt, nx, ny, bx, by, out, inp, w, V, T, B)


def oracle_weights(variant, u, w, H, L, V, T, init):
    periodic = not variant.endswith("np")
    out = init.copy()
    if variant.startswith("2DXY"):
        return oracle.apply(u, w, H, L, V, T, periodic, out=out)
    if variant.startswith("2DX"):
        return oracle.apply(u, w, H, L, 1, 0, periodic, zero_right_edge=not periodic, out=out)
    return oracle.apply(u, w, 1, 0, V, T, periodic, out=out)


CASES = [
    # variant, nx, ny, numTiles, H, L, V, T
    ("2DXp", 128, 64, 1, 5, 2, 1, 0),
    ("2DXp", 1024, 1024, 1, 5, 2, 1, 0),
    ("2DXp", 37, 53, 1, 9, 4, 1, 0),
    ("2DXp", 300, 40, 4, 4, 1, 1, 0),
    ("2DXp", 260, 20, 2, 11, 5, 1, 0),
    ("2DXnp", 128, 64, 2, 5, 2, 1, 0),
    ("2DXnp", 333, 17, 1, 4, 3, 1, 0),
    ("2DXnp", 200, 50, 5, 13, 6, 1, 0),
    ("2DYp", 128, 64, 1, 1, 0, 5, 2),
    ("2DYp", 64, 64, 4, 1, 0, 9, 4),
    ("2DYp", 97, 70, 2, 1, 0, 4, 0),
    ("2DYp", 130, 66, 3, 1, 0, 11, 6),
    ("2DYnp", 128, 64, 2, 1, 0, 5, 2),
    ("2DYnp", 51, 90, 3, 1, 0, 3, 2),
    ("2DXYp", 128, 64, 1, 5, 2, 5, 2),
    ("2DXYp", 256, 256, 4, 3, 1, 3, 1),
    ("2DXYp", 129, 77, 1, 4, 1, 2, 0),
    ("2DXYp", 300, 100, 2, 7, 3, 7, 3),
    ("2DXYp", 256, 64, 1, 7, 1, 7, 5),
    ("2DXYp", 300, 100, 2, 9, 4, 9, 4),
    ("2DXYp", 200, 64, 1, 3, 1, 5, 2),
    ("2DXYp", 200, 64, 1, 5, 2, 3, 1),
    ("2DXYnp", 300, 100, 1, 7, 3, 7, 3),
    ("2DXYnp", 256, 96, 1, 9, 4, 9, 4),
    ("2DXp", 1024, 64, 1, 13, 6, 1, 0),
    ("2DXnp", 520, 40, 2, 11, 5, 1, 0),
    ("2DYp", 256, 128, 2, 1, 0, 13, 6),
    ("2DYnp", 200, 120, 1, 1, 0, 11, 5),
    ("2DXYp", 64, 32, 1, 5, 0, 5, 4),
    ("2DXYp", 5, 4, 1, 5, 2, 5, 2),
    ("2DXYnp", 128, 64, 2, 5, 2, 5, 2),
    ("2DXYnp", 257, 131, 1, 3, 1, 3, 1),
    ("2DXYnp", 150, 80, 4, 4, 1, 3, 2),
    # more than four strips and none of them interior (odd widths: rows not
    # pair-aligned; a periodic grid lower than the stencil)
    ("2DYnp", 973, 123, 1, 1, 0, 5, 2),
    ("2DXYp", 1001, 64, 1, 5, 2, 5, 2),
    ("2DXp", 999, 40, 1, 5, 2, 1, 0),
    ("2DXYnp", 777, 50, 1, 3, 1, 3, 1),
    ("2DXYp", 700, 3, 1, 5, 2, 5, 2),
    ("2DYp", 1100, 4, 2, 1, 0, 7, 3),
    # 1 x 1 (a scaled copy)
    ("2DXYp", 300, 100, 1, 1, 0, 1, 0),
    ("2DXYnp", 256, 64, 2, 1, 0, 1, 0),
    ("2DXp", 1024, 64, 1, 1, 0, 1, 0),
]


@pytest.mark.parametrize("variant,nx,ny,nt,H,L,V,T", CASES)
def test_weights_device(gpu, variant, nx, ny, nt, H, L, V, T):
    torch = gpu
    seed = zlib.crc32(repr((variant, nx, ny, H, L, V, T)).encode()) & 0xFFFF
    u = synth.field(nx, ny, seed)
    w = synth.uniform(H * V, seed + 1, -2.0, 2.0)
    init = np.full_like(u, SENTINEL)
    du, dout, dw = dev(torch, u), dev(torch, init), dev(torch, w)
    h = api.cuSten_t()
    create_weights(variant, h, nt, nx, ny, dout, du, dw, H, L, V, T)
    getattr(api, f"cuStenCompute{variant}")(h, api.DEVICE)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)
    got = dout.cpu().numpy()
    ref = oracle_weights(variant, u, w, H, L, V, T, init)
    assert mismatches(got, ref) == 0, f"{mismatches(got, ref)} points differ"
    hh, ll, vv, tt = (H, L, 1, 0) if variant.startswith("2DX") and not variant.startswith("2DXY") \
        else ((1, 0, V, T) if variant.startswith("2DY") else (H, L, V, T))
    assert scaled_error(got, ref, u, w, hh, ll, vv, tt, not variant.endswith("np")) <= 1e-14


@pytest.mark.parametrize("variant", ["2DXp", "2DXnp", "2DYp", "2DYnp", "2DXYp", "2DXYnp"])
@pytest.mark.parametrize("mode", ["pageable", "pinned"])
@pytest.mark.parametrize("nt", [1, 2, 4])
def test_weights_host_streaming(gpu, variant, mode, nt):
    """Host-resident data is streamed tile by tile; results equal the oracle
    (and hence the device-resident path) bit for bit, for every numTiles."""
    torch = gpu
    nx, ny = 200, 96
    H, L, V, T = (5, 2, 5, 2)
    if variant.startswith("2DX") and not variant.startswith("2DXY"):
        V, T = 1, 0
    if variant.startswith("2DY"):
        H, L = 1, 0
    u = synth.field(nx, ny, 99)
    w = synth.uniform(H * V, 98, -1.0, 1.0)
    init = np.full_like(u, SENTINEL)
    if mode == "pinned":
        hin = torch.from_numpy(u).pin_memory()
        hout = torch.from_numpy(init).pin_memory()
    else:
        hin = torch.from_numpy(u.copy())
        hout = torch.from_numpy(init.copy())
    h = api.cuSten_t()
    create_weights(variant, h, nt, nx, ny, hout, hin, w, H, L, V, T)
    getattr(api, f"cuStenCompute{variant}")(h, api.HOST)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)
    ref = oracle_weights(variant, u, w, H, L, V, T, init)
    got = hout.numpy()
    assert mismatches(got, ref) == 0


def test_swap_two_steps(gpu):
    """Compute, Swap, Compute equals two applies with a manual pointer exchange."""
    torch = gpu
    nx, ny, nt = 256, 128, 2
    u = synth.field(nx, ny, 7)
    w = synth.uniform(9, 8, -0.5, 0.5)
    a, b, dw = dev(torch, u), dev(torch, np.zeros_like(u)), dev(torch, w)
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, nt, nx, ny, 16, 16, b, a, dw, 3, 1, 1, 3, 1, 1)
    api.cuStenCompute2DXYp(h, api.DEVICE)
    api.cuStenSwap2DXYp(h, b)
    api.cuStenCompute2DXYp(h, api.DEVICE)
    torch.cuda.synchronize()
    assert h.dataInput[0] == b.data_ptr() and h.dataOutput[0] == a.data_ptr()
    api.cuStenDestroy2DXYp(h)
    r1 = oracle.apply(u, w, 3, 1, 3, 1, True)
    r2 = oracle.apply(r1, w, 3, 1, 3, 1, True)
    assert bits_equal(a.cpu().numpy(), r2)


def test_handle_fields_match_reference_layout(gpu):
    torch = gpu
    nx, ny, nt = 256, 128, 4
    a = dev(torch, np.zeros((ny, nx)))
    b = dev(torch, np.zeros((ny, nx)))
    w = dev(torch, np.zeros(25))
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, nt, nx, ny, 32, 16, b, a, w, 5, 2, 2, 5, 2, 2)
    assert (h.nyTile, h.numSten, h.numStreams, h.xGrid, h.yGrid) == (32, 25, 3, 8, 2)
    assert (h.nxLocal, h.nyLocal) == (36, 20)
    base = a.data_ptr()
    for t in range(nt):
        assert h.dataInput[t] == base + t * 32 * nx * 8
        # custenCreateDestroy2DXYp.cu:194-228 periodic halo pointers
        top = ((t * 32 - 2) % ny) * nx * 8
        bot = (((t + 1) * 32) % ny) * nx * 8
        assert h.boundaryTop[t] == base + top and h.boundaryBottom[t] == base + bot
    assert h.numBoundaryTop == 2 * nx
    api.cuStenDestroy2DXYp(h)
    assert not h.plan


def test_f32_storage(gpu):
    """FP32 storage: bit-exact against the FP32 oracle; the error against the
    FP64 oracle is reported by bench.py."""
    torch = gpu
    for dir_, (H, L, V, T) in ((2, (5, 2, 5, 2)), (0, (5, 2, 1, 0)), (1, (1, 0, 5, 2))):
        for periodic in (1, 0):
            nx, ny = 300, 130
            u64 = synth.field(nx, ny, 55)
            u = u64.astype(np.float32)
            w64 = synth.uniform(H * V, 56, -1, 1)
            init = np.full_like(u, np.float32(SENTINEL))
            du, dout, dw = dev(torch, u), dev(torch, init), dev(torch, w64)
            h = api.cuSten_t()
            api.custenCreate2DF32(h, 0, 1, nx, ny, dir_, periodic, dout, du, dw, H, L, H - 1 - L,
                                  V, T, V - 1 - T)
            api.custenCompute2DF32(h, api.DEVICE)
            torch.cuda.synchronize()
            api.custenDestroy2DF32(h)
            ref = oracle.apply(u, w64.astype(np.float32), H, L, V, T, bool(periodic),
                               zero_right_edge=(dir_ == 0 and not periodic), out=init.copy())
            assert mismatches(dout.cpu().numpy(), ref) == 0


def test_full_size_xy_biharmonic(gpu):
    """8192^2 xy-periodic 13-point biharmonic (5x5 weights, 12 zero): bitwise
    against the oracle on the full grid, and translation equivariant."""
    torch = gpu
    n = 8192
    h_ = 2 * np.pi / n
    pat = np.array([[0, 0, 1, 0, 0], [0, 2, -8, 2, 0], [1, -8, 20, -8, 1],
                    [0, 2, -8, 2, 0], [0, 0, 1, 0, 0]], dtype=float)
    w = (pat / h_ ** 4).ravel()
    u = synth.field(n, n, 2)
    du, dout, dw = dev(torch, u), torch.empty((n, n), dtype=torch.float64, device="cuda"), dev(torch, w)
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, 1, n, n, 32, 32, dout, du, dw, 5, 2, 2, 5, 2, 2)
    api.cuStenCompute2DXYp(h, api.DEVICE)
    torch.cuda.synchronize()
    got = dout.cpu().numpy()
    ref = oracle.apply(u, w, 5, 2, 5, 2, True, nthreads=8)
    assert mismatches(got, ref) == 0
    # shift-equivariance on the device (size-independent property)
    du.copy_(torch.roll(du, shifts=(5, -3), dims=(0, 1)))
    api.cuStenCompute2DXYp(h, api.DEVICE)
    torch.cuda.synchronize()
    assert bits_equal(dout.cpu().numpy(), np.roll(ref, (5, -3), axis=(0, 1)))
    api.cuStenDestroy2DXYp(h)


@pytest.mark.parametrize("rpw,chunks", [(4, 1), (4, 3), (8, 1), (8, 7)])
def test_tile_configs_identical(gpu, rpw, chunks):
    torch = gpu
    nx, ny = 520, 300
    u = synth.field(nx, ny, 61)
    w = synth.uniform(25, 62, -1, 1)
    du, dout, dw = dev(torch, u), dev(torch, np.zeros_like(u)), dev(torch, w)
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, 1, nx, ny, 8, 8, dout, du, dw, 5, 2, 2, 5, 2, 2)
    api.custenSetTileConfig(h, rpw, chunks)
    api.cuStenCompute2DXYp(h, api.DEVICE)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYp(h)
    assert bits_equal(dout.cpu().numpy(), oracle.apply(u, w, 5, 2, 5, 2, True))
