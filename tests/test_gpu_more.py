"""Further GPU coverage: FP32 streaming, point functions streamed and
non-periodic, caller streams, the autotuner and its cache, batched 3x3 grids,
and 64-bit indexing on a grid of more than 2^31 points."""
import os

import numpy as np
import pytest

import oracle
from custen_amd import api, synth, workloads
from tests.helpers import bits_equal, mismatches

pytestmark = pytest.mark.gpu

SENTINEL = 4321.0


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("dir_", [0, 1, 2])
@pytest.mark.parametrize("periodic", [1, 0])
@pytest.mark.parametrize("nt", [1, 4])
def test_f32_host_streaming(gpu, dir_, periodic, nt):
    torch = gpu
    nx, ny = 260, 96
    H, L, V, T = {0: (5, 2, 1, 0), 1: (1, 0, 5, 2), 2: (5, 2, 5, 2)}[dir_]
    u = synth.field(nx, ny, 70 + dir_).astype(np.float32)
    w = synth.uniform(H * V, 71, -1.0, 1.0)
    init = np.full_like(u, np.float32(SENTINEL))
    hin = torch.from_numpy(u.copy()).pin_memory()
    hout = torch.from_numpy(init.copy()).pin_memory()
    h = api.cuSten_t()
    api.custenCreate2DF32(h, 0, nt, nx, ny, dir_, periodic, hout, hin, w, H, L, H - 1 - L, V, T,
                          V - 1 - T)
    api.custenCompute2DF32(h, api.HOST)
    torch.cuda.synchronize()
    api.custenDestroy2DF32(h)
    ref = oracle.apply(u, w.astype(np.float32), H, L, V, T, bool(periodic),
                       zero_right_edge=(dir_ == 0 and not periodic), out=init.copy())
    assert mismatches(hout.numpy(), ref) == 0


@pytest.mark.parametrize("indirect", [False, True])
def test_fun_np_host_streaming(gpu, indirect):
    torch = gpu
    u = synth.field(200, 128, 81)
    coe = synth.uniform(9, 82, -1.0, 1.0)
    init = np.full_like(u, SENTINEL)
    hin = torch.from_numpy(u.copy()).pin_memory()
    hout = torch.from_numpy(init.copy()).pin_memory()
    f = api.custenGetBuiltinFunction("xy_cube_minus", 0)
    h = api.cuSten_t()
    api.cuStenCreate2DXYnpFun(h, 0, 4, 200, 128, 16, 16, hout, hin, dev(torch, coe), 3, 1, 1, 3, 1,
                              1, f)
    api.custenSetForceIndirect(h, indirect)
    api.cuStenCompute2DXYnpFun(h, api.HOST)
    api.cuStenDestroy2DXYnpFun(h)
    ref = oracle.apply_fun(oracle.FUN_XY_CUBE_MINUS, 2, u, coe, 3, 1, 3, 1, False, out=init.copy())
    assert mismatches(hout.numpy(), ref) == 0


def test_caller_stream(gpu):
    torch = gpu
    nx, ny = 512, 256
    u = synth.field(nx, ny, 91)
    w = synth.uniform(9, 92, -1.0, 1.0)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        du = dev(torch, u)
        dout = torch.zeros_like(du)
        dw = dev(torch, w)
        h = api.cuSten_t()
        api.cuStenCreate2DXYp(h, 0, 1, nx, ny, 32, 32, dout, du, dw, 3, 1, 1, 3, 1, 1)
        api.custenSetStream(h, s.cuda_stream)
        api.cuStenCompute2DXYp(h, api.DEVICE)
        got = dout.cpu()  # ordered on the same stream
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYp(h)
    assert bits_equal(got.numpy(), oracle.apply(u, w, 3, 1, 3, 1, True))


def test_autotuner_choice(gpu):
    """A grid of >= 2^20 points is tuned on its first Compute; the choice is
    reported and does not change the result."""
    torch = gpu
    n = 1024
    u = synth.field(n, n, 93)
    w = synth.uniform(25, 94, -1.0, 1.0)
    du, dout, dw = dev(torch, u), dev(torch, np.zeros_like(u)), dev(torch, w)
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, 1, n, n, 32, 32, dout, du, dw, 5, 2, 2, 5, 2, 2)
    assert api.custenGetTileConfig(h) == (0, 0, 0)
    api.cuStenCompute2DXYp(h, api.DEVICE)
    torch.cuda.synchronize()
    rpw, chunks, cpl = api.custenGetTileConfig(h)
    assert rpw in (4, 8) and chunks >= 1 and cpl == 2
    api.cuStenDestroy2DXYp(h)
    assert bits_equal(dout.cpu().numpy(), oracle.apply(u, w, 5, 2, 5, 2, True))


def test_batch_3x3_and_validation(gpu):
    torch = gpu
    sizes = [130, 64, 257]
    w = synth.uniform(9, 95, -0.2, 0.2)
    fields = [synth.field(n, n, 96 + g) for g, n in enumerate(sizes)]
    ins = [dev(torch, f) for f in fields]
    outs = [torch.zeros_like(t) for t in ins]
    b = api.Batch(outs, ins, dev(torch, w), H=3, L=1, V=3, T=1)
    b.run(3)
    torch.cuda.synchronize()
    cur = outs if b.parity == 1 else ins
    for g, f in enumerate(fields):
        ref = f
        for _ in range(3):
            ref = oracle.apply(ref, w, 3, 1, 3, 1, True)
        assert bits_equal(cur[g].cpu().numpy(), ref)
    b.close()


def test_64bit_indexing_resident(gpu):
    """65536 x 32800 FP64 (> 2^31 points) resident in HBM: rows at the end of
    the grid, beyond 32-bit element offsets, match the oracle."""
    torch = gpu
    nx, ny = 65536, 32800
    pts = nx * ny
    assert pts > 2 ** 31
    w = workloads.laplacian9_weights(nx)
    din = torch.empty((ny, nx), dtype=torch.float64, device="cuda")
    host = np.empty((1024, nx))
    for r0 in range(0, ny, 1024):  # fill the device field slab by slab
        rows = min(1024, ny - r0)
        api.custenFillUniform(host, rows * nx, 3, -1.0, 1.0)
        host[:rows] += r0  # distinct slabs
        din[r0:r0 + rows].copy_(torch.from_numpy(host[:rows]))
    dout = torch.empty_like(din)
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, 1, nx, ny, 32, 32, dout, din, dev(torch, w), 3, 1, 1, 3, 1, 1)
    api.custenSetTileConfig(h, 8, 4)
    api.cuStenCompute2DXYp(h, api.DEVICE)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYp(h)
    for y0 in (ny - 3, 32767, 0):
        blk = din[torch.tensor(np.arange(y0 - 1, y0 + 4) % ny, device="cuda")].cpu().numpy()
        ref = oracle.apply(np.ascontiguousarray(blk), w, 3, 1, 3, 1, False)
        got = dout[torch.tensor(np.arange(y0, y0 + 3) % ny, device="cuda")].cpu().numpy()
        assert mismatches(got[:, 1:-1], np.ascontiguousarray(ref[1:4, 1:-1])) == 0


def test_autotuner_cache_file(gpu, tmp_path):
    """With CUSTEN_TUNE_CACHE set, the tuner appends its choice to the file and
    a later process reuses it (run in child processes, since the table is
    loaded once per process)."""
    import subprocess
    import sys

    cache = tmp_path / "tune.txt"
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import numpy as np, torch\n"
        "from custen_amd import api\n"
        "n = 1024\n"
        "a = torch.rand(n, n, dtype=torch.float64, device='cuda'); b = torch.empty_like(a)\n"
        "w = torch.rand(9, dtype=torch.float64, device='cuda')\n"
        "h = api.cuSten_t()\n"
        "api.cuStenCreate2DXYp(h, 0, 1, n, n, 32, 32, b, a, w, 3, 1, 1, 3, 1, 1)\n"
        "api.cuStenCompute2DXYp(h, api.DEVICE); torch.cuda.synchronize()\n"
        "print(api.custenGetTileConfig(h)); api.cuStenDestroy2DXYp(h)\n"
    ) % os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, CUSTEN_TUNE_CACHE=str(cache))
    r1 = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                        timeout=240)
    assert r1.returncode == 0, r1.stderr[-2000:]
    lines = cache.read_text().strip().splitlines()
    assert len(lines) == 1 and len(lines[0].split()) == 12
    r2 = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                        timeout=240)
    assert r2.returncode == 0, r2.stderr[-2000:]
    assert r1.stdout.strip() == r2.stdout.strip()
    assert len(cache.read_text().strip().splitlines()) == 1  # reused, not re-tuned


@pytest.mark.parametrize("band", [False, True])
def test_compute_captured_in_a_graph(gpu, band):
    """Compute launched into a caller's HIP graph (via custenSetStream on the
    capturing stream): every replay, with eager work in between that changes
    the input, matches the oracle on the input of that moment (strip kernel
    and band walker)."""
    torch = gpu
    n = 512
    u0 = synth.field(n, n, seed=71)
    w = synth.uniform(25, seed=72, a=-1.0, b=1.0)
    din = torch.from_numpy(u0.copy()).cuda()
    dout = torch.zeros_like(din)
    dw = torch.from_numpy(w).cuda()
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, 1, n, n, 32, 32, dout, din, dw, 5, 2, 2, 5, 2, 2)
    if band:
        api.custenSetBandConfig(h, 4, 2)
    s = torch.cuda.Stream()
    api.custenSetStream(h, s.cuda_stream)
    with torch.cuda.stream(s):
        api.cuStenCompute2DXYp(h, api.DEVICE)  # tunes / warms up outside the capture
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        api.cuStenCompute2DXYp(h, api.DEVICE)
    for rep in range(3):
        din.mul_(1.5).add_(0.25)  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        cur = din.cpu().numpy()
        ref = oracle.apply(cur, w, 5, 2, 5, 2, True)
        got = dout.cpu().numpy()
        assert bits_equal(got, ref), (rep, mismatches(got, ref))
    del graph
    api.cuStenDestroy2DXYp(h)


@pytest.mark.parametrize("H,L,V,T", [(7, 3, 7, 3), (9, 4, 9, 4), (3, 1, 5, 2), (13, 6, 1, 0)])
def test_f32_wide_shapes(gpu, H, L, V, T):
    torch = gpu
    nx, ny = 384, 96
    u = synth.field(nx, ny, seed=81).astype(np.float32)
    w = synth.uniform(H * V, seed=82, a=-1.0, b=1.0)
    din = torch.from_numpy(u).cuda()
    dout = torch.zeros_like(din)
    dw = torch.from_numpy(w).cuda()
    h = api.cuSten_t()
    api.custenCreate2DF32(h, 0, 1, nx, ny, 2 if V > 1 else 0, 1, dout, din, dw, H, L, H - 1 - L,
                          V, T, V - 1 - T)
    api.custenCompute2DF32(h, api.DEVICE)
    torch.cuda.synchronize()
    api.custenDestroy2DF32(h)
    ref = oracle.apply(u, w.astype(np.float32), H, L, V, T, True)
    got = dout.cpu().numpy()
    assert bits_equal(got, ref), mismatches(got, ref)


@pytest.mark.parametrize("nt", [16])
def test_host_streaming_config3_width(gpu, nt):
    """Config 3's row width (65536 columns) and tile count (numTiles = 16),
    xy periodic 9-point, host-pinned input and output streamed through HBM:
    bit for bit the oracle.  512 rows keep the oracle under a second."""
    torch = gpu
    nx, ny = 65536, 512
    u = synth.field(nx, ny, 3)
    w = synth.uniform(9, 31, -1.0, 1.0)
    hin = torch.from_numpy(u).pin_memory()
    hout = torch.zeros(ny, nx, dtype=torch.float64).pin_memory()
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, nt, nx, ny, 32, 32, hout, hin, w, 3, 1, 1, 3, 1, 1)
    assert h.nyTile == ny // nt
    api.cuStenCompute2DXYp(h, api.HOST)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYp(h)
    ref = oracle.apply(u, w, 3, 1, 3, 1, True, nthreads=8)
    assert mismatches(hout.numpy(), ref) == 0
