"""GPU parity of the WENO5 advection variant (cuStenCreate/Compute/Swap/
Destroy2DXYWENOADVp) against the CPU oracle, through the C ABI.

Bar: bit-exact equality with the oracle at every point, for both treatments of
the smoothness squares, for device-resident data (one launch) and for
host-resident data streamed in numTiles tiles."""
import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import bits_equal, mismatches

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def case(nx, ny, seed, zero_frac=0.1):
    phi = synth.field(nx, ny, seed=seed)
    u = synth.field(nx, ny, seed=seed + 1)
    v = synth.field(nx, ny, seed=seed + 2)
    # exact zeros take the u <= 0 side
    u[synth.field(nx, ny, seed=seed + 3, a=0.0, b=1.0) < zero_frac] = 0.0
    v[synth.field(nx, ny, seed=seed + 4, a=0.0, b=1.0) < zero_frac] = 0.0
    return phi, u, v


def run_resident(torch, phi, u, v, dx, dy, single, numTiles=1, rpw=0):
    ny, nx = phi.shape
    dphi, du, dv = dev(torch, phi), dev(torch, u), dev(torch, v)
    dout = torch.full_like(dphi, 7.0)
    h = api.cuSten_t()
    api.cuStenCreate2DXYWENOADVp(h, 0, numTiles, nx, ny, 32, 32, dx, dy, du, dv, dout, dphi)
    if not single:
        api.custenSetWenoSquares(h, False)
    if rpw:
        api.custenSetTileConfig(h, rpw, 1)
    api.cuStenCompute2DXYWENOADVp(h, api.DEVICE)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYWENOADVp(h)
    return dout.cpu().numpy()


@pytest.mark.parametrize("nx,ny", [(64, 64), (100, 37), (256, 128), (7, 9), (1024, 520)])
@pytest.mark.parametrize("single", [True, False])
def test_resident_matches_oracle(gpu, nx, ny, single):
    print(f"weno resident {nx}x{ny} single={single}", flush=True)
    phi, u, v = case(nx, ny, 100 + nx)
    dx, dy = 2 * np.pi / nx, 2 * np.pi / ny
    got = run_resident(gpu, phi, u, v, dx, dy, single)
    ref = oracle.weno(phi, u, v, dx, dy, single, nthreads=8)
    assert bits_equal(got, ref), mismatches(got, ref)


@pytest.mark.parametrize("rpw", [4, 8])
def test_rows_per_wave_identical(gpu, rpw):
    phi, u, v = case(300, 200, 7)
    got = run_resident(gpu, phi, u, v, 0.01, 0.02, True, rpw=rpw)
    ref = oracle.weno(phi, u, v, 0.01, 0.02, True, nthreads=8)
    assert bits_equal(got, ref), mismatches(got, ref)


def test_example_fields_kat(gpu):
    """The reference example's fields (examples/src/2d_xyWENOADV_p.cu) at 512^2:
    identical to the oracle and within truncation error of the exact term."""
    n = 512
    h = 2 * np.pi / n
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    phi, u, v = np.cos(X) * np.sin(Y), np.sin(Y), -np.sin(X)
    exact = u * (-np.sin(X) * np.sin(Y)) + v * (np.cos(X) * np.cos(Y))
    got = run_resident(gpu, phi, u, v, h, h, True, numTiles=4)
    assert bits_equal(got, oracle.weno(phi, u, v, h, h, True, nthreads=8))
    assert np.abs(got - exact).max() < 1e-10


@pytest.mark.parametrize("single", [True, False])
def test_tiny_gradients(gpu, single):
    """Differences scaled to ~1e-22 make the single-precision squares
    subnormal (or zero): the GPU's float arithmetic must keep them as the CPU
    does."""
    phi, u, v = case(96, 64, 31)
    got = run_resident(gpu, phi, u, v, 1e22, 3e21, single)
    ref = oracle.weno(phi, u, v, 1e22, 3e21, single)
    assert bits_equal(got, ref), mismatches(got, ref)


@pytest.mark.parametrize("numTiles", [1, 2, 4])
@pytest.mark.parametrize("pinned", [False, True])
def test_host_streaming_matches_oracle(gpu, numTiles, pinned):
    nx, ny = 320, 96
    print(f"weno streamed tiles={numTiles} pinned={pinned}", flush=True)
    phi, u, v = case(nx, ny, 55)
    ref = oracle.weno(phi, u, v, 0.05, 0.07, True, nthreads=8)
    if pinned:
        t = lambda a: gpu.from_numpy(a).pin_memory()  # noqa: E731
        hphi, hu, hv = t(phi), t(u), t(v)
        hout = gpu.full_like(hphi, 7.0).pin_memory()
    else:
        hphi, hu, hv = phi.copy(), u.copy(), v.copy()
        hout = np.full_like(phi, 7.0)
    h = api.cuSten_t()
    api.cuStenCreate2DXYWENOADVp(h, 0, numTiles, nx, ny, 32, 32, 0.05, 0.07, hu, hv, hout, hphi)
    api.cuStenCompute2DXYWENOADVp(h, api.HOST)
    api.cuStenDestroy2DXYWENOADVp(h)
    got = hout.numpy() if pinned else hout
    assert bits_equal(got, ref), mismatches(got, ref)


def test_swap_time_steps(gpu):
    """Three forward-Euler-like applies with Swap equal three oracle applies
    with the roles exchanged by hand."""
    nx, ny = 128, 80
    phi, u, v = case(nx, ny, 77)
    a, b = dev(gpu, phi), gpu.zeros(ny, nx, dtype=gpu.float64, device="cuda")
    du, dv = dev(gpu, u), dev(gpu, v)
    h = api.cuSten_t()
    api.cuStenCreate2DXYWENOADVp(h, 0, 2, nx, ny, 32, 32, 0.1, 0.1, du, dv, b, a)
    ref_in = phi.copy()
    bufs = [a, b]
    for step in range(3):
        api.cuStenCompute2DXYWENOADVp(h, api.DEVICE)
        ref_out = oracle.weno(ref_in, u, v, 0.1, 0.1, True)
        ref_in = ref_out
        api.cuStenSwap2DXYWENOADVp(h, bufs[(step + 1) % 2])
    gpu.cuda.synchronize()
    api.cuStenDestroy2DXYWENOADVp(h)
    got = bufs[1].cpu().numpy()  # three applies: a -> b -> a -> b
    assert bits_equal(got, ref_in), mismatches(got, ref_in)

