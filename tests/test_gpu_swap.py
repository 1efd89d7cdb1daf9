"""Swap on every variant: the handle's pointer tables after Create and after
Swap follow the reference's formulas, and Compute, Swap, Compute equals two
oracle applies with the roles of the two fields exchanged by hand.

Halo layouts (in rows of the base the pointers are set from):
  wrapped  tile 0 top = ny - T, tile t top = t*nyTile - T,
           tile t bottom = (t+1)*nyTile, last tile bottom = 0
           (custenCreateDestroy2DXYp.cu:194-228, 2DYp.cu:183-214,
           2DXYnp.cu:196-228 at Create; every variant's Swap, e.g.
           2DYnp.cu:248-279, 2DXYnp.cu:265-290)
  clamped  tile 0 top = 0, last tile bottom = ny - B; Create of 2DYnp and
           2DYnpFun only (custenCreateDestroy2DYnp.cu:183-214,
           2DYnpFun.cu:186-216)
"""
import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import mismatches

pytestmark = pytest.mark.gpu

SENTINEL = -3.5

# variant -> (kind, H, L, V, T, builtin function, oracle function id, direction)
VARIANTS = {
    "2DXp": ("w", 5, 2, 1, 0, None, None, 0),
    "2DXnp": ("w", 5, 2, 1, 0, None, None, 0),
    "2DYp": ("w", 1, 0, 5, 2, None, None, 1),
    "2DYnp": ("w", 1, 0, 5, 2, None, None, 1),
    "2DXYp": ("w", 5, 2, 5, 2, None, None, 2),
    "2DXYnp": ("w", 3, 1, 3, 1, None, None, 2),
    "2DXpFun": ("f", 3, 1, 1, 0, "x_central_difference", oracle.FUN_X_CENTRAL_DIFF, 0),
    "2DXnpFun": ("f", 3, 1, 1, 0, "x_central_difference", oracle.FUN_X_CENTRAL_DIFF, 0),
    "2DYpFun": ("f", 1, 0, 9, 4, "y_central_difference9", oracle.FUN_Y_WSUM, 1),
    "2DYnpFun": ("f", 1, 0, 9, 4, "y_central_difference9", oracle.FUN_Y_WSUM, 1),
    "2DXYpFun": ("f", 3, 1, 3, 1, "xy_cube_minus", oracle.FUN_XY_CUBE_MINUS, 2),
    "2DXYnpFun": ("f", 3, 1, 3, 1, "xy_cube_minus", oracle.FUN_XY_CUBE_MINUS, 2),
}


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def expected_halos(ny, nx, nt, T, B, clamped):
    nyt = ny // nt
    tops, bots = [], []
    for t in range(nt):
        top = (0 if clamped else ny - T) if t == 0 else t * nyt - T
        bot = ((ny - B) if clamped else 0) if t == nt - 1 else (t + 1) * nyt
        tops.append(top * nx)
        bots.append(bot * nx)
    return tops, bots


def check_tables(h, nt, nx, ny, T, B, inp, out, esize, halos, clamped):
    nyt = ny // nt
    for t in range(nt):
        assert h.dataInput[t] == inp + t * nyt * nx * esize
        assert h.dataOutput[t] == out + t * nyt * nx * esize
    if not halos:
        assert not h.boundaryTop and not h.boundaryBottom
        return
    tops, bots = expected_halos(ny, nx, nt, T, B, clamped)
    for t in range(nt):
        assert h.boundaryTop[t] == inp + tops[t] * esize, (t, "top")
        assert h.boundaryBottom[t] == inp + bots[t] * esize, (t, "bottom")
    assert h.numBoundaryTop == T * nx and h.numBoundaryBottom == B * nx


def oracle_step(variant, u, w, H, L, V, T, fid, direc, init):
    periodic = "np" not in variant
    out = init.copy()
    if fid is not None:
        return oracle.apply_fun(fid, direc, u, w, H, L, V, T, periodic, out=out)
    x_only = variant.startswith("2DX") and not variant.startswith("2DXY")
    return oracle.apply(u, w, H, L, V, T, periodic, zero_right_edge=x_only and not periodic,
                        out=out)


@pytest.mark.parametrize("nt", [1, 2, 3])
@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_swap_variant(gpu, variant, nt):
    torch = gpu
    kind, H, L, V, T, fname, fid, direc = VARIANTS[variant]
    R, B = H - 1 - L, V - 1 - T
    nx, ny = 160, 96
    u = synth.field(nx, ny, 4100 + nt)
    if kind == "w":
        w = synth.uniform(H * V, 4200, -1.0, 1.0)
    else:
        w = synth.uniform({"x_central_difference": 1}.get(fname, 9), 4201, -1.0, 1.0)
    binit = np.full_like(u, SENTINEL)
    a, b, dw = dev(torch, u), dev(torch, binit), dev(torch, w)
    h = api.cuSten_t()
    create = getattr(api, f"cuStenCreate{variant}")
    if kind == "w":
        if variant.startswith("2DXY"):
            create(h, 0, nt, nx, ny, 32, 8, b, a, dw, H, L, R, V, T, B)
        elif variant.startswith("2DX"):
            create(h, 0, nt, nx, ny, 32, 8, b, a, dw, H, L, R)
        else:
            create(h, 0, nt, nx, ny, 32, 8, b, a, dw, V, T, B)
    else:
        f = api.custenGetBuiltinFunction(fname, 0)
        if variant in ("2DXpFun", "2DXnpFun"):
            create(h, 0, nt, nx, ny, 32, 8, b, a, dw, H, L, R, w.size, f)
        elif variant == "2DYpFun":
            create(h, 0, nt, nx, ny, 32, 8, b, a, dw, V, T, B, w.size, f)
        elif variant == "2DYnpFun":
            create(h, 0, nt, nx, ny, 32, 8, b, a, dw, V, T, B, f)
        else:
            create(h, 0, nt, nx, ny, 32, 8, b, a, dw, H, L, R, V, T, B, f)
    halos = not (variant.startswith("2DX") and not variant.startswith("2DXY"))
    clamped = variant in ("2DYnp", "2DYnpFun")
    check_tables(h, nt, nx, ny, T, B, a.data_ptr(), b.data_ptr(), 8, halos, clamped)

    getattr(api, f"cuStenCompute{variant}")(h, api.DEVICE)
    getattr(api, f"cuStenSwap{variant}")(h, b)
    # after Swap: b is the input, a the output; halos re-based on b, wrapped
    check_tables(h, nt, nx, ny, T, B, b.data_ptr(), a.data_ptr(), 8, halos, False)
    getattr(api, f"cuStenCompute{variant}")(h, api.DEVICE)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)

    r1 = oracle_step(variant, u, w, H, L, V, T, fid, direc, binit)
    r2 = oracle_step(variant, r1, w, H, L, V, T, fid, direc, u)
    got = a.cpu().numpy()
    assert mismatches(got, r2) == 0, f"{mismatches(got, r2)} points differ"
    assert mismatches(b.cpu().numpy(), r1) == 0


@pytest.mark.parametrize("nt", [1, 3])
def test_swap_weno(gpu, nt):
    torch = gpu
    nx, ny = 128, 96
    phi = synth.field(nx, ny, 4300)
    u = synth.field(nx, ny, 4301)
    v = synth.field(nx, ny, 4302)
    a, b = dev(torch, phi), dev(torch, np.zeros_like(phi))
    du, dv = dev(torch, u), dev(torch, v)
    h = api.cuSten_t()
    api.cuStenCreate2DXYWENOADVp(h, 0, nt, nx, ny, 32, 32, 0.1, 0.2, du, dv, b, a)
    check_tables(h, nt, nx, ny, 3, 3, a.data_ptr(), b.data_ptr(), 8, True, False)
    api.cuStenCompute2DXYWENOADVp(h, api.DEVICE)
    api.cuStenSwap2DXYWENOADVp(h, b)
    check_tables(h, nt, nx, ny, 3, 3, b.data_ptr(), a.data_ptr(), 8, True, False)
    api.cuStenCompute2DXYWENOADVp(h, api.DEVICE)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYWENOADVp(h)
    r1 = oracle.weno(phi, u, v, 0.1, 0.2, True)
    r2 = oracle.weno(r1, u, v, 0.1, 0.2, True)
    assert mismatches(a.cpu().numpy(), r2) == 0


@pytest.mark.parametrize("nt", [1, 2, 3])
@pytest.mark.parametrize("dir_,periodic", [(0, 1), (0, 0), (1, 1), (1, 0), (2, 1), (2, 0)])
def test_swap_f32(gpu, dir_, periodic, nt):
    torch = gpu
    H, L, V, T = {0: (5, 2, 1, 0), 1: (1, 0, 5, 2), 2: (5, 2, 5, 2)}[dir_]
    R, B = H - 1 - L, V - 1 - T
    nx, ny = 192, 72
    u = synth.field(nx, ny, 4400 + dir_).astype(np.float32)
    w64 = synth.uniform(H * V, 4410, -1.0, 1.0)
    binit = np.full_like(u, np.float32(SENTINEL))
    a, b, dw = dev(torch, u), dev(torch, binit), dev(torch, w64)
    h = api.cuSten_t()
    api.custenCreate2DF32(h, 0, nt, nx, ny, dir_, periodic, b, a, dw, H, L, R, V, T, B)
    halos = dir_ != 0
    clamped = dir_ == 1 and not periodic
    check_tables(h, nt, nx, ny, T, B, a.data_ptr(), b.data_ptr(), 4, halos, clamped)
    api.custenCompute2DF32(h, api.DEVICE)
    api.custenSwap2DF32(h, b)
    check_tables(h, nt, nx, ny, T, B, b.data_ptr(), a.data_ptr(), 4, halos, False)
    api.custenCompute2DF32(h, api.DEVICE)
    torch.cuda.synchronize()
    api.custenDestroy2DF32(h)
    w32 = w64.astype(np.float32)
    zr = dir_ == 0 and not periodic
    r1 = oracle.apply(u, w32, H, L, V, T, bool(periodic), zero_right_edge=zr, out=binit.copy())
    r2 = oracle.apply(r1, w32, H, L, V, T, bool(periodic), zero_right_edge=zr, out=u.copy())
    assert mismatches(a.cpu().numpy(), r2) == 0
