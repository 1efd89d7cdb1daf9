"""Point-function ("Fun") variants on the GPU: the library's built-in device
functions run inlined and, forced through the function pointer, indirectly;
both are bit-identical to the oracle's restatement of the same function."""
import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import mismatches

pytestmark = pytest.mark.gpu

SENTINEL = -777.0


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run_fun(torch, variant, func_name, u, coe, H, L, V, T, nt, force_indirect, host=False):
    ny, nx = u.shape
    init = np.full_like(u, SENTINEL)
    if host:
        du, dout = torch.from_numpy(u.copy()).pin_memory(), torch.from_numpy(init).pin_memory()
    else:
        du, dout = dev(torch, u), dev(torch, init)
    dc = dev(torch, coe)
    f = api.custenGetBuiltinFunction(func_name, 0)
    assert f
    h = api.cuSten_t()
    R, B = H - 1 - L, V - 1 - T
    if variant in ("2DXpFun", "2DXnpFun"):
        getattr(api, f"cuStenCreate{variant}")(h, 0, nt, nx, ny, 32, 8, dout, du, dc, H, L, R,
                                               coe.size, f)
    elif variant == "2DYpFun":
        api.cuStenCreate2DYpFun(h, 0, nt, nx, ny, 32, 8, dout, du, dc, V, T, B, coe.size, f)
    elif variant == "2DYnpFun":
        api.cuStenCreate2DYnpFun(h, 0, nt, nx, ny, 32, 8, dout, du, dc, V, T, B, f)
    else:
        getattr(api, f"cuStenCreate{variant}")(h, 0, nt, nx, ny, 32, 8, dout, du, dc, H, L, R, V,
                                               T, B, f)
    api.custenSetForceIndirect(h, force_indirect)
    getattr(api, f"cuStenCompute{variant}")(h, api.HOST if host else api.DEVICE)
    torch.cuda.synchronize()
    getattr(api, f"cuStenDestroy{variant}")(h)
    return dout.cpu().numpy() if not host else dout.numpy(), init


FUN_CASES = [
    # variant, builtin name, oracle id, direction, nx, ny, nt, H, L, V, T, ncoe
    ("2DXYpFun", "xy_cube_minus", oracle.FUN_XY_CUBE_MINUS, 2, 256, 128, 1, 3, 1, 3, 1, 9),
    ("2DXYpFun", "xy_cube_minus", oracle.FUN_XY_CUBE_MINUS, 2, 130, 70, 2, 3, 1, 3, 1, 9),
    ("2DXYnpFun", "xy_cube_minus", oracle.FUN_XY_CUBE_MINUS, 2, 200, 64, 4, 3, 1, 3, 1, 9),
    ("2DXYpFun", "xy_weighted_sum", oracle.FUN_XY_WSUM, 2, 256, 96, 1, 5, 2, 5, 2, 25),
    ("2DXYnpFun", "xy_weighted_sum", oracle.FUN_XY_WSUM, 2, 129, 65, 1, 4, 1, 2, 1, 8),
    ("2DXpFun", "x_central_difference", oracle.FUN_X_CENTRAL_DIFF, 0, 256, 64, 1, 3, 1, 1, 0, 1),
    ("2DXnpFun", "x_central_difference", oracle.FUN_X_CENTRAL_DIFF, 0, 190, 40, 4, 3, 1, 1, 0, 1),
    ("2DYpFun", "y_central_difference9", oracle.FUN_Y_WSUM, 1, 64, 64, 4, 1, 0, 9, 4, 9),
    ("2DYnpFun", "y_central_difference9", oracle.FUN_Y_WSUM, 1, 96, 128, 2, 1, 0, 9, 4, 9),
]


@pytest.mark.parametrize("variant,name,fid,direc,nx,ny,nt,H,L,V,T,ncoe", FUN_CASES)
@pytest.mark.parametrize("indirect", [False, True])
def test_builtin_functions(gpu, variant, name, fid, direc, nx, ny, nt, H, L, V, T, ncoe, indirect):
    torch = gpu
    u = synth.field(nx, ny, 300 + nx)
    coe = synth.uniform(ncoe, 301, -1.0, 1.0)
    got, init = run_fun(torch, variant, name, u, coe, H, L, V, T, nt, indirect)
    ref = oracle.apply_fun(fid, direc, u, coe, H, L, V, T, variant.endswith("pFun") and
                           not variant.endswith("npFun"), out=init.copy())
    assert mismatches(got, ref) == 0


@pytest.mark.parametrize("indirect", [False, True])
def test_fun_host_streaming(gpu, indirect):
    torch = gpu
    u = synth.field(160, 96, 77)
    coe = synth.uniform(9, 78, -1.0, 1.0)
    got, init = run_fun(torch, "2DXYpFun", "xy_cube_minus", u, coe, 3, 1, 3, 1, 4, indirect,
                        host=True)
    ref = oracle.apply_fun(oracle.FUN_XY_CUBE_MINUS, 2, u, coe, 3, 1, 3, 1, True, out=init.copy())
    assert mismatches(got, ref) == 0


def test_misspelt_reference_entry_point(gpu):
    """cuSenCompute2DXpFun (the name the reference defines,
    2d_x_p_fun_kernel.cu:172) runs the same computation."""
    torch = gpu
    u = synth.field(128, 32, 5)
    coe = np.array([3.0])
    du, dout, dc = dev(torch, u), dev(torch, np.zeros_like(u)), dev(torch, coe)
    f = api.custenGetBuiltinFunction("x_central_difference", 0)
    h = api.cuSten_t()
    api.cuStenCreate2DXpFun(h, 0, 1, 128, 32, 32, 8, dout, du, dc, 3, 1, 1, 1, f)
    api.cuSenCompute2DXpFun(h, api.DEVICE)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXpFun(h)
    ref = oracle.apply_fun(oracle.FUN_X_CENTRAL_DIFF, 0, u, coe, 3, 1, 1, 0, True)
    assert mismatches(dout.cpu().numpy(), ref) == 0
