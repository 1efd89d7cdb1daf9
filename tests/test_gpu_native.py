"""Native callers of the C ABI, run as child processes (tests/native/):

  weights_caller  a C++ main shaped like examples/src/2d_xy_p.cu, linked
                  against libcusten.so: Create, Compute, Swap, Compute,
                  Destroy on managed memory, checkError after each call;
  fun_caller      a C++ main with its own __device__ point functions, fetched
                  with hipMemcpyFromSymbol as in examples/src/2d_xy_p_fun.cu
                  and cuPentCahnADI.cu:164-190,486-516, compiled with
                  -fgpu-rdc and linked against libcusten.a;
  point_caller    the same kind of caller with its functions compiled into
                  its own kernels (include/custen_point.h), linked against
                  libcusten.so without relocatable device code.

What each writes is compared bit for bit with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from custen_amd import synth
from tests.helpers import mismatches

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "bin")


def run(args, tmp_path):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (make -C tests/native)")
    print(f"native: {' '.join(map(str, args))}", flush=True)
    res = subprocess.run([exe, *map(str, args[1:])], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, f"{args[0]} exited {res.returncode}: {res.stdout}{res.stderr}"
    return res


@pytest.mark.parametrize("periodic,nt", [(1, 1), (1, 4), (0, 1), (0, 3)])
def test_weights_caller(gpu, tmp_path, periodic, nt):
    nx, ny, seed = 384, 240, 6100 + nt
    out = tmp_path / "out.bin"
    run(["weights_caller", nx, ny, nt, periodic, seed, out], tmp_path)
    raw = np.fromfile(out, dtype=np.uint8)
    n = nx * ny
    field = raw[: n * 8].view(np.float64).reshape(ny, nx)
    offs = raw[n * 8:].view(np.int64).reshape(nt, 4)

    u = synth.field(nx, ny, seed)
    w = synth.uniform(25, seed + 1, -1.0, 1.0)
    r1 = oracle.apply(u, w, 5, 2, 5, 2, bool(periodic), out=np.zeros_like(u))
    r2 = oracle.apply(r1, w, 5, 2, 5, 2, bool(periodic), out=u.copy())
    assert mismatches(field, r2) == 0

    # after Swap(dataOutput): inputs follow the old output buffer, halos are
    # re-based on it with the wrapped layout (custenCreateDestroy2DXYp.cu:253-301,
    # 2DXYnp.cu:265-290)
    nyt = ny // nt
    for t in range(nt):
        top = (ny - 2) if t == 0 else t * nyt - 2
        bot = 0 if t == nt - 1 else (t + 1) * nyt
        assert offs[t, 0] == t * nyt * nx
        assert offs[t, 2] == top * nx and offs[t, 3] == bot * nx


def test_fun_caller_own_device_functions(gpu, tmp_path):
    nx, ny, seed = 520, 136, 6200
    out = tmp_path / "fun.bin"
    run(["fun_caller", nx, ny, seed, out], tmp_path)
    got = np.fromfile(out, dtype=np.float64).reshape(3, ny, nx)
    u = synth.field(nx, ny, seed)
    coe = synth.uniform(9, seed + 1, -1.0, 1.0)
    ref_x = oracle.apply_fun(oracle.FUN_X_WSUM, 0, u, coe[:5], 5, 2, 1, 0, True)
    ref_xy = oracle.apply_fun(oracle.FUN_XY_WSUM, 2, u, coe, 3, 1, 3, 1, True)
    ref_nl = oracle.apply_fun(oracle.FUN_XY_CUBE_MINUS, 2, u, coe, 3, 1, 3, 1, False)
    assert mismatches(got[0], ref_x) == 0, "x periodic, caller's devArg1X"
    assert mismatches(got[1], ref_xy) == 0, "xy periodic, caller's devArg1XY"
    assert mismatches(got[2], ref_nl) == 0, "xy non-periodic, caller's nonLinRHS"


def test_point_caller_inlined_functions(gpu, tmp_path):
    nx, ny, seed = 520, 136, 6300
    out = tmp_path / "point.bin"
    run(["point_caller", nx, ny, seed, out], tmp_path)
    got = np.fromfile(out, dtype=np.float64).reshape(5, ny, nx)
    u = synth.field(nx, ny, seed)
    coe = synth.uniform(9, seed + 1, -1.0, 1.0)
    ref_x = oracle.apply_fun(oracle.FUN_X_WSUM, 0, u, coe[:5], 5, 2, 1, 0, True)
    ref_y = oracle.apply_fun(oracle.FUN_Y_WSUM, 1, u, coe[:5], 1, 0, 5, 2, False)
    ref_xy = oracle.apply_fun(oracle.FUN_XY_WSUM, 2, u, coe, 3, 1, 3, 1, True)
    ref_nl = oracle.apply_fun(oracle.FUN_XY_CUBE_MINUS, 2, u, coe, 3, 1, 3, 1, False)
    assert mismatches(got[0], ref_x) == 0, "x periodic"
    assert mismatches(got[1], ref_y) == 0, "y non-periodic"
    assert mismatches(got[2], ref_xy) == 0, "xy periodic"
    assert mismatches(got[3], ref_nl) == 0, "xy non-periodic nonLinRHS"
    assert mismatches(got[4], ref_xy) == 0, "xy periodic, host-streamed in 4 tiles"
