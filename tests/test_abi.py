"""The C-ABI library loads, exports every entry point include/cuSten.h
declares (the reference's 48 Create/Compute/Swap/Destroy + the misspelt
cuSenCompute2DXpFun + checkError, the lower-case aliases and the extensions),
and validates arguments before touching the GPU."""
import ctypes

import pytest

from custen_amd import api


def test_library_loads_and_exports_header_symbols():
    L = api.lib()
    names = api.header_symbols()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    # the reference surface: 12 variants x 4 entry points
    ref = [f"cuSten{op}2D{v}" for op in ("Create", "Compute", "Swap", "Destroy")
           for v in ("Xp", "Xnp", "XpFun", "XnpFun", "Yp", "Ynp", "YpFun", "YnpFun",
                     "XYp", "XYnp", "XYpFun", "XYnpFun")]
    assert len(ref) == 48 and all(n in names for n in ref)
    assert "cuSenCompute2DXpFun" in names and "checkError" in names


def test_aliases_share_addresses():
    L = api.lib()
    for v in ("Create2DXYp", "Compute2DXYpFun", "Swap2DYnp", "Destroy2DXnpFun"):
        a = ctypes.cast(getattr(L, "cuSten" + v), ctypes.c_void_p).value
        b = ctypes.cast(getattr(L, "custen" + v), ctypes.c_void_p).value
        assert a == b


def test_handle_size_matches_binding():
    assert api.lib().custenHandleSize() == ctypes.sizeof(api.cuSten_t)


@pytest.mark.parametrize("args,msg", [
    ((0, 1, 64, 64, 8, 8, 1, 2, 3, 5, 2, 1, 5, 2, 2), "horizontal stencil size"),
    ((0, 3, 64, 64, 8, 8, 1, 2, 3, 5, 2, 2, 5, 2, 2), "divisible by numTiles"),
    ((0, 1, 0, 64, 8, 8, 1, 2, 3, 5, 2, 2, 5, 2, 2), "nx and ny"),
    ((0, 1, 64, 64, 8, 8, 1, 1, 3, 5, 2, 2, 5, 2, 2), "alias"),
    ((0, 1, 64, 64, 8, 8, 1, 2, None, 5, 2, 2, 5, 2, 2), "weights"),
    ((0, 1, 64, 64, 8, 8, 1, 2, 3, 5, -1, 5, 5, 2, 2), "extents"),
])
def test_create_validation(args, msg):
    h = api.cuSten_t()
    with pytest.raises(api.CustenError) as ei:
        api.cuStenCreate2DXYp(h, *args)
    assert msg in str(ei.value)
    assert ei.value.code == 1
    assert not h.plan


@pytest.mark.parametrize("args,msg", [
    # (dx, dy, u, v, out, in)
    ((0.1, 0.1, None, 5, 2, 1), "null velocity"),
    ((0.1, 0.1, 2, 5, 2, 1), "alias the output"),
    ((0.1, 0.1, 4, 5, 1, 1), "must not alias"),
    ((0.0, 0.1, 4, 5, 2, 1), "dx and dy"),
    ((0.1, float("inf"), 4, 5, 2, 1), "dx and dy"),
])
def test_weno_create_validation(args, msg):
    h = api.cuSten_t()
    with pytest.raises(api.CustenError) as ei:
        api.cuStenCreate2DXYWENOADVp(h, 0, 1, 64, 64, 32, 32, *args)
    assert msg in str(ei.value)
    assert not h.plan


@pytest.mark.parametrize("m,batch,msg", [(3, 4, "m must be >= 4"), (8, 0, "batchCount")])
def test_pent_validation_without_gpu(m, batch, msg):
    with pytest.raises(api.CustenError) as ei:
        api.custenPentFactorBatch(1, 2, 3, 4, 5, m, batch)
    assert msg in str(ei.value)
    with pytest.raises(api.CustenError) as ei:
        api.custenPentSolveBatch(1, 2, 3, 4, 5, 6, m, batch)
    assert msg in str(ei.value)


def test_uncreated_handle_is_rejected():
    h = api.cuSten_t()
    with pytest.raises(api.CustenError):
        api.cuStenCompute2DXYp(h, api.DEVICE)
    with pytest.raises(api.CustenError):
        api.cuStenDestroy2DXYp(h)


def test_tile_and_band_config_need_a_handle():
    h = api.cuSten_t()
    for call in (lambda: api.custenSetBandConfig(h, 4, 2), lambda: api.custenGetBandMode(h),
                 lambda: api.custenSetTileConfig(h, 4, 2)):
        with pytest.raises(api.CustenError):
            call()


def test_builtin_names():
    with pytest.raises(api.CustenError):
        api.custenGetBuiltinFunction("no_such_function", 0)


def test_ch_validation_without_gpu():
    L = api.lib()
    p = api.ch_params(4)  # below the minimum grid
    import numpy as np
    c0 = np.zeros(16)
    assert not L.custenCHCreate(0, ctypes.byref(p), c0.ctypes.data)
    assert L.custenGetLastError() == 1
    api.clear_error()
    # three field buffers used in turn, two work fields, the factors
    assert L.custenCHFootprint(64) == 5 * 64 * 64 * 8 + 7 * 62 * 8
    assert not L.custenCHServerCreate(0, 0, 4)
    assert L.custenGetLastError() == 1
    api.clear_error()


def test_batch_validation_without_gpu():
    L = api.lib()
    nx = (ctypes.c_int * 1)(64)
    ptr = (ctypes.c_void_p * 1)(0)
    assert not L.custenBatchCreate(0, 1, nx, nx, ptr, ptr, None, 5, 2, 5, 2, 0)
    assert L.custenGetLastError() == 1
    api.clear_error()
    w = (ctypes.c_double * 49)()
    assert not L.custenBatchCreate(0, 1, nx, nx, ptr, ptr, w, 7, 3, 7, 3, 0)
    assert L.custenGetLastError() == 4  # shape not supported for batches
    api.clear_error()
    assert L.custenBatchSetTileConfig(None, 4, 2, 1) == 1
    assert L.custenBatchGetTileConfig(None, None, None, None) == 1
    api.clear_error()


def test_ch_restore_rejects_bad_files(tmp_path):
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"NOTACKPT" + bytes(48))
    with pytest.raises(api.CustenError, match="not a complete checkpoint"):
        api.CHSolver.restore(str(bad))
    short = tmp_path / "short.ckpt"
    import struct
    short.write_bytes(b"CUSTENCH" + struct.pack("<II4dq", 1, 16, 6.28, 1.0, 0.01, 0.1, 3) +
                      bytes(16 * 16 * 8))  # cOld missing
    with pytest.raises(api.CustenError, match="not a complete checkpoint"):
        api.CHSolver.restore(str(short))
    with pytest.raises(api.CustenError, match="cannot open"):
        api.CHSolver.restore(str(tmp_path / "missing.ckpt"))


def test_check_error_prints_and_exits(tmp_path):
    """checkError after a refused Create prints error.cu:46's text,
    "\\nError while '<action>': <error>\\nprogram terminated ...\\n\\n", and
    exits with EXIT_FAILURE (error.cu:41-50).  Run in a child process."""
    import re
    import subprocess
    import sys

    script = tmp_path / "child.py"
    script.write_text(
        "import ctypes, sys\n"
        f"sys.path.insert(0, {api._HERE!r} + '/..')\n"
        "from custen_amd import api\n"
        "L = api.lib()\n"
        "h = api.cuSten_t()\n"
        "L.cuStenCreate2DXYp(ctypes.byref(h), 0, 1, 0, 64, 8, 8, 1, 2, 3, 5, 2, 2, 5, 2, 2)\n"
        "sys.stderr.write(L.custenGetLastErrorString().decode() + '\\n')\n"
        "sys.stderr.flush()\n"
        "L.checkError(b'creating the handle')\n"
        "sys.exit(0)\n")
    res = subprocess.run([sys.executable, str(script)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 1, (res.returncode, res.stdout, res.stderr)
    m = re.fullmatch(r"\nError while 'creating the handle': (.+)\nprogram terminated \.\.\.\n\n",
                     res.stdout)
    assert m, repr(res.stdout)
    # either the library's own record of the refused Create or, on a host
    # without a GPU, the HIP runtime's pending error
    recorded = res.stderr.strip().splitlines()[-1]
    assert "nx and ny must be >= 1" in recorded
    assert m.group(1) == recorded or "device" in m.group(1).lower()

