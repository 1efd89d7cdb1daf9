"""ctypes binding of the custen C ABI (include/cuSten.h).

The functions here keep the reference's names and argument order
(cuSten/src/struct/cuSten_struct_functions.h, cuSten/src/kernels/stencil_kernels.h),
so a Python caller writes the same sequence a C caller of the reference does:

    h = cuSten_t()
    cuStenCreate2DXYp(h, 0, numTiles, nx, ny, BX, BY, out, inp, weights, 5, 2, 2, 5, 2, 2)
    cuStenCompute2DXYp(h, DEVICE)
    cuStenDestroy2DXYp(h)

Data arguments may be torch tensors (device or host), numpy arrays (host),
raw integer addresses or None.  Where the C library records an error, the
wrapper raises CustenError (the C entry point `checkError` keeps the
reference's print-and-exit behaviour for C callers).

The native library must be built (custen_amd.build); there is no fallback.
"""
from __future__ import annotations

import ctypes
import os
import re

DEVICE = 0
HOST = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
# CUSTEN_LIBRARY selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("CUSTEN_LIBRARY") or os.path.join(_HERE, "lib", "libcusten.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")
HEADER_PATH = os.path.join(INCLUDE_DIR, "cuSten.h")

c_int = ctypes.c_int
c_bool = ctypes.c_bool
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
c_uint64 = ctypes.c_uint64


class CustenError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"custen error {code}: {msg}")
        self.code = code
        self.msg = msg


class cuSten_t(ctypes.Structure):
    """Mirror of cuSten_t (include/cuSten.h; reference layout
    cuSten/src/struct/cuSten_struct_type.h:84-122 plus the trailing plan)."""

    _fields_ = [
        ("deviceNum", c_int),
        ("numStreams", c_int),
        ("numTiles", c_int),
        ("nx", c_int),
        ("ny", c_int),
        ("nyTile", c_int),
        ("numSten", c_int),
        ("numStenLeft", c_int),
        ("numStenRight", c_int),
        ("numStenTop", c_int),
        ("numStenBottom", c_int),
        ("numStenHoriz", c_int),
        ("numStenVert", c_int),
        ("BLOCK_X", c_int),
        ("BLOCK_Y", c_int),
        ("xGrid", c_int),
        ("yGrid", c_int),
        ("mem_shared", c_int),
        ("dataInput", ctypes.POINTER(c_void_p)),
        ("dataOutput", ctypes.POINTER(c_void_p)),
        ("uVel", ctypes.POINTER(c_void_p)),
        ("vVel", ctypes.POINTER(c_void_p)),
        ("weights", c_void_p),
        ("coe", c_void_p),
        ("coeDx", c_double),
        ("coeDy", c_double),
        ("numCoe", c_int),
        ("nxLocal", c_int),
        ("nyLocal", c_int),
        ("boundaryTop", ctypes.POINTER(c_void_p)),
        ("boundaryBottom", ctypes.POINTER(c_void_p)),
        ("numBoundaryTop", c_int),
        ("numBoundaryBottom", c_int),
        ("streams", ctypes.POINTER(c_void_p)),
        ("events", ctypes.POINTER(c_void_p)),
        ("devFunc", c_void_p),
        ("plan", c_void_p),
    ]


class custenCHParams(ctypes.Structure):
    _fields_ = [
        ("n", c_int),
        ("lx", c_double),
        ("D", c_double),
        ("gamma", c_double),
        ("dtOverDx", c_double),
    ]


_H = ctypes.POINTER(cuSten_t)
_P = c_void_p

# name -> (restype, argtypes)
_X_CREATE = [_H, c_int, c_int, c_int, c_int, c_int, c_int, _P, _P, _P, c_int, c_int, c_int]
_XFUN_CREATE = _X_CREATE + [c_int, _P]
_YNPFUN_CREATE = _X_CREATE + [_P]
_XY_CREATE = [_H, c_int, c_int, c_int, c_int, c_int, c_int, _P, _P, _P,
              c_int, c_int, c_int, c_int, c_int, c_int]
_XYFUN_CREATE = _XY_CREATE + [_P]

_VARIANTS = {
    "2DXnp": _X_CREATE, "2DXp": _X_CREATE,
    "2DXnpFun": _XFUN_CREATE, "2DXpFun": _XFUN_CREATE,
    "2DYnp": _X_CREATE, "2DYp": _X_CREATE,
    "2DYnpFun": _YNPFUN_CREATE, "2DYpFun": _XFUN_CREATE,
    "2DXYnp": _XY_CREATE, "2DXYp": _XY_CREATE,
    "2DXYnpFun": _XYFUN_CREATE, "2DXYpFun": _XYFUN_CREATE,
}

_SIGS: dict[str, tuple] = {}
for _v, _cargs in _VARIANTS.items():
    for _pre in ("cuSten", "custen"):
        _SIGS[f"{_pre}Create{_v}"] = (None, _cargs)
        _SIGS[f"{_pre}Compute{_v}"] = (None, [_H, c_bool])
        _SIGS[f"{_pre}Swap{_v}"] = (None, [_H, _P])
        _SIGS[f"{_pre}Destroy{_v}"] = (None, [_H])
_WENO_CREATE = [_H, c_int, c_int, c_int, c_int, c_int, c_int, c_double, c_double, _P, _P, _P, _P]
for _pre in ("cuSten", "custen"):
    _SIGS[f"{_pre}Create2DXYWENOADVp"] = (None, _WENO_CREATE)
    _SIGS[f"{_pre}Compute2DXYWENOADVp"] = (None, [_H, c_bool])
    _SIGS[f"{_pre}Swap2DXYWENOADVp"] = (None, [_H, _P])
    _SIGS[f"{_pre}Destroy2DXYWENOADVp"] = (None, [_H])
_SIGS.update({
    "custenSetWenoSquares": (None, [_H, c_int]),
    "custenPentFactorBatch": (c_int, [_P, _P, _P, _P, _P, c_int, c_int, c_void_p]),
    "custenPentSolveBatch": (c_int, [_P, _P, _P, _P, _P, _P, c_int, c_int, c_void_p]),
    "cuSenCompute2DXpFun": (None, [_H, c_bool]),
    "checkError": (None, [c_char_p]),
    "custenGetLastError": (c_int, []),
    "custenGetLastErrorString": (c_char_p, []),
    "custenClearError": (None, []),
    "custenVersion": (c_char_p, []),
    "custenHandleSize": (c_size_t, []),
    "custenGetBuiltinFunction": (c_void_p, [c_char_p, c_int]),
    "custenSetForceIndirect": (None, [_H, c_int]),
    "custenCreate2DF32": (None, [_H, c_int, c_int, c_int, c_int, c_int, c_int, _P, _P, _P,
                                 c_int, c_int, c_int, c_int, c_int, c_int]),
    "custenCompute2DF32": (None, [_H, c_bool]),
    "custenSwap2DF32": (None, [_H, _P]),
    "custenDestroy2DF32": (None, [_H]),
    "custenSetStream": (None, [_H, c_void_p]),
    "custenSetTileConfig": (None, [_H, c_int, c_int]),
    "custenSetTileConfigEx": (None, [_H, c_int, c_int, c_int]),
    "custenGetTileConfig": (c_int, [_H, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                    ctypes.POINTER(c_int)]),
    "custenSetBandConfig": (None, [_H, c_int, c_int]),
    "custenGetBandMode": (c_int, [_H]),
    "custenBatchCreate": (c_void_p, [c_int, c_int, _P, _P, _P, _P, _P, c_int, c_int, c_int,
                                     c_int, c_int]),  # noqa: E501
    "custenBatchRun": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "custenBatchParity": (c_int, [c_void_p]),
    "custenBatchSetTileConfig": (c_int, [c_void_p, c_int, c_int, c_int]),
    "custenBatchGetTileConfig": (c_int, [c_void_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                         ctypes.POINTER(c_int)]),
    "custenBatchDestroy": (None, [c_void_p]),
    "custenFillUniform": (None, [_P, c_size_t, c_uint64, c_double, c_double]),
    "custenCHCreate": (c_void_p, [c_int, ctypes.POINTER(custenCHParams), _P]),
    "custenCHRun": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "custenCHGetField": (c_int, [c_void_p, _P, _P]),
    "custenCHDestroy": (None, [c_void_p]),
    "custenCHServerCreate": (c_void_p, [c_int, c_size_t, c_int]),
    "custenCHServerSubmit": (c_int, [c_void_p, ctypes.POINTER(custenCHParams), c_int, _P]),
    "custenCHServerRunUntilIdle": (c_int, [c_void_p, c_int]),
    "custenCHServerStep": (c_int, [c_void_p]),
    "custenCHServerQuery": (c_int, [c_void_p, c_int, ctypes.POINTER(c_int),
                                    ctypes.POINTER(c_double), ctypes.POINTER(c_double)]),
    "custenCHServerResult": (c_int, [c_void_p, c_int, _P]),
    "custenCHFootprint": (c_size_t, [c_int]),
    "custenCHStepsDone": (c_int, [c_void_p]),
    "custenCHSnapshot": (c_int, [c_void_p, c_char_p]),
    "custenCHCheckpoint": (c_int, [c_void_p, c_char_p]),
    "custenCHRestore": (c_void_p, [c_int, c_char_p]),
    "custenCHServerSubmitResume": (c_int, [c_void_p, ctypes.POINTER(custenCHParams), c_int, _P,
                                           _P]),
    "custenCHServerDestroy": (None, [c_void_p]),
    "custenCHAdvanceNoAlloc": (c_int, [c_int, ctypes.POINTER(custenCHParams), _P, _P, c_int,
                                       c_void_p]),
})

_lib = None


def lib() -> ctypes.CDLL:
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"custen native library not found at {LIB_PATH}; build it with "
            "`python -m custen_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def header_symbols() -> list[str]:
    """Every function name declared in include/*.h (aliases expanded)."""
    names = set()
    for fname in sorted(os.listdir(INCLUDE_DIR)):
        if not fname.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE_DIR, fname)).read()
        body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"^\s*(?:void|int|size_t|double\*|const char\*|custen\w+\*)\s+"
                                r"(\w+)\s*\(", body, flags=re.M))
        alias_list = re.search(r"#define CUSTEN_ALIAS_LIST\(X\)(.*?)\n\n", text, flags=re.S)
        if alias_list:
            for suffix in re.findall(r"X\((\w+)\)", alias_list.group(1)):
                names.add("custen" + suffix)
    return sorted(names)


def _addr(x) -> int | None:
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):  # torch tensor
        return x.data_ptr()
    if hasattr(x, "ctypes"):  # numpy array
        return x.ctypes.data
    if isinstance(x, ctypes._Pointer) or isinstance(x, c_void_p):
        return ctypes.cast(x, c_void_p).value
    raise TypeError(f"cannot take the address of {type(x)!r}")


def check(what: str = "") -> None:
    L = lib()
    code = L.custenGetLastError()
    if code != 0:
        msg = L.custenGetLastErrorString().decode()
        L.custenClearError()
        raise CustenError(code, f"{what}: {msg}" if what else msg)


def clear_error() -> None:
    lib().custenClearError()


def _handle(h: cuSten_t):
    return ctypes.byref(h)


def _make_create(name: str, nptr: int):
    """Create wrapper: the first three pointer arguments after BLOCK_Y are data
    pointers; Fun variants end with the function pointer."""

    def create(h: cuSten_t, *args):
        args = list(args)
        # positions (after the handle) of pointer arguments: 6, 7, 8 (out, in, weights/coe)
        for i in (6, 7, 8):
            args[i] = _addr(args[i])
        if name.endswith("Fun"):
            args[-1] = _addr(args[-1])
        getattr(lib(), name)(_handle(h), *args)
        check(name)

    create.__name__ = name
    return create


def _make_compute(name):
    def compute(h: cuSten_t, offload: int = DEVICE):
        getattr(lib(), name)(_handle(h), bool(offload))
        check(name)

    compute.__name__ = name
    return compute


def _make_swap(name):
    def swap(h: cuSten_t, dataInput):
        getattr(lib(), name)(_handle(h), _addr(dataInput))
        check(name)

    swap.__name__ = name
    return swap


def _make_destroy(name):
    def destroy(h: cuSten_t):
        getattr(lib(), name)(_handle(h))
        check(name)

    destroy.__name__ = name
    return destroy


for _v in _VARIANTS:
    for _pre in ("cuSten", "custen"):
        globals()[f"{_pre}Create{_v}"] = _make_create(f"{_pre}Create{_v}", 3)
        globals()[f"{_pre}Compute{_v}"] = _make_compute(f"{_pre}Compute{_v}")
        globals()[f"{_pre}Swap{_v}"] = _make_swap(f"{_pre}Swap{_v}")
        globals()[f"{_pre}Destroy{_v}"] = _make_destroy(f"{_pre}Destroy{_v}")
cuSenCompute2DXpFun = _make_compute("cuSenCompute2DXpFun")


def _make_weno_create(name):
    def create(h: cuSten_t, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y, dx, dy, u, v,
               dataOutput, dataInput):
        getattr(lib(), name)(_handle(h), deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
                             float(dx), float(dy), _addr(u), _addr(v), _addr(dataOutput),
                             _addr(dataInput))
        check(name)

    create.__name__ = name
    return create


for _pre in ("cuSten", "custen"):
    globals()[f"{_pre}Create2DXYWENOADVp"] = _make_weno_create(f"{_pre}Create2DXYWENOADVp")
    globals()[f"{_pre}Compute2DXYWENOADVp"] = _make_compute(f"{_pre}Compute2DXYWENOADVp")
    globals()[f"{_pre}Swap2DXYWENOADVp"] = _make_swap(f"{_pre}Swap2DXYWENOADVp")
    globals()[f"{_pre}Destroy2DXYWENOADVp"] = _make_destroy(f"{_pre}Destroy2DXYWENOADVp")


def _stream(stream):
    if stream is None:
        return None
    return stream if isinstance(stream, int) else stream.cuda_stream


def custenPentFactorBatch(ds, dl, d, du, dw, m: int, batchCount: int, stream=None) -> None:
    """In-place factorisation of interleaved pentadiagonal systems (device arrays)."""
    lib().custenPentFactorBatch(_addr(ds), _addr(dl), _addr(d), _addr(du), _addr(dw), m,
                                batchCount, _stream(stream))
    check("custenPentFactorBatch")


def custenPentSolveBatch(ds, dl, d, du, dw, b, m: int, batchCount: int, stream=None) -> None:
    """Solve interleaved pentadiagonal systems in place of b (device arrays)."""
    lib().custenPentSolveBatch(_addr(ds), _addr(dl), _addr(d), _addr(du), _addr(dw), _addr(b),
                               m, batchCount, _stream(stream))
    check("custenPentSolveBatch")


def custenSetWenoSquares(h: cuSten_t, singlePrecision: bool) -> None:
    lib().custenSetWenoSquares(_handle(h), int(bool(singlePrecision)))
    check("custenSetWenoSquares")


def custenCreate2DF32(h, deviceNum, numTiles, nx, ny, dir_, periodic, dataOutput, dataInput,
                      weights, numStenHoriz, numStenLeft, numStenRight, numStenVert, numStenTop,
                      numStenBottom):
    lib().custenCreate2DF32(_handle(h), deviceNum, numTiles, nx, ny, dir_, periodic,
                            _addr(dataOutput), _addr(dataInput), _addr(weights), numStenHoriz,
                            numStenLeft, numStenRight, numStenVert, numStenTop, numStenBottom)
    check("custenCreate2DF32")


custenCompute2DF32 = _make_compute("custenCompute2DF32")
custenSwap2DF32 = _make_swap("custenSwap2DF32")
custenDestroy2DF32 = _make_destroy("custenDestroy2DF32")


def custenGetBuiltinFunction(name: str, deviceNum: int = 0) -> int:
    addr = lib().custenGetBuiltinFunction(name.encode(), deviceNum)
    check("custenGetBuiltinFunction")
    return addr


def custenSetForceIndirect(h: cuSten_t, force: bool) -> None:
    lib().custenSetForceIndirect(_handle(h), int(bool(force)))
    check("custenSetForceIndirect")


def custenSetTileConfig(h: cuSten_t, rowsPerWave: int, chunks: int, colsPerLane: int = 0) -> None:
    lib().custenSetTileConfigEx(_handle(h), rowsPerWave, chunks, colsPerLane)
    check("custenSetTileConfigEx")


def custenGetTileConfig(h: cuSten_t):
    r, c, l = c_int(), c_int(), c_int()
    lib().custenGetTileConfig(_handle(h), ctypes.byref(r), ctypes.byref(c), ctypes.byref(l))
    check("custenGetTileConfig")
    return r.value, c.value, l.value


def custenSetBandConfig(h: cuSten_t, rowsPerWave: int, chunksPerBand: int) -> None:
    lib().custenSetBandConfig(_handle(h), rowsPerWave, chunksPerBand)
    check("custenSetBandConfig")


def custenGetBandMode(h: cuSten_t) -> int:
    r = lib().custenGetBandMode(_handle(h))
    check("custenGetBandMode")
    return r


def custenSetStream(h: cuSten_t, stream: int) -> None:
    lib().custenSetStream(_handle(h), stream)
    check("custenSetStream")


def custenFillUniform(out, n: int, seed: int, a: float, b: float) -> None:
    lib().custenFillUniform(_addr(out), n, seed, a, b)


# ---------------------------------------------------------------------------
# Cahn-Hilliard ADI (include/custen_ch.h)
# ---------------------------------------------------------------------------
def ch_params(n: int, lx: float = 2 * 3.141592653589793, D: float = 1.0, gamma: float = 0.01,
              dt_over_dx: float = 0.1) -> custenCHParams:
    import math

    if lx == 2 * 3.141592653589793:
        lx = 2 * math.pi
    return custenCHParams(n, lx, D, gamma, dt_over_dx)


class CHSolver:
    """One resident Cahn-Hilliard simulation (custenCH*)."""

    def __init__(self, params: custenCHParams, c0, device: int = 0):
        self.n = params.n
        self._p = params
        self.h = lib().custenCHCreate(device, ctypes.byref(self._p), _addr(c0))
        if not self.h:
            check("custenCHCreate")
            raise CustenError(2, "custenCHCreate failed")

    @classmethod
    def restore(cls, path: str, device: int = 0) -> "CHSolver":
        """A solver continuing from a checkpoint file (custenCHRestore)."""
        h = lib().custenCHRestore(device, os.fsencode(path))
        if not h:
            check("custenCHRestore")
            raise CustenError(1, "custenCHRestore failed")
        self = cls.__new__(cls)
        self.h = h
        self._p = None
        # the grid size is the checkpoint's (header bytes 12..15, u32 n)
        with open(path, "rb") as f:
            head = f.read(16)
        self.n = int.from_bytes(head[12:16], "little")
        return self

    @property
    def steps_done(self) -> int:
        return lib().custenCHStepsDone(self.h)

    def checkpoint(self, path: str) -> None:
        if lib().custenCHCheckpoint(self.h, os.fsencode(path)) != 0:
            check("custenCHCheckpoint")

    def snapshot(self, path: str) -> None:
        """Write the current field c as a .npy file (custenCHSnapshot)."""
        if lib().custenCHSnapshot(self.h, os.fsencode(path)) != 0:
            check("custenCHSnapshot")

    def run(self, steps: int, mode: int = 0, stream: int = 0) -> None:
        rc = lib().custenCHRun(self.h, steps, mode, stream)
        if rc != 0:
            check("custenCHRun")

    def field(self):
        import numpy as np

        c = np.empty((self.n, self.n))
        co = np.empty((self.n, self.n))
        rc = lib().custenCHGetField(self.h, _addr(c), _addr(co))
        if rc != 0:
            check("custenCHGetField")
        return c, co

    def close(self) -> None:
        if self.h:
            lib().custenCHDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ch_advance_no_alloc(params: custenCHParams, c, c_old, steps: int, stream: int = 0,
                        device: int = 0) -> None:
    """custenCHAdvanceNoAlloc: advance the simulation held in the device
    arrays c (current) and c_old (previous) by `steps` fused steps without any
    device allocation (n <= 1024)."""
    rc = lib().custenCHAdvanceNoAlloc(device, ctypes.byref(params), _addr(c), _addr(c_old), steps,
                                      stream)
    if rc != 0:
        check("custenCHAdvanceNoAlloc")
        raise CustenError(rc, "custenCHAdvanceNoAlloc failed")


class CHServer:
    """Continuous-batching request server (custenCHServer*)."""

    def __init__(self, budget_bytes: int, max_slots: int, device: int = 0):
        self.h = lib().custenCHServerCreate(device, budget_bytes, max_slots)
        if not self.h:
            check("custenCHServerCreate")
        self._n = {}

    def submit(self, params: custenCHParams, steps: int, c0, c_old=None) -> int:
        """Queue a request; with c_old, resume from the time levels (c0, c_old)."""
        if c_old is None:
            rid = lib().custenCHServerSubmit(self.h, ctypes.byref(params), steps, _addr(c0))
        else:
            rid = lib().custenCHServerSubmitResume(self.h, ctypes.byref(params), steps,
                                                   _addr(c0), _addr(c_old))
        if rid < 0:
            check("custenCHServerSubmit")
        self._n[rid] = params.n
        return rid

    def step(self) -> int:
        left = lib().custenCHServerStep(self.h)
        if left < 0:
            check("custenCHServerStep")
        return left

    def run_until_idle(self, max_iterations: int = 0) -> int:
        it = lib().custenCHServerRunUntilIdle(self.h, max_iterations)
        if it < 0:
            check("custenCHServerRunUntilIdle")
        return it

    def query(self, rid: int):
        st = c_int()
        ls = c_double()
        la = c_double()
        if lib().custenCHServerQuery(self.h, rid, ctypes.byref(st), ctypes.byref(ls),
                                     ctypes.byref(la)) != 0:
            check("custenCHServerQuery")
        return st.value, ls.value, la.value

    def result(self, rid: int):
        import numpy as np

        n = self._n[rid]
        out = np.empty((n, n))
        if lib().custenCHServerResult(self.h, rid, _addr(out)) != 0:
            check("custenCHServerResult")
        return out

    def close(self) -> None:
        if self.h:
            lib().custenCHServerDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ch_footprint(n: int) -> int:
    return lib().custenCHFootprint(n)


# ---------------------------------------------------------------------------
# Batched grids (include/custen_batch.h)
# ---------------------------------------------------------------------------
class Batch:
    """Many periodic xy grids advanced together (custenBatch*)."""

    def __init__(self, outs, ins, weights, H=5, L=2, V=5, T=2, f32=False, device=0):
        n = len(ins)
        self._keep = (outs, ins, weights)
        nx = (c_int * n)(*[int(t.shape[1]) for t in ins])
        ny = (c_int * n)(*[int(t.shape[0]) for t in ins])
        po = (c_void_p * n)(*[_addr(t) for t in outs])
        pi = (c_void_p * n)(*[_addr(t) for t in ins])
        self.h = lib().custenBatchCreate(device, n, nx, ny, po, pi, _addr(weights), H, L, V, T,
                                         1 if f32 else 0)
        if not self.h:
            check("custenBatchCreate")
            raise CustenError(2, "custenBatchCreate failed")

    def run(self, steps: int, mode: int = 0, stream: int = 0) -> None:
        if lib().custenBatchRun(self.h, steps, mode, stream) != 0:
            check("custenBatchRun")

    @property
    def parity(self) -> int:
        return lib().custenBatchParity(self.h)

    def set_tile_config(self, rows_per_wave: int, chunks: int, band: bool = False) -> None:
        if lib().custenBatchSetTileConfig(self.h, rows_per_wave, chunks, 1 if band else 0) != 0:
            check("custenBatchSetTileConfig")

    def tile_config(self):
        r, c, b = c_int(), c_int(), c_int()
        if lib().custenBatchGetTileConfig(self.h, ctypes.byref(r), ctypes.byref(c),
                                          ctypes.byref(b)) != 0:
            check("custenBatchGetTileConfig")
        return r.value, c.value, b.value

    def close(self) -> None:
        if self.h:
            lib().custenBatchDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
