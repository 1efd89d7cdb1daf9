"""ctypes wrapper of the CPU oracle (oracle/custen_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product package
custen_amd never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcusten_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i = ctypes.c_int
_l = ctypes.c_long

FUN_X_CENTRAL_DIFF = 1
FUN_Y_WSUM = 2
FUN_XY_WSUM = 3
FUN_XY_CUBE_MINUS = 4
FUN_X_WSUM = 5


class ChState(ctypes.Structure):
    _fields_ = [
        ("n", _i), ("m", _i),
        ("dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("sigma_lin", ctypes.c_double), ("sigma_nl", ctypes.c_double),
        ("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double),
        ("d", ctypes.c_double), ("e", ctypes.c_double),
        ("omega", ctypes.c_double * 4),
        ("wlin", ctypes.c_double * 25),
        ("wnl", ctypes.c_double * 9),
        ("ds", _dp), ("dl", _dp), ("diag", _dp), ("du", _dp), ("dw", _dp),
        ("inv1", _dp), ("inv2", _dp),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = ctypes.CDLL(LIB_PATH)
        L.cso_apply_f64.argtypes = [_dp, _dp, _l, _l, _dp, _i, _i, _i, _i, _i, _i, _i]
        L.cso_apply_f32.argtypes = [_fp, _fp, _l, _l, _fp, _i, _i, _i, _i, _i, _i, _i]
        L.cso_apply_fun_f64.argtypes = [_i, _i, _dp, _dp, _l, _l, _dp, _i, _i, _i, _i, _i, _i, _i]
        L.cso_ch_init.argtypes = [ctypes.POINTER(ChState), _i, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double]
        L.cso_ch_init.restype = _i
        L.cso_ch_free.argtypes = [ctypes.POINTER(ChState)]
        L.cso_ch_step.argtypes = [ctypes.POINTER(ChState), _dp, _dp, _dp]
        L.cso_ch_cyclic_solve.argtypes = [ctypes.POINTER(ChState), _dp, _l]
        L.cso_pent_factor.argtypes = [_dp, _dp, _dp, _dp, _dp, _i]
        L.cso_pent_solve.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _i]
        L.cso_fill_uniform.argtypes = [_dp, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_double,
                                       ctypes.c_double]
        L.cso_pent_factor_batch.argtypes = [_dp, _dp, _dp, _dp, _dp, _i, _l, _i]
        L.cso_pent_solve_batch.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _i, _l, _i]
        L.cso_weno_f64.argtypes = [_dp, _dp, _dp, _dp, _l, _l, ctypes.c_double, ctypes.c_double,
                                   _i, _i]
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def apply(inp, w, H, L, V, T, periodic, zero_right_edge=False, out=None, nthreads=1):
    """Weight stencil on a (ny, nx) array; float64 or float32 storage."""
    inp = np.ascontiguousarray(inp)
    ny, nx = inp.shape
    if out is None:
        out = np.zeros_like(inp)
    if inp.dtype == np.float64:
        w = np.ascontiguousarray(w, dtype=np.float64)
        lib().cso_apply_f64(_d(inp), _d(out), nx, ny, _d(w), H, L, V, T, int(periodic),
                            int(zero_right_edge), nthreads)
    else:
        w = np.ascontiguousarray(w, dtype=np.float32)
        lib().cso_apply_f32(_f(inp), _f(out), nx, ny, _f(w), H, L, V, T, int(periodic),
                            int(zero_right_edge), nthreads)
    return out


def apply_fun(fun_id, direction, inp, coe, H, L, V, T, periodic, out=None, nthreads=1):
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    ny, nx = inp.shape
    if out is None:
        out = np.zeros_like(inp)
    coe = np.ascontiguousarray(coe, dtype=np.float64)
    lib().cso_apply_fun_f64(fun_id, direction, _d(inp), _d(out), nx, ny, _d(coe), coe.size, H, L,
                            V, T, int(periodic), nthreads)
    return out


def pent_factor_batch(ds, dl, d, du, dw, nthreads=1):
    """Factorisation of interleaved (m, batch) diagonal arrays
    (cuPentBatch.cu:35-113); returns the factored copies."""
    arrs = [np.array(a, dtype=np.float64, copy=True, order="C") for a in (ds, dl, d, du, dw)]
    m, batch = arrs[0].shape
    lib().cso_pent_factor_batch(*[_d(a) for a in arrs], m, batch, nthreads)
    return arrs


def pent_solve_batch(ds, dl, d, du, dw, b, nthreads=1):
    """Solve interleaved (m, batch) systems with factored diagonals
    (cuPentBatch.cu:119-198); returns the solution array."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (ds, dl, d, du, dw)]
    x = np.array(b, dtype=np.float64, copy=True, order="C")
    m, batch = x.shape
    lib().cso_pent_solve_batch(*[_d(a) for a in arrs], _d(x), m, batch, nthreads)
    return x


def weno(inp, u, v, dx, dy, single_squares=True, out=None, nthreads=1):
    """WENO5 upwind advection u * d/dx + v * d/dy of a periodic (ny, nx) field
    (2d_xyADVWENO_p_kernel.cu:51-401).  coeDx = 1/dx and coeDy = 1/dy, as the
    reference's Create computes them."""
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    ny, nx = inp.shape
    assert u.shape == inp.shape and v.shape == inp.shape
    if out is None:
        out = np.zeros_like(inp)
    lib().cso_weno_f64(_d(inp), _d(u), _d(v), _d(out), nx, ny, 1.0 / dx, 1.0 / dy,
                       int(single_squares), nthreads)
    return out


class CH:
    """Cahn-Hilliard ADI oracle for one n x n grid."""

    def __init__(self, n, lx=2 * np.pi, D=1.0, gamma=0.01, dt_over_dx=0.1):
        self.st = ChState()
        rc = lib().cso_ch_init(ctypes.byref(self.st), n, lx, D, gamma, dt_over_dx)
        if rc != 0:
            raise ValueError(f"cso_ch_init failed ({rc})")
        self.n = n
        self.work = np.zeros(3 * n * n)

    def __del__(self):
        try:
            lib().cso_ch_free(ctypes.byref(self.st))
        except Exception:
            pass

    def arrays(self):
        m = self.st.m
        get = lambda p: np.ctypeslib.as_array(p, shape=(m,)).copy()
        return {k: get(getattr(self.st, k)) for k in ("ds", "dl", "diag", "du", "dw", "inv1", "inv2")}

    def step(self, c, cold, steps=1):
        for _ in range(steps):
            lib().cso_ch_step(ctypes.byref(self.st), _d(c), _d(cold), _d(self.work))

    def cyclic_solve(self, x, stride=1):
        lib().cso_ch_cyclic_solve(ctypes.byref(self.st), _d(x), stride)


def pent_factor_solve(ds, dl, d, du, dw, b):
    ds, dl, d, du, dw = (np.array(v, dtype=np.float64) for v in (ds, dl, d, du, dw))
    b = np.array(b, dtype=np.float64)
    m = d.size
    lib().cso_pent_factor(_d(ds), _d(dl), _d(d), _d(du), _d(dw), m)
    lib().cso_pent_solve(_d(ds), _d(dl), _d(d), _d(du), _d(dw), _d(b), m)
    return b


def fill_uniform(n, seed, a, b):
    out = np.empty(n)
    lib().cso_fill_uniform(_d(out), n, seed, a, b)
    return out
