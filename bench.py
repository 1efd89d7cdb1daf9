"""Benchmark of the cuSten stencil path on one MI355X.

Default workload (one "step" = one stencil apply over the whole grid):
  2D xy-periodic 13-point biharmonic (5x5 weights, 12 of them zero), FP64,
  8192 x 8192, inputs U[-1,1) from SplitMix64 seed 2, resident in HBM.

Prints ONE JSON line: the metric in GPts/s, the roofline of the stencil kernel
(algorithmic 16 B/point over the kernel's average launch time, measured with
HIP events on the kernel's own stream), the CPU oracle baseline on a bounded
sample of the same workload, and secondary figures under "extra".

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload stencil|batch|ch]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "stencil GPts/s + % HBM peak; batched-vs-plain CH solves/s, p50/p99 latency"
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


def biharmonic_weights(n: int) -> np.ndarray:
    h = 2 * np.pi / n
    pat = np.array([[0, 0, 1, 0, 0], [0, 2, -8, 2, 0], [1, -8, 20, -8, 1],
                    [0, 2, -8, 2, 0], [0, 0, 1, 0, 0]], dtype=np.float64)
    return (pat / h ** 4).ravel()


def laplacian9_weights(n: int) -> np.ndarray:
    h = 2 * np.pi / n
    return (np.array([1, 4, 1, 4, -20, 4, 1, 4, 1], dtype=np.float64) / (6 * h * h))


class Timer:
    """HIP-event timing on a given torch stream."""

    def __init__(self, torch, stream):
        self.torch = torch
        self.stream = stream

    def run(self, fn, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(self.stream)
        for _ in range(steps):
            fn()
        e1.record(self.stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps  # ms per step


def make_handle(api, torch, stream, kind, n, dout, din, w_or_coe, func=None, f32=False):
    h = api.cuSten_t()
    if f32:
        api.custenCreate2DF32(h, 0, 1, n, n, 2, 1 if kind.endswith("p") else 0, dout, din,
                              w_or_coe, 5, 2, 2, 5, 2, 2)
    elif kind == "xy_p_bih":
        api.cuStenCreate2DXYp(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe, 5, 2, 2, 5, 2, 2)
    elif kind == "xy_np_bih":
        api.cuStenCreate2DXYnp(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe, 5, 2, 2, 5, 2, 2)
    elif kind == "xy_p_lap":
        api.cuStenCreate2DXYp(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe, 3, 1, 1, 3, 1, 1)
    elif kind == "xy_np_lap":
        api.cuStenCreate2DXYnp(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe, 3, 1, 1, 3, 1, 1)
    elif kind == "xy_p_fun":
        api.cuStenCreate2DXYpFun(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe, 3, 1, 1, 3, 1, 1,
                                 func)
    elif kind == "xy_np_fun":
        api.cuStenCreate2DXYnpFun(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe, 3, 1, 1, 3, 1, 1,
                                  func)
    else:
        raise ValueError(kind)
    api.custenSetStream(h, stream.cuda_stream)
    return h


def compute_fn(api, kind, h, f32=False):
    if f32:
        return lambda: api.custenCompute2DF32(h, api.DEVICE)
    name = {"xy_p_bih": "2DXYp", "xy_np_bih": "2DXYnp", "xy_p_lap": "2DXYp",
            "xy_np_lap": "2DXYnp", "xy_p_fun": "2DXYpFun", "xy_np_fun": "2DXYnpFun"}[kind]
    fn = getattr(api, f"cuStenCompute{name}")
    return lambda: fn(h, api.DEVICE)


def destroy(api, kind, h, f32=False):
    if f32:
        api.custenDestroy2DF32(h)
        return
    name = {"xy_p_bih": "2DXYp", "xy_np_bih": "2DXYnp", "xy_p_lap": "2DXYp",
            "xy_np_lap": "2DXYnp", "xy_p_fun": "2DXYpFun", "xy_np_fun": "2DXYnpFun"}[kind]
    getattr(api, f"cuStenDestroy{name}")(h)


def cpu_baseline(n: int, w: np.ndarray, seconds: float = 10.0):
    """Oracle (fma chain, OpenMP over rows) on 8192 x 1024 slabs of the same
    workload, repeated for about `seconds`."""
    import oracle
    from custen_amd import synth

    threads = min(16, os.cpu_count() or 1)
    rows = 1024
    slab = synth.uniform(n * rows, 2, -1.0, 1.0).reshape(rows, n)
    out = np.empty_like(slab)
    oracle.apply(slab, w, 5, 2, 5, 2, True, out=out, nthreads=threads)  # warm
    reps = 0
    t0 = time.perf_counter()
    while True:
        oracle.apply(slab, w, 5, 2, 5, 2, True, out=out, nthreads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    pts = reps * n * rows
    return {
        "value": pts / el / 1e9,
        "unit": "GPts/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{reps} applies of an {n}x{rows} periodic slab of the same FP64 5x5 "
                  f"biharmonic workload (oracle fma chain, {threads} OpenMP threads, "
                  f"{el:.1f} s)",
    }


def bench_stencil(args):
    import torch

    from custen_amd import api, synth

    n = 8192
    stream = torch.cuda.current_stream()
    timer = Timer(torch, stream)
    host = np.empty((n, n))
    api.custenFillUniform(host, n * n, 2, -1.0, 1.0)
    din = torch.from_numpy(host).cuda()
    dout = torch.empty_like(din)
    wb = torch.from_numpy(biharmonic_weights(n)).cuda()
    pts = n * n

    h = make_handle(api, torch, stream, "xy_p_bih", n, dout, din, wb)
    ms = timer.run(compute_fn(api, "xy_p_bih", h), args.steps, args.warmup)
    destroy(api, "xy_p_bih", h)
    gpts = pts / (ms * 1e-3) / 1e9
    achieved = 16.0 * pts / (ms * 1e-3) / 1e9

    extra = {}
    # secondary stencils on the same grid (fewer steps)
    k2 = max(5, args.steps // 4)
    wl = torch.from_numpy(laplacian9_weights(n)).cuda()
    coe = torch.from_numpy(laplacian9_weights(n)).cuda()
    fptr = api.custenGetBuiltinFunction("xy_cube_minus", 0)
    for kind, w_ in (("xy_np_bih", wb), ("xy_p_lap", wl), ("xy_np_lap", wl),
                     ("xy_p_fun", coe), ("xy_np_fun", coe)):
        h = make_handle(api, torch, stream, kind, n, dout, din, w_,
                        func=fptr if "fun" in kind else None)
        t = timer.run(compute_fn(api, kind, h), k2, 2)
        extra[kind] = {"ms": round(t, 4), "GPts/s": round(pts / (t * 1e-3) / 1e9, 2),
                       "GB/s": round(16.0 * pts / (t * 1e-3) / 1e9, 1)}
        if "fun" in kind:
            api.custenSetForceIndirect(h, True)
            t = timer.run(compute_fn(api, kind, h), 3, 1)
            extra[kind + "_indirect_call"] = {"ms": round(t, 4),
                                              "GPts/s": round(pts / (t * 1e-3) / 1e9, 2)}
        destroy(api, kind, h)

    # FP32 storage
    din32 = din.float()
    dout32 = torch.empty_like(din32)
    h = make_handle(api, torch, stream, "xy_p_bih", n, dout32, din32, wb, f32=True)
    t = timer.run(compute_fn(api, "xy_p_bih", h, f32=True), k2, 2)
    destroy(api, "xy_p_bih", h, f32=True)
    rows = 256
    import oracle
    ref = oracle.apply(host[:rows + 4].copy(), biharmonic_weights(n), 5, 2, 5, 2, False)
    got = dout32[:rows + 4].double().cpu().numpy()
    inner = (slice(2, rows + 2), slice(2, n - 2))
    err_abs = float(np.max(np.abs(got[inner] - ref[inner])))
    scale = float(np.max(np.abs(ref[inner])))
    extra["xy_p_bih_f32"] = {"ms": round(t, 4), "GPts/s": round(pts / (t * 1e-3) / 1e9, 2),
                             "GB/s": round(8.0 * pts / (t * 1e-3) / 1e9, 1),
                             "max_abs_err_vs_fp64_oracle": err_abs,
                             "max_rel_err_vs_fp64_oracle": err_abs / scale,
                             "err_sample_rows": rows}

    # device copy ceiling (read + write of the same 512 MiB)
    tc = timer.run(lambda: dout.copy_(din), k2, 2)
    extra["copy_ceiling_GBps"] = round(2 * 8.0 * pts / (tc * 1e-3) / 1e9, 1)

    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        try:
            with open(PROFILE_TRAFFIC) as f:
                traffic = json.load(f).get("xy_p_bih_8192", {}).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    result = {
        "metric": METRIC,
        "value": round(gpts, 3),
        "unit": "GPts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 U[-1,1), seed 2)",
        "config": {"workload": "xy-periodic 13-point biharmonic (5x5 weights) FP64 stencil "
                               "apply, 8192x8192, device-resident",
                   "nx": n, "ny": n, "stencil": "5x5", "bytes_per_point": 16},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu_baseline(n, biharmonic_weights(n), seconds=args.cpu_seconds),
        "extra": extra,
    }
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="stencil", choices=["stencil"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the stencil path has no multi-GPU code "
              "(sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    res = bench_stencil(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
