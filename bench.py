"""Benchmark of the cuSten stencil path on one MI355X.

Headline (one "step" = one stencil apply over the whole grid):
  2D xy-periodic 13-point biharmonic (5x5 weights, 12 of them zero), FP64,
  8192 x 8192, inputs U[-1,1) from SplitMix64 seed 2, resident in HBM.

Prints ONE JSON line: the metric in GPts/s, the roofline of the stencil kernel
(algorithmic 16 B/point over the kernel's average launch time, measured with
HIP events on the stream the kernel is launched on), the CPU oracle baseline on
a bounded sample of the same workload, and under "extra" the secondary
figures: other stencils of the 8192^2 configuration (np, 9-point Laplacian,
point function, FP32 storage with its error against the FP64 oracle), the
1024^2 x-periodic configuration, WENO5 advection at 8192^2, batched uneven
grids against the per-grid loop, batched pentadiagonal factor and solve, and
the served Cahn-Hilliard stream against the plain per-request loop.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
  python bench.py --workload stream    # 65536 x 32768 host-resident tiled streaming alone
                                       # (also part of the default run's "extra")
  python bench.py --workload weno|pent|batch|ch|x1024|fun   # one secondary benchmark alone
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Timer:
    """HIP-event timing on a given torch stream."""

    def __init__(self, torch, stream):
        self.torch = torch
        self.stream = stream

    def run(self, fn, steps, warmup, prewarm_s=0.0):
        """ms per step of `steps` calls of fn, after `warmup` untimed calls.
        prewarm_s: first call fn for about that many seconds (untimed), so
        that the GPU clocks are back up after a host-side pause."""
        torch = self.torch
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < prewarm_s:
            for _ in range(20):
                fn()
            torch.cuda.synchronize()
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


_NAMES = {"xy_p_bih": "2DXYp", "xy_np_bih": "2DXYnp", "xy_p_lap": "2DXYp",
          "xy_np_lap": "2DXYnp", "xy_p_fun": "2DXYpFun", "xy_np_fun": "2DXYnpFun"}


def make_handle(api, stream, kind, n, dout, din, w_or_coe, func=None, f32=False):
    h = api.cuSten_t()
    if f32:
        api.custenCreate2DF32(h, 0, 1, n, n, 2, 1, dout, din, w_or_coe, 5, 2, 2, 5, 2, 2)
    elif kind in ("xy_p_bih", "xy_np_bih"):
        getattr(api, f"cuStenCreate{_NAMES[kind]}")(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe,
                                                     5, 2, 2, 5, 2, 2)
    elif kind in ("xy_p_lap", "xy_np_lap"):
        getattr(api, f"cuStenCreate{_NAMES[kind]}")(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe,
                                                     3, 1, 1, 3, 1, 1)
    else:
        getattr(api, f"cuStenCreate{_NAMES[kind]}")(h, 0, 1, n, n, 32, 32, dout, din, w_or_coe,
                                                     3, 1, 1, 3, 1, 1, func)
    api.custenSetStream(h, stream.cuda_stream)
    return h


def compute_fn(api, kind, h, f32=False):
    if f32:
        return lambda: api.custenCompute2DF32(h, api.DEVICE)
    fn = getattr(api, f"cuStenCompute{_NAMES[kind]}")
    return lambda: fn(h, api.DEVICE)


def timestep_fn(api, kind, h, a, b, f32=False):
    """One time step: Compute, then Swap so that the output becomes the next
    input (fields a and b alternate; a is the input when the loop starts)."""
    comp = compute_fn(api, kind, h, f32)
    swap = api.custenSwap2DF32 if f32 else getattr(api, f"cuStenSwap{_NAMES[kind]}")
    state = {"in": a}

    def step():
        comp()
        nxt = b if state["in"] is a else a
        swap(h, nxt)
        state["in"] = nxt
    return step


def destroy(api, kind, h, f32=False):
    if f32:
        api.custenDestroy2DF32(h)
    else:
        getattr(api, f"cuStenDestroy{_NAMES[kind]}")(h)


def cpu_baseline(n: int, w: np.ndarray, seconds: float, threads: int = 0):
    """Oracle (fma chain, OpenMP over rows) on 8192 x 1024 slabs of the same
    workload, repeated for about `seconds`; threads 0 = up to 16 host cores."""
    import oracle
    from custen_amd import synth

    threads = threads or min(16, os.cpu_count() or 1)
    rows = 1024 if threads > 1 else 128
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
    return {
        "value": round(reps * n * rows / el / 1e9, 4),
        "unit": "GPts/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{reps} applies of an {n}x{rows} periodic slab of the same FP64 5x5 "
                  f"biharmonic workload (oracle fma chain, {threads} OpenMP threads, "
                  f"{el:.1f} s)",
    }


# ---------------------------------------------------------------------------
def bench_stencil_8192(args, torch, api, extra):
    from custen_amd import workloads
    import oracle

    n = 8192
    pts = n * n
    stream = torch.cuda.current_stream()
    timer = Timer(torch, stream)
    host = np.empty((n, n))
    api.custenFillUniform(host, pts, 2, -1.0, 1.0)
    din = torch.from_numpy(host).cuda()
    dout = torch.empty_like(din)
    wbh = workloads.biharmonic_weights(n)
    wb = torch.from_numpy(wbh).cuda()

    h = make_handle(api, stream, "xy_p_bih", n, dout, din, wb)
    if args.tile.startswith("band:"):
        api.custenSetBandConfig(h, *(int(v) for v in args.tile[5:].split(",")))
    elif args.tile:
        api.custenSetTileConfig(h, *(int(v) for v in args.tile.split(",")))
    api.cuStenCompute2DXYp(h, api.DEVICE)  # autotunes the tile configuration once
    torch.cuda.synchronize()
    # parity spot check of this apply against the oracle (rows 0..255)
    ref = oracle.apply(host[:260].copy(), wbh, 5, 2, 5, 2, False)
    got = dout[:260].cpu().numpy()
    inner = (slice(2, 258), slice(2, n - 2))
    extra["headline_rows_checked_vs_oracle"] = 256
    extra["headline_mismatching_points"] = int(np.count_nonzero(
        got[inner].view(np.uint64) != ref[inner].view(np.uint64)))
    # timed: a time-stepping loop, each step one apply and then a swap of the
    # two fields (each step reads the field the previous one wrote, so no
    # input is read twice in a row); untimed first: ~100 ms of steps so that
    # the GPU clocks have ramped up whether or not the autotuner ran, then the
    # W warm-up steps
    fn = timestep_fn(api, "xy_p_bih", h, din, dout)
    ms = timer.run(fn, args.steps, args.warmup, prewarm_s=0.1)
    # the same apply repeated on one unchanged input, for comparison
    same = timer.run(compute_fn(api, "xy_p_bih", h), args.steps, args.warmup, prewarm_s=0.05)
    extra["headline_same_input_us"] = round(same * 1e3, 1)
    rpw, chunks, cpl = api.custenGetTileConfig(h)
    band = api.custenGetBandMode(h)
    extra["headline_tile_config"] = {"kernel": "band walker" if band else "strip",
                                     "rows_per_wave": rpw,
                                     ("chunks_per_band" if band else "chunks_per_group"): chunks,
                                     "columns_per_lane": cpl}
    destroy(api, "xy_p_bih", h)
    din.copy_(torch.from_numpy(host))  # the variants below start from the original field
    log(f"xy_p_bih 8192^2: {ms * 1e3:.1f} us/apply")
    # torch's device-to-device copy of the same 512 MiB, for comparison (also
    # the calibration dispatch for the rocprof traffic counters)
    tc = timer.run(lambda: dout.copy_(din), max(5, args.steps // 4), 2, prewarm_s=0.05)
    extra["torch_copy_GBps"] = round(2 * 8.0 * pts / (tc * 1e-3) / 1e9, 1)
    extra["copy_ceiling"] = copy_ceiling(args, torch, api, timer, din, dout, ms)
    if not args.no_extra:
        bench_stencil_variants(args, torch, api, extra, din, dout, wb, ref)
    del din, dout
    torch.cuda.empty_cache()
    return ms


def copy_ceiling(args, torch, api, timer, din, dout, headline_ms):
    """The same time-stepping loop (apply, then swap) with a 1 x 1 stencil of
    weight 1.0: a copy of the 8192^2 field through the library's own kernels
    (the autotuner picks the configuration, in practice the band walker), so
    it moves the headline's 16 B/point with none of its arithmetic.  The
    headline's share of this rate is reported beside it."""
    n = din.shape[0]
    pts = n * n
    stream = torch.cuda.current_stream()
    w1 = torch.ones(1, dtype=torch.float64, device="cuda")
    h = api.cuSten_t()
    api.cuStenCreate2DXYp(h, 0, 1, n, n, 32, 32, dout, din, w1, 1, 0, 0, 1, 0, 0)
    api.custenSetStream(h, stream.cuda_stream)
    api.cuStenCompute2DXYp(h, api.DEVICE)  # autotunes
    torch.cuda.synchronize()
    exact = bool(torch.equal(dout, din))
    step = timestep_fn(api, "xy_p_bih", h, din, dout)
    t = timer.run(step, args.steps, args.warmup, prewarm_s=0.05)
    rpw, chunks, _ = api.custenGetTileConfig(h)
    band = api.custenGetBandMode(h)
    api.cuStenDestroy2DXYp(h)
    rate = 16.0 * pts / (t * 1e-3) / 1e9
    head = 16.0 * pts / (headline_ms * 1e-3) / 1e9
    log(f"copy ceiling (1x1 apply): {t * 1e3:.1f} us, {rate:.0f} GB/s")
    return {"us": round(t * 1e3, 1), "GBps": round(rate, 1),
            "frac_of_hbm_peak": round(rate / HBM_PEAK_GBS, 4),
            "headline_frac_of_copy_ceiling": round(head / rate, 4),
            "kernel": "band walker" if band else "strip", "rows_per_wave": rpw,
            "chunks": chunks, "output_equals_input": exact,
            "what": "xy-periodic 1x1 stencil, weight 1.0, 8192^2 FP64, apply + swap loop"}


def bench_stencil_variants(args, torch, api, extra, din, dout, wb, ref):
    """The other stencils of the 8192^2 configuration on the headline's input."""
    from custen_amd import workloads

    n = din.shape[0]
    pts = n * n
    stream = torch.cuda.current_stream()
    timer = Timer(torch, stream)
    inner = (slice(2, 258), slice(2, n - 2))
    k2 = max(20, args.steps // 2)
    wl = torch.from_numpy(workloads.laplacian9_weights(n)).cuda()
    fptr = api.custenGetBuiltinFunction("xy_cube_minus", 0)
    for kind, w_ in (("xy_np_bih", wb), ("xy_p_lap", wl), ("xy_np_lap", wl),
                     ("xy_p_fun", wl), ("xy_np_fun", wl)):
        h = make_handle(api, stream, kind, n, dout, din, w_,
                        func=fptr if "fun" in kind else None)
        fn = compute_fn(api, kind, h)
        fn()
        torch.cuda.synchronize()
        t = timer.run(fn, k2, 2, prewarm_s=0.05)
        extra[kind + "_8192"] = {"us": round(t * 1e3, 1), "GPts/s": round(pts / (t * 1e-3) / 1e9, 1),
                                 "GB/s": round(16.0 * pts / (t * 1e-3) / 1e9, 1)}
        if "fun" in kind:
            api.custenSetForceIndirect(h, True)
            t = timer.run(fn, 3, 1)
            extra[kind + "_8192_indirect_call"] = {"us": round(t * 1e3, 1),
                                                   "GPts/s": round(pts / (t * 1e-3) / 1e9, 1)}
        destroy(api, kind, h)

    # the largest built-in xy shapes, centred, periodic and not (synthetic
    # weights; the autotuner picks the configuration on the first Compute)
    from custen_amd import synth
    for H in (7, 9):
        wbig = torch.from_numpy(synth.uniform(H * H, 40 + H, -0.02, 0.02)).cuda()
        for per in ("p", "np"):
            name = "2DXYp" if per == "p" else "2DXYnp"
            h = api.cuSten_t()
            getattr(api, f"cuStenCreate{name}")(h, 0, 1, n, n, 32, 32, dout, din, wbig,
                                                 H, H // 2, H // 2, H, H // 2, H // 2)
            api.custenSetStream(h, stream.cuda_stream)
            comp = getattr(api, f"cuStenCompute{name}")
            fn = lambda: comp(h, api.DEVICE)  # noqa: E731
            fn()
            torch.cuda.synchronize()
            t = timer.run(fn, k2, 2, prewarm_s=0.05)
            extra[f"xy_{per}_{H}x{H}_8192"] = {"us": round(t * 1e3, 1),
                                               "GPts/s": round(pts / (t * 1e-3) / 1e9, 1),
                                               "GB/s": round(16.0 * pts / (t * 1e-3) / 1e9, 1)}
            getattr(api, f"cuStenDestroy{name}")(h)

    # FP32 storage, error against the FP64 oracle on the first 256 rows
    din32 = din.float()
    dout32 = torch.empty_like(din32)
    h = make_handle(api, stream, "xy_p_bih", n, dout32, din32, wb, f32=True)
    compute_fn(api, "xy_p_bih", h, f32=True)()
    torch.cuda.synchronize()
    got = dout32[:260].double().cpu().numpy()
    # time-stepping loop (apply + swap), as for the FP64 headline
    t = timer.run(timestep_fn(api, "xy_p_bih", h, din32, dout32, f32=True), k2, 2,
                  prewarm_s=0.05)
    destroy(api, "xy_p_bih", h, f32=True)
    err_abs = float(np.max(np.abs(got[inner] - ref[inner])))
    scale = float(np.max(np.abs(ref[inner])))
    extra["xy_p_bih_f32_8192"] = {"us": round(t * 1e3, 1), "GPts/s": round(pts / (t * 1e-3) / 1e9, 1),
                                  "GB/s": round(8.0 * pts / (t * 1e-3) / 1e9, 1),
                                  "max_abs_err_vs_fp64_oracle": err_abs,
                                  "max_err_rel_to_max_abs_output": err_abs / scale}
    del din32, dout32

    bench_weno_8192(torch, api, stream, timer, din, dout, k2, extra)



def bench_fun_8192(args, torch, api, extra):
    """The point-function ("Fun") variants on the 8192^2 field, xy periodic,
    3 x 3: the built-ins run inlined on the stencil kernels; through their
    function pointer (forced) they take the call path a caller's own function
    takes; and the native caller in tests/native (its own __device__ functions,
    linked -fgpu-rdc against libcusten.a) times its functions the same way."""
    import subprocess

    from custen_amd import workloads

    n = 8192
    pts = n * n
    stream = torch.cuda.current_stream()
    timer = Timer(torch, stream)
    host = np.empty((n, n))
    api.custenFillUniform(host, pts, 2, -1.0, 1.0)
    din = torch.from_numpy(host).cuda()
    dout = torch.empty_like(din)
    wl = torch.from_numpy(workloads.laplacian9_weights(n)).cuda()
    res = {}
    k = max(5, args.steps // 4)
    for name in ("xy_cube_minus", "xy_weighted_sum"):
        f = api.custenGetBuiltinFunction(name, 0)
        for indirect in (False, True):
            h = make_handle(api, stream, "xy_p_fun", n, dout, din, wl, func=f)
            api.custenSetForceIndirect(h, indirect)
            fn = compute_fn(api, "xy_p_fun", h)
            fn()
            torch.cuda.synchronize()
            t = timer.run(fn, k, 2, prewarm_s=0.05)
            destroy(api, "xy_p_fun", h)
            res[f"{name}_{'pointer_call' if indirect else 'inlined'}"] = {
                "us": round(t * 1e3, 1), "GPts/s": round(pts / (t * 1e-3) / 1e9, 1)}
    del din, dout
    torch.cuda.empty_cache()
    exe = os.path.join(ROOT, "tests", "native", "bin", "fun_caller")
    if os.path.exists(exe):
        try:
            out = subprocess.run([exe, "time", str(n), "20"], capture_output=True, text=True,
                                 timeout=300)
            res["native_caller_own_functions"] = json.loads(out.stdout.strip().splitlines()[-1])
        except (subprocess.SubprocessError, ValueError, IndexError) as e:
            res["native_caller_own_functions"] = {"error": str(e)}
    exe = os.path.join(ROOT, "tests", "native", "bin", "point_caller")
    if os.path.exists(exe):
        try:
            out = subprocess.run([exe, "time", str(n), "20"], capture_output=True, text=True,
                                 timeout=300)
            res["native_caller_functions_compiled_in"] = json.loads(
                out.stdout.strip().splitlines()[-1])
        except (subprocess.SubprocessError, ValueError, IndexError) as e:
            res["native_caller_functions_compiled_in"] = {"error": str(e)}
    extra["xy_p_fun_8192_paths"] = res
    log(f"fun 8192^2: {res}")


def bench_weno_8192(torch, api, stream, timer, din, dout, steps, extra):
    """WENO5 advection of the 8192^2 field by synthetic u, v (32 B/pt: read the
    field, u and v, write the result), both treatments of the smoothness
    squares and both tile heights; a 256-row band is checked against the
    oracle on a periodic 8192 x 256 slab of its own."""
    import oracle
    from custen_amd import synth

    n = din.shape[0]
    pts = n * n
    hu = np.empty(pts)
    hv = np.empty(pts)
    api.custenFillUniform(hu, pts, 6, -1.0, 1.0)
    api.custenFillUniform(hv, pts, 7, -1.0, 1.0)
    du = torch.from_numpy(hu.reshape(n, n)).cuda()
    dv = torch.from_numpy(hv.reshape(n, n)).cuda()
    del hu, hv
    dx = 2 * np.pi / n
    res = {}
    for single in (True, False):
        for rpw in (4, 8):
            h = api.cuSten_t()
            api.cuStenCreate2DXYWENOADVp(h, 0, 1, n, n, 32, 32, dx, dx, du, dv, dout, din)
            api.custenSetStream(h, stream.cuda_stream)
            api.custenSetWenoSquares(h, single)
            api.custenSetTileConfig(h, rpw, 1)
            fn = lambda: api.cuStenCompute2DXYWENOADVp(h, api.DEVICE)  # noqa: E731
            fn()
            torch.cuda.synchronize()
            t = timer.run(fn, steps, 2, prewarm_s=0.05)
            api.cuStenDestroy2DXYWENOADVp(h)
            res[f"{'single' if single else 'double'}_squares_rpw{rpw}"] = {
                "us": round(t * 1e3, 1), "GPts/s": round(pts / (t * 1e-3) / 1e9, 1),
                "GB/s": round(32.0 * pts / (t * 1e-3) / 1e9, 1)}
    # parity on a slab: a separate periodic 8192 x 256 problem
    rows = 256
    phi = synth.uniform(n * rows, 8, -1.0, 1.0).reshape(rows, n)
    us = du[:rows].cpu().numpy().copy()
    vs = dv[:rows].cpu().numpy().copy()
    sphi, sout = torch.from_numpy(phi).cuda(), torch.empty(rows, n, dtype=torch.float64,
                                                           device="cuda")
    h = api.cuSten_t()
    api.cuStenCreate2DXYWENOADVp(h, 0, 1, n, rows, 32, 32, dx, dx, du[:rows], dv[:rows], sout,
                                 sphi)
    api.cuStenCompute2DXYWENOADVp(h, api.DEVICE)
    torch.cuda.synchronize()
    api.cuStenDestroy2DXYWENOADVp(h)
    ref = oracle.weno(phi, us, vs, dx, dx, True, nthreads=min(16, os.cpu_count() or 1))
    res["slab_mismatching_points_vs_oracle"] = int(np.count_nonzero(
        sout.cpu().numpy().view(np.uint64) != ref.view(np.uint64)))
    extra["xy_p_weno_8192"] = res
    log(f"weno 8192^2: {res}")
    del du, dv, sphi, sout


def bench_pent(args, torch, api, extra):
    """Batched pentadiagonal factor + solve, 65536 interleaved systems of 1024
    unknowns with their own diagonals.  Algorithmic bytes per unknown: factor
    72 (read 5 diagonals, write 4), solve 72 (forward: read ds, dl, d, b and
    write b; backward: read du, dw, b and write b)."""
    import oracle

    m, batch = 1024, 65536
    n = m * batch
    stream = torch.cuda.current_stream()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)

    def rnd(a, b):
        return torch.rand(m, batch, dtype=torch.float64, device="cuda", generator=gen) * (b - a) + a

    pristine = [rnd(-1, 1), rnd(-1, 1), rnd(6, 9), rnd(-1, 1), rnd(-1, 1)]
    rhs = rnd(-1, 1)
    work = [torch.empty_like(a) for a in pristine]
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = 5
    fac_ms = 0.0
    for r in range(reps + 1):
        for w, p in zip(work, pristine):
            w.copy_(p)
        e0.record(stream)
        api.custenPentFactorBatch(*work, m, batch, stream=stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        if r > 0:
            fac_ms += e0.elapsed_time(e1) / reps
    b = rhs.clone()
    sol = lambda: api.custenPentSolveBatch(*work, b, m, batch,  # noqa: E731
                                           stream=stream.cuda_stream)
    sol_ms = Timer(torch, stream).run(sol, reps, 1, prewarm_s=0.05)
    # parity on the first 256 systems: factor again from pristine, one solve
    for w, p in zip(work, pristine):
        w.copy_(p)
    b.copy_(rhs)
    api.custenPentFactorBatch(*work, m, batch, stream=stream.cuda_stream)
    api.custenPentSolveBatch(*work, b, m, batch, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    k = 256
    host = [p[:, :k].cpu().numpy().copy() for p in pristine]
    fac = oracle.pent_factor_batch(*host, nthreads=min(16, os.cpu_count() or 1))
    ref = oracle.pent_solve_batch(*fac, rhs[:, :k].cpu().numpy().copy())
    got = b[:, :k].cpu().numpy()
    extra["pent_batch"] = {
        "m": m, "systems": batch,
        "factor_ms": round(fac_ms, 3), "factor_GB/s": round(72.0 * n / (fac_ms * 1e-3) / 1e9, 1),
        "solve_ms": round(sol_ms, 3), "solve_GB/s": round(72.0 * n / (sol_ms * 1e-3) / 1e9, 1),
        "systems_checked_vs_oracle": k,
        "mismatching_unknowns": int(np.count_nonzero(got.view(np.uint64) != ref.view(np.uint64))),
    }
    log(f"pent batch: {extra['pent_batch']}")
    del pristine, work, rhs, b
    torch.cuda.empty_cache()


def bench_x_1024(args, torch, api, extra):
    """1024^2 x-periodic 5-point d2/dx2 (cache-resident: effective GB/s)."""
    from custen_amd import synth

    n = 1024
    h_ = 2 * np.pi / n
    x = np.arange(n) * h_
    u = np.sin(x)[None, :] * np.cos(x)[:, None] + synth.uniform(n * n, 1, -1e-3, 1e-3).reshape(n, n)
    w = np.array([-1, 16, -30, 16, -1], dtype=np.float64) / (12 * h_ * h_)
    stream = torch.cuda.current_stream()
    din = torch.from_numpy(u).cuda()
    dout = torch.empty_like(din)
    dw = torch.from_numpy(w).cuda()
    h = api.cuSten_t()
    api.cuStenCreate2DXp(h, 0, 1, n, n, 32, 32, dout, din, dw, 5, 2, 2)
    api.custenSetStream(h, stream.cuda_stream)
    api.cuStenCompute2DXp(h, api.DEVICE)
    torch.cuda.synchronize()
    t = Timer(torch, stream).run(lambda: api.cuStenCompute2DXp(h, api.DEVICE), 200, 10,
                                 prewarm_s=0.05)
    res = {"us": round(t * 1e3, 2), "GPts/s": round(n * n / (t * 1e-3) / 1e9, 1),
           "effective_GB/s": round(16.0 * n * n / (t * 1e-3) / 1e9, 1)}
    # the same 100 applies captured once in a HIP graph and replayed: the
    # per-apply host launch cost drops out
    gs = torch.cuda.Stream()
    api.custenSetStream(h, gs.cuda_stream)
    with torch.cuda.stream(gs):
        for _ in range(3):
            api.cuStenCompute2DXp(h, api.DEVICE)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=gs):
        for _ in range(100):
            api.cuStenCompute2DXp(h, api.DEVICE)
    graph.replay()
    torch.cuda.synchronize()

    def replay():
        # replay() launches on the current stream: make that the timed one
        with torch.cuda.stream(gs):
            graph.replay()

    tg = Timer(torch, gs).run(replay, 10, 2) / 100
    res["graph_us"] = round(tg * 1e3, 2)
    res["graph_effective_GB/s"] = round(16.0 * n * n / (tg * 1e-3) / 1e9, 1)
    # the graph's output equals one plain apply (same input every time): the
    # eager result goes to its own buffer, the graph's output buffer is
    # filled with NaN before one more replay, so an empty or wrong graph fails
    api.custenSetStream(h, stream.cuda_stream)
    api.cuStenCompute2DXp(h, api.DEVICE)
    torch.cuda.synchronize()
    eager = dout.clone()
    dout.fill_(float("nan"))
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    res["graph_matches_eager"] = bool(torch.equal(eager, dout))
    res["graph_applies_per_replay"] = 100
    del graph
    api.cuStenDestroy2DXp(h)
    extra["x_p_5pt_1024"] = res


def bench_batch(args, torch, api, extra):
    """512 uneven periodic grids, 13-point biharmonic, 100 steps: one launch per
    step (HIP graph) against the per-grid launch loop."""
    from custen_amd import workloads

    sizes = workloads.batch_sizes(512, 4)
    pts = sum(s * s for s in sizes)
    steps = 100
    w = torch.from_numpy(workloads.biharmonic_weights(2048) * 1e-12).cuda()
    out = {"grids": len(sizes), "points": pts, "steps": steps}
    for f32 in (False, True):
        dt = torch.float32 if f32 else torch.float64
        ins, outs = [], []
        for g, s in enumerate(sizes):
            hbuf = np.empty((s, s))
            api.custenFillUniform(hbuf, s * s, 1000 + g, -1.0, 1.0)
            t = torch.from_numpy(hbuf).to(device="cuda", dtype=dt)
            ins.append(t)
            outs.append(torch.empty_like(t))
        b = api.Batch(outs, ins, w, f32=f32)
        b.run(2, mode=0)  # tunes and builds the graph for 2 steps
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        timer = Timer(torch, stream)
        t_fused = timer.run(lambda: b.run(steps, mode=0, stream=stream.cuda_stream), 1, 1)
        t_loop = timer.run(lambda: b.run(steps, mode=1, stream=stream.cuda_stream), 1, 0)
        bpp = 8 if f32 else 16
        key = "f32" if f32 else "f64"
        out[key] = {
            "fused_ms": round(t_fused, 2), "loop_ms": round(t_loop, 2),
            "fused_GPts/s": round(pts * steps / (t_fused * 1e-3) / 1e9, 1),
            "fused_GB/s": round(bpp * pts * steps / (t_fused * 1e-3) / 1e9, 1),
            "loop_GPts/s": round(pts * steps / (t_loop * 1e-3) / 1e9, 1),
            "speedup_vs_loop": round(t_loop / t_fused, 2),
        }
        log(f"batch {key}: fused {t_fused:.1f} ms, loop {t_loop:.1f} ms")
        b.close()
        del ins, outs
        torch.cuda.empty_cache()
    extra["batched_grids"] = out


def bench_ch(args, torch, api, extra):
    """Served Cahn-Hilliard (continuous batching) against the plain loop."""
    from custen_amd import synth, workloads

    count = args.ch_requests
    steps = args.ch_steps
    reqs = workloads.ch_requests(count, rate_per_s=1.0)  # sizes and seeds; times rescaled below
    ics = {r: synth.uniform(n * n, seed, -0.1, 0.1).reshape(n, n)
           for r, (_, n, seed) in enumerate(reqs)}

    # plain: one request at a time, reference-shaped launch-and-synchronise loop
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_res = {}
    for r, (_, n, _) in enumerate(reqs):
        s = api.CHSolver(api.ch_params(n), ics[r])
        s.run(steps, mode=1)
        c, _ = s.field()
        plain_res[r] = c
        s.close()
    t_plain = time.perf_counter() - t0
    plain_rate = count / t_plain
    log(f"ch plain: {count} requests in {t_plain:.2f} s")
    # one request at a time on the fused kernels (no batching)
    t0 = time.perf_counter()
    for r, (_, n, _) in enumerate(reqs):
        s = api.CHSolver(api.ch_params(n), ics[r])
        s.run(steps, mode=0)
        s.field()
        s.close()
    t_single = time.perf_counter() - t0

    out = {"requests": count, "steps": steps, "sizes": sorted(set(n for _, n, _ in reqs)),
           "plain_solves_per_s": round(plain_rate, 2), "plain_s": round(t_plain, 3),
           "plain": "reference-shaped launch loop: 13 launches + synchronisations per step, "
                    "one request at a time",
           "fused_one_at_a_time_solves_per_s": round(count / t_single, 2)}
    budget = int(args.ch_budget_gb * (1 << 30))
    for load in (1.0, 4.0, 16.0):
        arrivals = workloads.ch_requests(count, rate_per_s=load * plain_rate)
        srv = api.CHServer(budget_bytes=budget, max_slots=args.ch_slots)
        pending = list(enumerate(arrivals))
        submit_at = {}
        ids = {}
        t0 = time.perf_counter()
        work = 0  # requests queued or resident in the server
        while pending or work > 0:
            now = time.perf_counter() - t0
            while pending and pending[0][1][0] <= now:
                r, (ta, n, _) = pending.pop(0)
                ids[r] = srv.submit(api.ch_params(n), steps, ics[r])
                submit_at[r] = (ta, time.perf_counter() - t0)
                work += 1
            if work > 0:
                work = srv.step()
            elif pending:
                time.sleep(max(0.0, min(0.002, pending[0][1][0] - (time.perf_counter() - t0))))
        makespan = time.perf_counter() - t0
        lat = []
        mism = 0
        mass = 0.0
        for r in ids:
            st, l_sub, _ = srv.query(ids[r])
            ta, ts = submit_at[r]
            lat.append(ts - ta + l_sub)
            res = srv.result(ids[r])
            mism += int(np.count_nonzero(res.view(np.uint64) != plain_res[r].view(np.uint64)))
            mass = max(mass, abs(res.mean() - ics[r].mean()))
        srv.close()
        lat = np.array(lat)
        # a Poisson stream: the completion rate is bounded by the offered
        # rate (the makespan spans the arrivals), so it measures the arrival
        # stream, not the server's capacity (served_saturated below)
        out[f"served_load{load}"] = {
            "offered_rate_per_s": round(load * plain_rate, 2),
            "arrival_limited_solves_per_s": round(count / makespan, 2),
            "last_arrival_s": round(arrivals[-1][0], 4),
            "makespan_s": round(makespan, 4),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "mismatching_points_vs_plain": mism,
            "max_mass_drift": mass,
        }
        log(f"ch served load {load}: {count / makespan:.2f} solves/s")
    # saturated throughput: everything submitted at t = 0, timed until every
    # result is in the caller's arrays (as the plain loop's field() copies)
    srv = api.CHServer(budget_bytes=budget, max_slots=args.ch_slots)
    t0 = time.perf_counter()
    sat_ids = [srv.submit(api.ch_params(n), steps, ics[r]) for r, (_, n, _) in enumerate(reqs)]
    srv.run_until_idle()
    sat_res = [srv.result(i) for i in sat_ids]
    t_sat = time.perf_counter() - t0
    srv.close()
    sat_mism = sum(int(np.count_nonzero(sat_res[r].view(np.uint64) != plain_res[r].view(np.uint64)))
                   for r in range(count))
    out["served_saturated"] = {"solves_per_s": round(count / t_sat, 2),
                               "speedup_vs_plain": round((count / t_sat) / plain_rate, 2),
                               "mismatching_points_vs_plain": sat_mism,
                               "what": "capacity: all requests submitted at once, timed from "
                                       "the first submit to the last result copied out"}
    out["cpu_oracle"] = ch_cpu_baseline(reqs, ics, steps)
    extra["served_cahn_hilliard"] = out


def ch_cpu_baseline(reqs, ics, steps, sample_steps=2):
    """The oracle's restatement of serialCahnADI.c on the host cores: every
    request advanced `sample_steps` steps, single-threaded (as the reference's
    serial program runs) and with one thread per request on up to 16 cores;
    solves/s scaled to `steps` steps per request."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor

    def advance(r):
        n = reqs[r][1]
        c = ics[r].ravel().copy()
        co = c.copy()
        oracle.CH(n).step(c, co, steps=sample_steps)

    idx = list(range(len(reqs)))
    t0 = time.perf_counter()
    for r in idx:
        advance(r)
    t1 = time.perf_counter() - t0
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(advance, idx))
    tp = time.perf_counter() - t0
    scale = steps / sample_steps
    return {"single_thread_solves_per_s": round(len(idx) / (t1 * scale), 4),
            "threads_over_requests": threads,
            "parallel_solves_per_s": round(len(idx) / (tp * scale), 4),
            "sample": f"{len(idx)} requests x {sample_steps} steps, scaled to {steps} steps"}


def host_cpu() -> dict:
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def bench_stream(args, extra=None):
    """65536 x 32768 FP64 xy-periodic 9-point stencil with host-resident pinned
    in/out, streamed through HBM in numTiles strips.  With `extra`, the figures
    go there; otherwise they are printed as one JSON line."""
    import torch

    import oracle
    from custen_amd import api, workloads

    nx, ny = 65536, 32768
    pts = nx * ny
    log("allocating 2 x 17.2 GB pinned host memory")
    hin = torch.empty((ny, nx), dtype=torch.float64, pin_memory=True)
    hout = torch.empty((ny, nx), dtype=torch.float64, pin_memory=True)
    hin_np = hin.numpy()
    api.custenFillUniform(hin_np, pts, 3, -1.0, 1.0)
    w = workloads.laplacian9_weights(nx)
    res = {"nx": nx, "ny": ny, "bytes_host_link": 2 * 8 * pts}
    for nt in (1, 4, 16):
        h = api.cuSten_t()
        api.cuStenCreate2DXYp(h, 0, nt, nx, ny, 32, 32, hout, hin, w, 3, 1, 1, 3, 1, 1)
        t0 = time.perf_counter()
        api.cuStenCompute2DXYp(h, api.HOST)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        api.cuStenDestroy2DXYp(h)
        res[f"numTiles{nt}"] = {"s": round(t, 3), "GPts/s": round(pts / t / 1e9, 2),
                                "host_link_GB/s": round(16.0 * pts / t / 1e9, 1)}
        log(f"stream numTiles={nt}: {t:.2f} s")
    # parity: rows around tile seams and the periodic wrap against the oracle
    # (interior columns; the oracle is applied non-periodically to a 5-row
    # block gathered with the periodic row wrap)
    out_np = hout.numpy()
    mism = 0
    for y0 in (0, 2047, 8191, ny - 3):
        blk = np.ascontiguousarray(np.take(hin_np, np.arange(y0 - 1, y0 + 4) % ny, axis=0))
        ref = oracle.apply(blk, w, 3, 1, 3, 1, False)
        got = out_np[np.arange(y0, y0 + 3) % ny]
        mism += int(np.count_nonzero(got[:, 1:-1].view(np.uint64) !=
                                     ref[1:4, 1:-1].view(np.uint64)))
    res["mismatching_points_sampled_rows"] = mism
    # host link ceilings: 1 GiB each way alone, and both ways at once on two
    # streams (the aggregate bounds a streamed apply, which moves the field in
    # and the result out)
    n1 = (1 << 30) // 8
    dev_a = torch.empty(n1, dtype=torch.float64, device="cuda")
    dev_b = torch.zeros(n1, dtype=torch.float64, device="cuda")
    src = hin.view(-1)[:n1]
    dst = hout.view(-1)[:n1]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()

    def timed(fn):
        best = 1e30
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    def h2d():
        with torch.cuda.stream(s1):
            dev_a.copy_(src, non_blocking=True)

    def d2h():
        with torch.cuda.stream(s2):
            dst.copy_(dev_b, non_blocking=True)

    def both():
        h2d()
        d2h()

    gib = (1 << 30) / 1e9
    res["h2d_GBps_1GiB"] = round(gib / timed(h2d), 1)
    res["d2h_GBps_1GiB"] = round(gib / timed(d2h), 1)
    agg = 2 * gib / timed(both)
    res["both_directions_aggregate_GBps"] = round(agg, 1)
    for nt in (1, 4, 16):
        r = res[f"numTiles{nt}"]
        # share of the host-link bound (all bytes at the measured aggregate rate)
        r["host_link_efficiency"] = round((16.0 * pts / 1e9 / agg) / r["s"], 3)
    del dev_a, dev_b, src, dst, hin, hout, hin_np, out_np
    torch.cuda.empty_cache()
    if extra is not None:
        extra["stream_65536x32768"] = res
    else:
        print(json.dumps({"workload": "stream", **res}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="stencil",
                    choices=["stencil", "stream", "weno", "pent", "batch", "ch", "x1024", "fun"],
                    help="stencil: the headline line (default); the others run one "
                         "secondary benchmark alone and print its figures")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the 65536 x 32768 host-streamed configuration (34 GB of "
                         "pinned host memory) in the default run")
    ap.add_argument("--tile", default="",
                    help="rows_per_wave,chunks,columns_per_lane of the strip kernel, or "
                         "band:rows_per_wave,chunks_per_band for the band walker, for the "
                         "headline apply (default: autotuned)")
    ap.add_argument("--ch-requests", type=int, default=24)
    ap.add_argument("--ch-steps", type=int, default=200)
    ap.add_argument("--ch-slots", type=int, default=64)
    ap.add_argument("--ch-budget-gb", type=float, default=64.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the stencil path has no multi-GPU code; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "stream":
        bench_stream(args)
        return
    if args.workload in ("weno", "pent", "batch", "ch", "x1024", "fun"):
        import torch

        from custen_amd import api

        extra = {}
        if args.workload == "weno":
            n = 8192
            host = np.empty((n, n))
            api.custenFillUniform(host, n * n, 2, -1.0, 1.0)
            din = torch.from_numpy(host).cuda()
            dout = torch.empty_like(din)
            stream = torch.cuda.current_stream()
            bench_weno_8192(torch, api, stream, Timer(torch, stream), din, dout,
                            max(5, args.steps // 4), extra)
        elif args.workload == "pent":
            bench_pent(args, torch, api, extra)
        elif args.workload == "ch":
            bench_ch(args, torch, api, extra)
        elif args.workload == "x1024":
            bench_x_1024(args, torch, api, extra)
        elif args.workload == "fun":
            bench_fun_8192(args, torch, api, extra)
        else:
            bench_batch(args, torch, api, extra)
        print(json.dumps(extra), flush=True)
        return

    import torch

    from custen_amd import api, workloads

    n = 8192
    extra = {}
    ms = bench_stencil_8192(args, torch, api, extra)
    if not args.no_extra:
        bench_x_1024(args, torch, api, extra)
        bench_batch(args, torch, api, extra)
        bench_pent(args, torch, api, extra)
        bench_ch(args, torch, api, extra)
        if not args.no_stream:
            bench_stream(args, extra)
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        try:
            with open(PROFILE_TRAFFIC) as f:
                traffic = json.load(f).get("xy_p_bih_8192", {}).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    pts = n * n
    achieved = 16.0 * pts / (ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": round(pts / (ms * 1e-3) / 1e9, 3),
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
        "cpu_baseline": cpu_baseline(n, workloads.biharmonic_weights(n), args.cpu_seconds),
        "extra": extra,
    }
    extra["cpu_host"] = host_cpu()
    extra["cpu_baseline_single_thread"] = cpu_baseline(
        n, workloads.biharmonic_weights(n), max(0.2, args.cpu_seconds / 3), threads=1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
