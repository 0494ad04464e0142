#!/usr/bin/env python3
"""ALMA parameter-sweep benchmark on one MI355X.

Workload: 200,000-bar f32 random walk (log-return sd 0.015, seed 2) x 1,250
ALMA combos (periods 10..=250 step 10 x offsets 0.05..=0.95 step 0.1 x sigmas
2..=10 step 2) = 250M outputs, row-major [combo][t], inputs and outputs
resident in HBM.  A step is one enqueue-only sweep launch
(vta_alma_sweep_launch).  Both kernel forms are timed: the split-BF16 MFMA
contraction (reported value, when it meets 1e-5) and the plain FP32 sliding
dot.  Kernel times come from HIP events recorded on the library's stream.

Prints ONE JSON line (metric, value, unit, ... roofline, cpu_baseline).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import vectorta_amd as vta  # noqa: E402
from vectorta_amd import datagen  # noqa: E402
from vectorta_amd._lib import check, lib  # noqa: E402

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_VALU_PEAK_TFLOPS = 157.3

# split-BF16 sweep kernel launched for each value of the alma_mfma_variant knob
# (variant 4 also launches alma_batch_mfma_bf16x3_wg16_klo for its K <= 128 groups)
MFMA_VARIANT_KERNEL = {1: "alma_batch_mfma_bf16x3_wg16", 2: "alma_batch_mfma_bf16x3_wg8",
                       3: "alma_batch_mfma_bf16x3_wg16_t2", 4: "alma_batch_mfma_bf16x3_wg16",
                       5: "alma_batch_mfma_bf16x3_r8", 6: "alma_batch_mfma_bf16x3",
                       7: "alma_batch_mfma_bf16x3_r32", 8: "alma_sweep_bf16x3",
                       9: "alma_sweep_bf16x3_d", 10: "alma_sweep_bf16x3_r4",
                       11: "alma_sweep_bf16x3_p"}
# split-F16 sweep kernel per variant (it runs the diagonal form only: below 8,
# the default 9)
F16_VARIANT_KERNEL = {8: "alma_sweep_f16x3", 9: "alma_sweep_f16x3_d", 10: "alma_sweep_f16x3_r4",
                      11: "alma_sweep_f16x3_p"}


def builtin_mfma_variant(T, C):
    """The library's built-in variant for a T-bar x C-combo sweep (its rule
    in alma.cpp, mf_default_variant)."""
    if C <= 64:
        return 5
    return 2 if T * C <= 134217728 else 9


def form_kernel(form, variant):
    """Kernel a sweep form runs at a knob value (T % 4 == 0)."""
    if form == "fp32":
        return "alma_batch_f32"
    if form == "mfma_f16x3":
        return F16_VARIANT_KERNEL[variant if variant >= 8 else 9]
    return MFMA_VARIANT_KERNEL[variant]


def settle(fn, sync, seconds):
    """Runs `fn` back to back, untimed, for `seconds`: the engine clock of a
    power-capped launch stream dips in its first milliseconds and settles
    within a fraction of a second, so the timed launches that follow see the
    sustained clock a stream of such launches runs at."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(8):
            fn()
        sync()


METRIC = "ALMA sweep outputs/s at <=1e-5 rel err vs FP64 (MFMA vs FP32); p50/p99 latency"


class HipEvents:
    """hipEvent timing on an explicit stream (libamdhip64 via ctypes)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return float(ms.value)

    def destroy(self, e):
        self.h.hipEventDestroy(e)


def profiled_traffic(summary: str, kernel: str):
    """HBM bytes per launch of `kernel` recorded in profiles/<summary> (two
    rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel
    calibration); None if absent."""
    path = os.path.join(ROOT, "profiles", summary)
    key = f"traffic_bytes_per_launch[{kernel}] = "
    try:
        with open(path) as f:
            for line in f:
                line = line.strip()
                if line.startswith(key):
                    return int(line[len(key):])
    except OSError:
        pass
    return None


def algorithmic_work(combos, T):
    """Bytes: outputs written + prices read + f32 weights read; FLOPs: 2*p per valid output."""
    out_b = len(combos) * T * 4
    w_b = sum(c.period for c in combos) * 4
    flops = sum(2.0 * c.period * (T - c.period + 1) for c in combos)
    return out_b + T * 4 + w_b, flops


def time_form(alma, plan, d_x, d_out, steps, warmup, ev):
    """Returns (wall ms per step, mean launch ms, per-launch ms array).  The mean
    comes from `steps` launches enqueued back to back between two HIP events on
    the library's stream (the GPU never waits on the host inside the bracket);
    the per-launch array, for p50/p99, from a second pass with an event pair
    around every launch."""
    mean, ks, wall = _time_launches(lambda: plan.launch(d_x.ptr, 0, d_out.ptr), alma.ctx.stream,
                                    steps, warmup, ev, sync=alma.synchronize)
    return wall, mean, ks


SETTLE_S = 1.0   # set from --settle


def cpu_baseline(x, combos, threads, reps=10):
    """The f64 oracle (C restatement of alma_row_scalar) over the whole sweep on
    `threads` cores: the median of `reps` runs after a warm-up; returns the
    last run's rows (the error reference) and the median seconds."""
    import oracle
    sel = [(c.period, c.offset, c.sigma) for c in combos]
    xd = x.astype(np.float64)
    oracle.alma_batch(xd[:4096], sel[:2], nthreads=threads)  # warm-up
    ts = []
    ref = None
    for _ in range(reps):
        ref = None
        t0 = time.perf_counter()
        ref = oracle.alma_batch(xd, sel, nthreads=threads)
        ts.append(time.perf_counter() - t0)
    return ref, float(np.median(ts))


def cpu_baseline_avx2(x, combos, threads, reps=10):
    """The AVX2+FMA row (alma_row_avx2, the reference's nightly-avx build) over
    the whole sweep: median seconds of `reps` runs after a warm-up, None
    without AVX2."""
    import oracle
    sel = [(c.period, c.offset, c.sigma) for c in combos]
    xd = x.astype(np.float64)
    if oracle.alma_batch_avx2(xd[:4096], sel[:2], nthreads=threads) is None:
        return None
    oracle.alma_batch_avx2(xd, sel[:threads], nthreads=threads)  # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.alma_batch_avx2(xd, sel, nthreads=threads)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline_one_core(x, combos, every=10, reps=10):
    """The same oracle on one core over every `every`-th combo of the sweep
    (a bounded sample): median seconds of `reps` runs after a warm-up and the
    combos timed."""
    import oracle
    sel = [(c.period, c.offset, c.sigma) for c in combos][::every]
    xd = x.astype(np.float64)
    oracle.alma_batch(xd, sel[:2], nthreads=1)  # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.alma_batch(xd, sel, nthreads=1)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), len(sel)


def max_rel_err(g, ref):
    m = np.isfinite(ref)
    nan_ok = bool(np.array_equal(np.isnan(g), np.isnan(ref)))
    err = 0.0
    for r in range(ref.shape[0]):  # row-wise keeps the temporaries small
        mr = m[r]
        e = np.abs(g[r, mr].astype(np.float64) - ref[r, mr]) / np.abs(ref[r, mr])
        err = max(err, float(e.max()) if e.size else 0.0)
    return err, nan_ok


def calibrate_copy(nbytes: int = 256 << 20, reps: int = 3) -> None:
    """Plain streaming kernels of known traffic -- the references that scale
    the FETCH_SIZE / WRITE_SIZE counters.  torch.neg(out=) on 16-byte-aligned
    buffers issues 16-byte-per-lane loads; on views shifted by one float it
    cannot vectorise and issues 4-byte-per-lane loads (the width the indicator
    kernels use).  Each reads and writes nbytes (the shifted one 4 bytes
    less).  A device-to-device hipMemcpy runs on the copy engines, which those
    counters do not see."""
    import torch
    n = nbytes // 4
    src = torch.ones(n + 1, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for label, s_, d_ in (("vec16", src[:n], dst[:n]), ("vec4", src[1:], dst[1:])):
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            torch.neg(s_, out=d_)
            b.record()
            torch.cuda.synchronize()
            ms = a.elapsed_time(b)
            print(f"[calib] neg {label} {s_.numel() * 4} bytes: {ms:.4f} ms, "
                  f"{2 * s_.numel() * 4 / ms / 1e6:.1f} GB/s (read + write)", flush=True)
        assert float(d_[0]) == -1.0
    # write-only streams: the ceiling for a kernel that only writes (the ALMA
    # sweep writes 1.0 GB and reads 2 MB)
    big = torch.empty(1 << 28, dtype=torch.float32, device="cuda")   # 1 GiB
    for _ in range(reps + 2):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        big.fill_(1.0)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        print(f"[calib] fill {big.numel() * 4} bytes: {ms:.4f} ms, "
              f"{big.numel() * 4 / ms / 1e6:.1f} GB/s (write only)", flush=True)


def cpu_median(fn, reps=10):
    """Median seconds of `reps` calls of fn() after one warm-up call."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _cpu_threads():
    return CPU_THREADS or len(os.sched_getaffinity(0))


CPU_THREADS = 0      # set from --cpu-threads; 0 = every host core this process may use
CPU_EXTRA = True     # --no-cpu turns the extra workloads' baselines off too


def bench_many_series_ohlc(steps, warmup, ev):
    """Fused EMA(20) + RSI(14) + ATR(14) over 4,096 uneven tickers right-aligned
    in a [50,000 x 4,096] time-major matrix (one enqueue-only launch per step).
    Algorithmic bytes = the valid high/low/close bars read + all three output
    matrices written."""
    o, h, lo, c, lens = datagen.many_series_ohlc(4096, 50_000, seed=3)
    rows, cols = c.shape
    ctx = vta.HipContext(0)
    plan = vta.ManySeriesOhlcPlan(ctx, c, h, lo)
    del o
    mean, ks, _ = _time_launches(lambda: plan.launch(20, 14, 14), ctx.stream, steps, warmup, ev,
                                 sync=ctx.synchronize)
    alg = 3 * 4 * int(lens.sum()) + 3 * 4 * rows * cols
    t = mean * 1e-3
    res = {"workload": "ema20_rsi14_atr14_many_series_4096x50000", "kernel": "ms_ema_rsi_atr_f32",
           "kernel_ms_mean": mean, "kernel_ms_p50": float(np.percentile(ks, 50)),
           "kernel_ms_p99": float(np.percentile(ks, 99)), "outputs": 3 * rows * cols,
           "outputs_per_s": 3 * rows * cols / t, "algorithmic_bytes": alg,
           "traffic": profiled_traffic("ema_rsi_atr_many_series_4096x50000_traffic.txt",
                                       "ms_ema_rsi_atr_f32"),
           "achieved_GBs": alg / t / 1e9, "frac_of_hbm_peak": alg / t / 1e9 / HBM_PEAK_GBS}
    if CPU_EXTRA:
        # the oracle's scalar EMA, Wilder RSI and ATR per ticker, each ticker's
        # NaN-padded column as one contiguous f64 series (the reference's CPU
        # path runs series by series), OpenMP over tickers
        import oracle
        C, H, L = (np.ascontiguousarray(a.T, dtype=np.float64) for a in (c, h, lo))
        thr = _cpu_threads()

        def run(n, threads):
            oracle.series_major("ema", C[:n], 20, nthreads=threads)
            oracle.series_major("rsi", C[:n], 14, nthreads=threads)
            oracle.series_major("atr", H[:n], 14, low=L[:n], close=C[:n], nthreads=threads)

        dt = cpu_median(lambda: run(cols, thr))
        sample = 256
        dt1 = cpu_median(lambda: run(sample, 1), reps=3)
        res["cpu_baseline"] = {
            "value": 3 * rows * cols / dt, "unit": "outputs/s", "cores": thr, "kind": "port",
            "sample": f"all {cols} tickers, f64 C restatements of ema/rsi/atr scalar, one "
                      f"contiguous series per task, median of 10 runs {dt:.2f} s",
            "one_core": {"value": 3 * rows * sample / dt1, "unit": "outputs/s", "cores": 1,
                         "sample": f"first {sample} tickers, median of 3 runs {dt1:.2f} s"}}
        del C, H, L
    return res


def bench_rolling(steps, warmup, ev):
    """Bollinger Bands (20, 2.0), percentile_nearest_rank (100, 70) and rolling
    MAD (64) over 1,024 random-walk series x 50,000 bars with 1% +/-5-sigma
    jumps, time-major, device resident; one enqueue-only launch per step.
    Algorithmic bytes = every bar read once + every output written."""
    x = datagen.many_series_with_jumps(1024, 50_000, seed=4)
    rows, cols = x.shape
    ctx = vta.HipContext(0)
    plan = vta.RollingManySeriesPlan(ctx, x)
    x_host = x
    res = {"workload": "bb20_pnr100p70_mad64_many_series_1024x50000"}
    from vectorta_amd import autotune as at
    # PNR window 100 and MAD window 64 run the register-window kernels (window
    # rounded up to 16 slots); the LDS sorted-window form is timed beside them
    for name, fn, nout, kern, lds_kern in (
            ("bb", lambda: plan.launch_bb(20, 2.0, 2.0), 3, "bb_ms1p_f32", None),
            ("pnr", lambda: plan.launch_pnr(100, 70.0), 1, "pnr_ms1p_reg112", "pnr_ms1p_f32"),
            ("mad", lambda: plan.launch_mad(64), 1, "mad_ms1p_len64", "mad_ms1p_f32")):
        mean, ks, _ = _time_launches(fn, ctx.stream, steps, warmup, ev, sync=ctx.synchronize)
        t = mean * 1e-3
        alg = 4 * rows * cols * (1 + nout)
        res[name] = {"kernel": kern, "kernel_ms_mean": mean,
                     "kernel_ms_p50": float(np.percentile(ks, 50)),
                     "kernel_ms_p99": float(np.percentile(ks, 99)),
                     "series_bars_per_s": rows * cols / t, "algorithmic_bytes": alg,
                     "traffic": profiled_traffic("rolling_windows_1024x50000_traffic.txt", kern),
                     "achieved_GBs": alg / t / 1e9,
                     "frac_of_hbm_peak": alg / t / 1e9 / HBM_PEAK_GBS}
        if lds_kern:
            at.set_tuning(ctx, at.ROLL_FORM, 1)
            lds_mean, _, _ = _time_launches(fn, ctx.stream, steps, warmup, ev,
                                            sync=ctx.synchronize)
            at.set_tuning(ctx, at.ROLL_FORM, 0)
            res[name]["lds_form"] = {"kernel": lds_kern, "kernel_ms_mean": lds_mean}
    if CPU_EXTRA:
        # the oracle's scalar bollinger_bands / pnr / medium_ad per series on a
        # bounded sample of contiguous f64 series (medium_ad sorts every
        # window, ~2 us per bar on one core), OpenMP over series
        import oracle
        X = np.ascontiguousarray(x_host.T, dtype=np.float64)
        thr = _cpu_threads()
        for name, kind, prm, n in (("bb", "bb", (20, 2.0, 2.0), 512),
                                   ("pnr", "pnr", (100, 70.0), 256),
                                   ("mad", "medium_ad", (64,), 64)):
            n = min(n, cols)
            n1 = max(1, n // 16)
            dt = cpu_median(lambda: oracle.series_major(kind, X[:n], *prm, nthreads=thr))
            dt1 = cpu_median(lambda: oracle.series_major(kind, X[:n1], *prm, nthreads=1),
                             reps=3)
            res[name]["cpu_baseline"] = {
                "value": rows * n / dt, "unit": "series_bars/s", "cores": thr, "kind": "port",
                "sample": f"first {n} of {cols} series x {rows} bars, median of 10 runs "
                          f"{dt:.3f} s",
                "one_core": {"value": rows * n1 / dt1, "unit": "series_bars/s", "cores": 1,
                             "sample": f"first {n1} series, median of 3 runs {dt1:.3f} s"}}
        del X
    return res


def _time_launches(fn, stream, steps, warmup, ev, sync=None):
    """(mean launch ms, per-launch ms array, wall ms per step) of enqueue-only
    launches `fn` on `stream`: the mean from `steps` launches back to back
    between two HIP events, the per-launch figures from a second pass with an
    event pair around each launch.  Before the warm-up launches the stream is
    settled (settle(): SETTLE_S seconds of untimed launches)."""
    if sync and SETTLE_S > 0:
        settle(fn, sync, SETTLE_S)
    for _ in range(warmup):
        fn()
    if sync:
        sync()
    a, b = ev.create(), ev.create()
    t0 = time.perf_counter()
    ev.record(a, stream)
    for _ in range(steps):
        fn()
    ev.record(b, stream)
    total = ev.elapsed_ms(a, b)
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ev.destroy(a)
    ev.destroy(b)
    ks = []
    for _ in range(steps):
        a, b = ev.create(), ev.create()
        ev.record(a, stream)
        fn()
        ev.record(b, stream)
        ks.append(ev.elapsed_ms(a, b))
        ev.destroy(a)
        ev.destroy(b)
    return total / steps, np.array(ks), wall


def bench_captured_batch(steps, warmup, ev):
    """A batch of 16 small one-series requests on resident inputs (EMA, RSI,
    ADX, SMA over 32 periods and Bollinger Bands over 4, on four 10,000-bar
    OHLC series): the enqueue-only calls launched eagerly vs the same calls
    captured once into a hipGraph and replayed.  Reported per batch: host
    wall time of enqueue + synchronise, and GPU time between two HIP events."""
    L = lib()
    P = ctypes.c_void_p
    T, fv, ns = 10_000, 0, 4
    ctx = vta.HipContext(0)
    rng = np.random.Generator(np.random.PCG64(77))
    per = np.arange(4, 132, 4, dtype=np.int32)                 # 32 periods
    bper = np.array([10, 20, 30, 40], dtype=np.int32)
    d_p, d_bp = ctx.to_device(per), ctx.to_device(bper)
    d_bu, d_bd = ctx.to_device(np.full(4, 2.0)), ctx.to_device(np.full(4, 2.0))
    series, outs = [], []
    for s in range(ns):
        c = datagen.gbm_close(T, 0.01, seed=100 + s)
        _, h, lo, c = datagen.ohlc_from_close(c, rng)
        series.append((ctx.to_device(c), ctx.to_device(h), ctx.to_device(lo)))
        outs.append([ctx.alloc(per.size * T * 4) for _ in range(4)] +
                    [ctx.alloc(bper.size * T * 4) for _ in range(3)])
    C, BC, mp = int(per.size), int(bper.size), int(per.max())

    def batch():
        h = ctx.handle
        for (dc, dh, dl), o in zip(series, outs):
            check(L.vta_ema_batch_device(h, P(dc.ptr), P(d_p.ptr), None, T, fv, C, P(o[0].ptr)))
            check(L.vta_rsi_batch_device(h, P(dc.ptr), P(d_p.ptr), mp, T, fv, C, P(o[1].ptr)))
            check(L.vta_adx_batch_device(h, P(dh.ptr), P(dl.ptr), P(dc.ptr), P(d_p.ptr), mp, T,
                                         fv, C, P(o[2].ptr)))
            check(L.vta_sma_batch_device(h, P(dc.ptr), P(d_p.ptr), T, C, fv, mp, P(o[3].ptr)))
            check(L.vta_bollinger_bands_batch_device(
                h, P(dc.ptr), T, fv, T, P(d_bp.ptr), P(d_bu.ptr), P(d_bd.ptr), BC,
                int(bper.max()), P(o[4].ptr), P(o[5].ptr), P(o[6].ptr)))

    ctx.preload()
    ctx.begin_capture()
    batch()
    graph = ctx.end_capture()
    res = {"workload": "captured_batch_16_requests_10k_bars", "requests": 4 * ns,
           "graph_nodes": graph.num_nodes,
           "outputs_per_batch": ns * T * (4 * C + 3 * BC)}
    for name, fn in (("eager", batch), ("graph", graph.launch)):
        for _ in range(warmup):
            fn()
        ctx.synchronize()
        walls = []
        for _ in range(steps):
            t0 = time.perf_counter()
            fn()
            ctx.synchronize()
            walls.append((time.perf_counter() - t0) * 1e3)
        a, b = ev.create(), ev.create()
        ev.record(a, ctx.stream)
        for _ in range(steps):
            fn()
        ev.record(b, ctx.stream)
        gpu_ms = ev.elapsed_ms(a, b) / steps
        ev.destroy(a)
        ev.destroy(b)
        res[name] = {"wall_ms_p50": float(np.percentile(walls, 50)),
                     "wall_ms_p99": float(np.percentile(walls, 99)),
                     "gpu_ms_mean_back_to_back": gpu_ms}
    res["wall_speedup"] = res["eager"]["wall_ms_p50"] / res["graph"]["wall_ms_p50"]
    graph.destroy()
    for bufs in [d_p, d_bp, d_bu, d_bd] + [b for s in series for b in s] + \
            [b for o in outs for b in o]:
        bufs.free()
    return res


def bench_packed_batch(steps, warmup, ev):
    """64 one-series requests of one indicator (lengths log-uniform 1k-50k
    bars, grids log-uniform 1-128 periods, seeded), inputs resident: one
    vta_<ind>_batch_device call per request (back to back on one stream) vs
    one packed launch (vta_rec_pack_launch); and 64 ALMA sweeps of 16-256
    combos, one sweep-plan launch each vs one vta_alma_pack launch.  Outputs
    are bit-identical; both are timed between two HIP events over `steps`
    repetitions."""
    L = lib()
    P = ctypes.c_void_p
    rng = np.random.Generator(np.random.PCG64(31))
    n = 64
    lens = np.exp(rng.uniform(np.log(1_000), np.log(50_000), n)).astype(int)
    grids = np.exp(rng.uniform(0, np.log(128), n)).astype(int)
    res = {"workload": "packed_batch_64_requests_1k_50k_bars", "requests": n,
           "outputs_per_batch": int((lens * grids).sum())}
    for kind in ("ema", "adx"):
        ctx = vta.HipContext(0)
        jobs, bufs = [], []
        for i in range(n):
            T, C = int(lens[i]), int(grids[i])
            c = datagen.gbm_close(T, 0.01, seed=600 + i)
            _, h, lo, c = datagen.ohlc_from_close(c, rng)
            per = np.arange(2, 2 + C, dtype=np.int32) * 2
            d = {"c": ctx.to_device(c), "h": ctx.to_device(h), "l": ctx.to_device(lo),
                 "p": ctx.to_device(per), "out": ctx.alloc(C * T * 4)}
            job = {"close": d["c"], "periods": d["p"], "out": d["out"], "series_len": T,
                   "first_valid": 0, "n_combos": C, "max_period": int(per.max())}
            if kind == "adx":
                job["high"], job["low"] = d["h"], d["l"]
            jobs.append(job)
            bufs.append(d)
        pack = vta.RecPack(ctx, kind, jobs)

        def direct():
            for j, d in zip(jobs, bufs):
                if kind == "ema":
                    check(L.vta_ema_batch_device(ctx.handle, P(d["c"].ptr), P(d["p"].ptr), None,
                                                 j["series_len"], 0, j["n_combos"],
                                                 P(d["out"].ptr)))
                else:
                    check(L.vta_adx_batch_device(ctx.handle, P(d["h"].ptr), P(d["l"].ptr),
                                                 P(d["c"].ptr), P(d["p"].ptr), j["max_period"],
                                                 j["series_len"], 0, j["n_combos"],
                                                 P(d["out"].ptr)))
        r = {}
        for name, fn in (("direct_calls", direct), ("packed", pack.launch)):
            mean, ks, wall = _time_launches(fn, ctx.stream, steps, warmup, ev,
                                            sync=ctx.synchronize)
            r[name] = {"gpu_ms_mean": mean, "gpu_ms_p50": float(np.percentile(ks, 50)),
                       "wall_ms_per_batch": wall}
        r["speedup"] = r["direct_calls"]["gpu_ms_mean"] / r["packed"]["gpu_ms_mean"]
        res[kind] = r
        pack.close()
        for d in bufs:
            for b in d.values():
                b.free()
    # ALMA: 64 sweeps of 16..256 combos (periods 10..250, offsets 0.05..0.95)
    # over the same lengths, split-F16 (AUTO): one sweep-plan launch per request
    # vs one vta_alma_pack launch
    alma = vta.CudaAlma(0)
    jobs, bufs, plans, outs = [], [], [], 0
    for i in range(n):
        T = int(lens[i])
        np_, no = int(rng.integers(4, 17)), int(rng.integers(4, 17))
        sw = vta.AlmaBatchRange(period=(10, 10 + 16 * (np_ - 1), 16),
                                offset=(0.05, 0.95, 0.9 / (no - 1)), sigma=(6.0, 6.0, 0.0))
        C = len(vta.expand_grid(sw))
        x = datagen.gbm_close(T, 0.01, seed=700 + i)
        d = {"x": alma.ctx.to_device(x), "out": alma.ctx.alloc(C * T * 4)}
        jobs.append({"prices": d["x"], "out": d["out"], "sweep": sw, "series_len": T,
                     "first_valid": 0})
        plans.append(alma.sweep_plan(sw, T, kernel="mfma_f16x3"))
        bufs.append(d)
        outs += C * T
    pack = vta.AlmaPack(alma.ctx, jobs)

    def direct_alma():
        for pl, d in zip(plans, bufs):
            pl.launch(d["x"].ptr, 0, d["out"].ptr)

    r = {"kernel": pack.kernel, "outputs_per_batch": outs}
    for name, fn in (("direct_calls", direct_alma), ("packed", pack.launch)):
        mean, ks, wall = _time_launches(fn, alma.ctx.stream, steps, warmup, ev,
                                        sync=alma.synchronize)
        r[name] = {"gpu_ms_mean": mean, "gpu_ms_p50": float(np.percentile(ks, 50)),
                   "wall_ms_per_batch": wall}
    r["speedup"] = r["direct_calls"]["gpu_ms_mean"] / r["packed"]["gpu_ms_mean"]
    res["alma"] = r
    pack.close()
    for pl in plans:
        pl.close()
    for d in bufs:
        for b in d.values():
            b.free()
    return res


def bench_tune_shapes(reps):
    """Autotuner sweep over served shapes: for series lengths 2^12..2^20 and
    ALMA grids of 8..2,000 combos (periods spread over 10..250), every
    split-BF16 kernel variant timed by vta_autotune; and the chunked-scan
    wave target for EMA and ADX batches at two shapes.  Prints the best
    variant per shape (the table the library's built-in choice follows)."""
    from vectorta_amd import autotune as at
    alma = vta.CudaAlma(0)
    ctx = alma.ctx
    grids = {8: ((10, 250, 240), (0.25, 0.85, 0.6), (3.0, 6.0, 3.0)),
             64: ((10, 250, 80), (0.05, 0.95, 0.3), (2.0, 8.0, 2.0)),
             512: ((10, 250, 34), (0.05, 0.95, 0.128), (2.0, 9.0, 1.0)),
             2000: ((10, 250, 16), (0.05, 0.95, 0.1), (2.0, 10.0, 2.0))}
    res = {"workload": "autotune_shape_sweep", "alma": []}
    for T in (1 << 12, 1 << 14, 1 << 16, 1 << 18, 1 << 20):
        x = datagen.gbm_close(T, 0.015, seed=5)
        d_x = ctx.to_device(x)
        for target, (p, o, s) in grids.items():
            sweep = vta.AlmaBatchRange(period=p, offset=o, sigma=s)
            plan = vta.AlmaSweepPlan(alma, sweep, T, kernel="mfma")
            d_o = ctx.alloc(plan.n_combos * T * 4)
            best, times = at.measure(ctx, at.ALMA_MFMA_VARIANT, at.CANDIDATES[at.ALMA_MFMA_VARIANT],
                                     lambda: plan.launch(d_x.ptr, 0, d_o.ptr), reps)
            row = {"series_len": T, "combos": plan.n_combos, "n_tiles": (T + 31) // 32,
                   "best": best, "ms": times}
            res["alma"].append(row)
            print(f"[tune] alma T={T} C={plan.n_combos} best={best} "
                  f"{ {k: round(v, 4) for k, v in times.items() if v} }", file=sys.stderr,
                  flush=True)
            plan.close()
            d_o.free()
        d_x.free()
    at.set_tuning(ctx, at.ALMA_MFMA_VARIANT, 0)
    res["scan"] = []
    ema, adx = vta.CudaEma(0), vta.CudaAdx(0)
    for T, C in ((16_384, 64), (200_000, 1_250)):
        c = datagen.gbm_close(T, 0.01, seed=6)
        rng = np.random.Generator(np.random.PCG64(6))
        _, h, lo, c = datagen.ohlc_from_close(c, rng)
        per = np.arange(2, 2 + C, dtype=np.int32)
        for name, op in (("ema", ema), ("adx", adx)):
            cx = op.ctx
            d_c, d_h, d_l, d_p = (cx.to_device(a) for a in (c, h, lo, per))
            d_o = cx.alloc(C * T * 4)
            P = ctypes.c_void_p
            if name == "ema":
                fn = lambda: check(lib().vta_ema_batch_device(
                    cx.handle, P(d_c.ptr), P(d_p.ptr), None, T, 0, C, P(d_o.ptr)))
            else:
                fn = lambda: check(lib().vta_adx_batch_device(
                    cx.handle, P(d_h.ptr), P(d_l.ptr), P(d_c.ptr), P(d_p.ptr), int(per.max()), T,
                    0, C, P(d_o.ptr)))
            best, times = at.measure(cx, at.SCAN_WAVES, at.CANDIDATES[at.SCAN_WAVES], fn, reps)
            res["scan"].append({"indicator": name, "series_len": T, "combos": C, "best": best,
                                "ms": times})
            print(f"[tune] {name} T={T} C={C} best={best} "
                  f"{ {k: round(v, 4) for k, v in times.items() if v} }", file=sys.stderr,
                  flush=True)
            at.set_tuning(cx, at.SCAN_WAVES, 0)
            for b in (d_c, d_h, d_l, d_p, d_o):
                b.free()
    return res


def bench_moving_averages(steps, warmup, ev):
    """SMA / WMA / SWMA / SMMA / Wilders.  One series x 1,250 periods (2..=1251)
    over the 200,000-bar sweep series (the ALMA sweep's output size, 1.0 GB),
    and time-major 4,096 x 50,000 many-series at one period; inputs resident,
    one enqueue-only launch per step.  Algorithmic bytes = outputs written +
    inputs read once."""
    x = datagen.alma_sweep_series()
    T = x.size
    res = {"workload": "ma_batch_200k_x_1250_and_many_series_4096x50000"}
    ops = {"sma": vta.CudaSma(0), "wma": vta.CudaWma(0), "swma": vta.CudaSwma(0),
           "smma": vta.CudaSmma(0), "wilders": vta.CudaWilders(0)}
    ctx = ops["sma"].ctx
    periods = np.arange(2, 1252, dtype=np.int32)
    C = periods.size
    d_x = ctx.to_device(x)
    d_p = ctx.to_device(periods)
    d_o = ctx.alloc(C * T * 4)
    kern = {"sma": "fir_batch_f32", "wma": "fir_batch_f32", "swma": "fir_batch_f32",
            "smma": "smma_batch_f32", "wilders": "wilders_batch_f32"}
    d_w = ctx.to_device((periods - 1).astype(np.int32))     # warm = first_valid + p - 1
    for name, op in ops.items():
        if name == "smma":
            fn = (lambda op=op: op.smma_batch_device(d_x, d_p, d_w, 0, T, C, d_o))
        elif name == "wilders":
            fn = (lambda op=op: op.wilders_batch_device(d_x, d_p, None, d_w, T, 0, C, d_o))
        elif name == "swma":
            fn = (lambda op=op: op.swma_batch_device(d_x, d_p, d_w, T, C, int(periods.max()),
                                                     d_o))
        else:
            fn = (lambda op=op, name=name: getattr(op, f"{name}_batch_device")(
                d_x, d_p, T, C, 0, int(periods.max()), d_o))
        # every op launches on the stream of its own context
        mean, ks, _ = _time_launches(fn, op.ctx.stream, steps, warmup, ev, sync=op.synchronize)
        op.synchronize()
        t = mean * 1e-3
        alg = 4 * (C * T + T)
        res[f"{name}_batch"] = {"kernel": kern[name], "combos": C, "series_len": T,
                                "kernel_ms_mean": mean,
                                "kernel_ms_p50": float(np.percentile(ks, 50)),
                                "outputs_per_s": C * T / t, "algorithmic_bytes": alg,
                                "achieved_GBs": alg / t / 1e9,
                                "frac_of_hbm_peak": alg / t / 1e9 / HBM_PEAK_GBS}
        if name in ("sma", "wma", "swma"):
            res[f"{name}_batch"]["traffic"] = profiled_traffic(
                "moving_averages_traffic.txt", f"fir_batch_f32[{name}]")
    op.synchronize()       # reports a warm-index mismatch, if any
    d_o.free()
    d_x.free()
    d_p.free()
    d_w.free()
    rows, cols = 50_000, 4096
    rng = np.random.Generator(np.random.PCG64(9))
    tm = (100.0 * np.exp(np.cumsum(rng.normal(0, 0.01, (rows, cols)), axis=0))).astype(np.float32)
    d_tm = ctx.to_device(tm)
    fvs = np.zeros(cols, dtype=np.int32)
    d_fv = ctx.to_device(fvs)
    # per-series metadata prepared once (every bar finite): clean flags for
    # SMMA / Wilders, first non-finite bar (= rows) for SMA / WMA / SWMA
    d_clean = ctx.to_device(np.ones(cols, dtype=np.int32))
    d_z = ctx.to_device(np.full(cols, rows, dtype=np.int32))
    d_o = ctx.alloc(rows * cols * 4)
    del tm
    from vectorta_amd.rec import ma_many_series_prepared_device
    for name, p in (("sma", 14), ("wma", 30), ("swma", 14), ("smma", 14), ("wilders", 14)):
        op = ops[name]
        meta = d_clean if name in ("smma", "wilders") else d_z
        fn = (lambda op=op, name=name, p=p, meta=meta: ma_many_series_prepared_device(
            op.ctx, name, d_tm, d_fv, meta, p, cols, rows, d_o))
        mean, ks, _ = _time_launches(fn, op.ctx.stream, steps, warmup, ev, sync=op.synchronize)
        op.synchronize()
        t = mean * 1e-3
        alg = 4 * 2 * rows * cols
        res[f"{name}_many_series"] = {"period": p, "kernel_ms_mean": mean,
                                      "kernel_ms_p50": float(np.percentile(ks, 50)),
                                      "outputs_per_s": rows * cols / t,
                                      "algorithmic_bytes": alg,
                                      "achieved_GBs": alg / t / 1e9,
                                      "frac_of_hbm_peak": alg / t / 1e9 / HBM_PEAK_GBS}
    d_o.free()
    d_tm.free()
    d_fv.free()
    d_clean.free()
    d_z.free()
    return res


def bench_recurrence_batches(steps, warmup, ev):
    """EMA / RSI / ATR / ADX one series x 1,250 periods (2..=1251) over 200,000
    bars (1.0 GB of outputs per launch), inputs resident, enqueue-only
    `_batch_device` launches (the chunked scan; ADX its two-stage form).
    Algorithmic bytes = outputs written + input bars read once (1 or 3 series)."""
    rng = np.random.Generator(np.random.PCG64(12))
    c = datagen.alma_sweep_series()
    _, h, lo, c = datagen.ohlc_from_close(c, rng)
    T = c.size
    periods = np.arange(2, 1252, dtype=np.int32)
    C = periods.size
    ops = {"ema": vta.CudaEma(0), "rsi": vta.CudaRsi(0), "atr": vta.CudaAtr(0),
           "adx": vta.CudaAdx(0)}
    ctx = ops["ema"].ctx
    d_c, d_h, d_l = ctx.to_device(c), ctx.to_device(h), ctx.to_device(lo)
    d_p = ctx.to_device(periods)
    d_o = ctx.alloc(C * T * 4)
    L = lib()
    P = ctypes.c_void_p
    fns = {
        "ema": lambda: L.vta_ema_batch_device(ops["ema"].ctx.handle, P(d_c.ptr), P(d_p.ptr), None,
                                              T, 0, C, P(d_o.ptr)),
        "rsi": lambda: L.vta_rsi_batch_device(ops["rsi"].ctx.handle, P(d_c.ptr), P(d_p.ptr),
                                              int(periods.max()), T, 0, C, P(d_o.ptr)),
        "atr": lambda: L.vta_atr_batch_device(ops["atr"].ctx.handle, P(d_h.ptr), P(d_l.ptr),
                                              P(d_c.ptr), P(d_p.ptr), int(periods.max()), T, 0,
                                              C, P(d_o.ptr)),
        "adx": lambda: L.vta_adx_batch_device(ops["adx"].ctx.handle, P(d_h.ptr), P(d_l.ptr),
                                              P(d_c.ptr), P(d_p.ptr), int(periods.max()), T, 0,
                                              C, P(d_o.ptr)),
    }
    res = {"workload": "recurrence_batches_200k_x_1250"}
    for name, fn in fns.items():
        op = ops[name]
        mean, ks, _ = _time_launches(lambda fn=fn: check(fn()), op.ctx.stream, steps, warmup, ev,
                                     sync=op.synchronize)
        nin = 3 if name in ("atr", "adx") else 1
        alg = 4 * (C * T + nin * T)
        res[name] = {"combos": C, "series_len": T, "kernel_ms_mean": mean,
                     "kernel_ms_p50": float(np.percentile(ks, 50)),
                     "outputs_per_s": C * T / (mean * 1e-3), "algorithmic_bytes": alg,
                     "traffic_output_pass": profiled_traffic(
                         "recurrence_batches_200k_x1250_traffic.txt",
                         f"{name}_batch_f32[pass3]"),
                     "achieved_GBs": alg / (mean * 1e-3) / 1e9,
                     "frac_of_hbm_peak": alg / (mean * 1e-3) / 1e9 / HBM_PEAK_GBS}
    for b in (d_c, d_h, d_l, d_p, d_o):
        b.free()
    if CPU_EXTRA:
        # the oracle's scalar ema / rsi / atr (the batch path's row) / adx,
        # one row per period, on a bounded sample of the periods (every 5th),
        # OpenMP over rows, plus a one-core sample (every 125th)
        import oracle
        f = [a.astype(np.float64) for a in (h, lo, c)]
        thr = _cpu_threads()
        sample, one = periods[::5], periods[::125]
        for name in fns:
            dt = cpu_median(lambda: oracle.batch_periods(name, f[2], sample, h=f[0], l=f[1],
                                                         nthreads=thr), reps=3)
            dt1 = cpu_median(lambda: oracle.batch_periods(name, f[2], one, h=f[0], l=f[1],
                                                          nthreads=1), reps=3)
            res[name]["cpu_baseline"] = {
                "value": sample.size * T / dt, "unit": "outputs/s", "cores": thr,
                "kind": "port",
                "sample": f"every 5th period ({sample.size} of {C}) x {T} bars, median of 3 "
                          f"runs {dt:.3f} s",
                "one_core": {"value": one.size * T / dt1, "unit": "outputs/s", "cores": 1,
                             "sample": f"every 125th period ({one.size}), median of 3 runs "
                                       f"{dt1:.3f} s"}}
    return res


def bench_double_alma(steps, warmup, ev):
    """Double-ALMA crossover backtest reduction: 220 fast x 265 slow ALMA periods
    (58,300 pairs) over the 200,000-bar sweep series, prices resident.  A step is
    one enqueue of the f64 rows of the 290 distinct periods, the pair walk and
    the segment reduction; per-pair statistics stay on the device until read.
    The reference publishes 158.71 ms for this workload on an RTX 4090 + 9950X
    (its demo application, README.md:9) -- different hardware, context only."""
    x = datagen.alma_sweep_series()
    ctx = vta.HipContext(0)
    plan = vta.DoubleAlmaPlan(ctx, x, vta.AlmaBatchRange(**datagen.DOUBLE_ALMA_FAST),
                              vta.AlmaBatchRange(**datagen.DOUBLE_ALMA_SLOW))
    mean, ks, _ = _time_launches(plan.launch, ctx.stream, steps, warmup, ev, sync=ctx.synchronize)
    st, _ = plan.results()
    plan.close()
    t = mean * 1e-3
    res = {"workload": "double_alma_crossover_58300_pairs_x_200000", "ms_mean": mean,
           "ms_p50": float(np.median(ks)), "ms_p99": float(np.percentile(ks, 99)),
           "pair_bars_per_s": 58_300 * x.size / t,
           "mean_trades_per_pair": float(st[:, 0].mean()),
           "reference_published_ms": 158.71}
    if CPU_EXTRA:
        # the oracle on a bounded sample: every 4th fast x every 4th slow
        # period, their f64 rows (alma_row_scalar order) and the pair walk,
        # OpenMP over rows and over pairs
        import oracle
        xd = x.astype(np.float64)
        fp = list(range(5, 225, 4))
        sp = list(range(30, 295, 4))
        keys = sorted(set(fp) | set(sp))
        ix = {k: i for i, k in enumerate(keys)}
        pf = [ix[a] for a in fp for b in sp]
        ps = [ix[b] for a in fp for b in sp]
        thr = _cpu_threads()

        def run():
            rows = oracle.alma_batch(xd, [(k, 0.85, 6.0) for k in keys], nthreads=thr)
            oracle.cross_pairs(rows, xd, pf, ps, nthreads=thr)

        dt = cpu_median(run, reps=3)
        res["cpu_baseline"] = {
            "value": len(pf) * x.size / dt, "unit": "pair_bars/s", "cores": thr, "kind": "port",
            "sample": f"every 4th fast x every 4th slow period ({len(pf)} of 58,300 pairs, "
                      f"{len(keys)} rows) x {x.size} bars, rows and pair walk, median of 3 "
                      f"runs {dt:.3f} s"}
    return res


def bench_serving(rates=(200, 1000, 2000, 5000), duration_s=4.0, n_slots=8, seed=6):
    """Served mix: Poisson arrivals at each rate for duration_s seconds; each
    request = indicator uniform over {alma, ema, rsi, bollinger_bands, adx},
    series length log-uniform 1k..1M bars sliced from 8 resident-on-host OHLC
    random walks, grid size log-uniform 1..2,000 combos, 30% naming an already
    seen (cached) series.  The client releases each result as soon as it is
    done.  Latency = completion - submission on the server's clock."""
    import queue
    import threading
    pool = datagen.ohlc_pool(8, 1_000_000, seed=5)
    out = {"workload": "served_mix_alma_ema_rsi_bb_adx", "n_slots": n_slots, "rates": []}
    for rate in rates:
        n = int(rate * duration_s)
        specs = datagen.served_requests(n, seed=seed + int(rate))
        rng = np.random.Generator(np.random.PCG64(seed + 7 * int(rate)))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
        srv = vta.IndicatorServer(0, n_slots=n_slots, hbm_budget=96 << 30)
        q = queue.Queue()
        done = []

        def waiter():
            while True:
                rid = q.get()
                if rid is None:
                    return
                r = srv.wait(rid)
                done.append(r)
                srv.release(rid)

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for spec, at in zip(specs, arrivals):
            dt = at - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            q.put(datagen.submit_spec(srv, spec, pool))
        q.put(None)
        th.join()
        st = srv.stats()
        srv.close()
        ok = [r for r in done if r.ok]
        lat = np.array([r.latency_s for r in ok]) * 1e3
        span = max(r.t_done for r in ok) - min(r.t_submit for r in ok)
        vals = sum(r.n_values for r in ok)
        out["rates"].append({
            "offered_req_per_s": rate, "requests": n, "failed": n - len(ok),
            "latency_ms_p90": float(np.percentile(lat, 90)),
            "achieved_req_per_s": len(ok) / span, "latency_ms_p50": float(np.percentile(lat, 50)),
            "latency_ms_p99": float(np.percentile(lat, 99)), "outputs_per_s": vals / span,
            "peak_in_flight": st["peak_in_flight"], "peak_hbm_in_use_GB": st["peak_in_use"] / 1e9,
            "packs": st["packs"], "packed_requests": st["packed_requests"]})
        print(f"[bench] serving {rate} req/s: {out['rates'][-1]}", file=sys.stderr, flush=True)
    return out


def bench_alma_host_boundary(reps=3):
    """The sweep through the host-buffer entry points, wall clock: alma_batch_dev
    (host prices in, device-resident output allocated by the call) and
    alma_batch_into_host_f32 (output copied back over PCIe into pageable host
    memory)."""
    x = datagen.alma_sweep_series()
    sweep = vta.AlmaBatchRange(**datagen.ALMA_SWEEP_RANGE)
    alma = vta.CudaAlma(0)
    n = len(vta.expand_grid(sweep)) * x.size
    host = np.empty(n, dtype=np.float32)
    host[:] = 0.0                                   # touch the pages once
    dev_ms, host_ms = [], []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        arr = alma.alma_batch_dev(x, sweep)
        dev_ms.append((time.perf_counter() - t0) * 1e3)
        arr.free()
        t0 = time.perf_counter()
        alma.alma_batch_into_host_f32(x, sweep, host)
        host_ms.append((time.perf_counter() - t0) * 1e3)
    dev_ms, host_ms = dev_ms[1:], host_ms[1:]
    return {"workload": "alma_sweep_200k_x_1250_host_boundary",
            "batch_dev_ms_median": float(np.median(dev_ms)),
            "batch_dev_outputs_per_s": n / (np.median(dev_ms) * 1e-3),
            "into_host_ms_median": float(np.median(host_ms)),
            "into_host_outputs_per_s": n / (np.median(host_ms) * 1e-3)}


def _smi_sample():
    """(socket power W, mean GFX clock MHz over the XCDs) from amd-smi, or None."""
    import subprocess
    try:
        txt = subprocess.run(["amd-smi", "metric", "-g", "0", "--power", "--clock"],
                             capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    power, clocks, in_gfx = None, [], False
    for line in txt.splitlines():
        t = line.strip()
        if t.startswith("SOCKET_POWER:"):
            try:
                power = float(t.split()[1])
            except (IndexError, ValueError):
                pass
        elif t.startswith("GFX_") and t.endswith(":"):
            in_gfx = True
        elif t.startswith("MEM_") and t.endswith(":"):
            in_gfx = False
        elif in_gfx and t.startswith("CLK:"):
            try:
                clocks.append(float(t.split()[1]))
            except (IndexError, ValueError):
                pass
            in_gfx = False
    if power is None or not clocks:
        return None
    return power, sum(clocks) / len(clocks)


def bench_sustained(seconds=6.0, forms=("mfma", "mfma_f16x3", "fp32")):
    """Each sweep form launched back to back for `seconds` (host syncs every 50
    launches), with amd-smi read every half second from the second second on:
    the sustained launch time and the socket power and engine clock it runs at."""
    import threading
    x = datagen.alma_sweep_series()
    sweep = vta.AlmaBatchRange(**datagen.ALMA_SWEEP_RANGE)
    T, C = x.size, len(vta.expand_grid(sweep))
    alma = vta.CudaAlma(0)
    d_x = alma.ctx.to_device(x)
    d_out = alma.ctx.alloc(T * C * 4)
    idle = _smi_sample()
    res = {"workload": "alma_sweep_200k_x_1250", "seconds": seconds,
           "idle": {"socket_W": idle[0], "gfx_MHz": idle[1]} if idle else None, "forms": {}}
    for form in forms:
        plan = alma.sweep_plan(sweep, T, kernel=form)
        plan.launch(d_x.ptr, 0, d_out.ptr)
        alma.synchronize()
        samples, stop = [], threading.Event()

        def sampler():
            if stop.wait(1.0):
                return
            while not stop.is_set():
                v = _smi_sample()
                if v:
                    samples.append(v)
                stop.wait(0.5)

        th = threading.Thread(target=sampler)
        th.start()
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            for _ in range(50):
                plan.launch(d_x.ptr, 0, d_out.ptr)
            alma.synchronize()
            n += 50
        dt = time.perf_counter() - t0
        stop.set()
        th.join()
        plan.close()
        r = {"launches": n, "us_per_launch": dt / n * 1e6, "outputs_per_s": n * C * T / dt,
             "samples": len(samples)}
        if samples:
            r["socket_W_mean"] = float(np.mean([p for p, _ in samples]))
            r["gfx_MHz_mean"] = float(np.mean([c for _, c in samples]))
        res["forms"][form] = r
        print(f"[bench] sustained {form}: {r}", file=sys.stderr, flush=True)
    d_out.free()
    return res


def bench_alma_single(steps, ev):
    """Plain ALMA on one 100,000-bar series (period 9): launch-latency bound."""
    x = datagen.alma_single_series()
    alma = vta.CudaAlma(0, policy="fp32")
    sweep = vta.AlmaBatchRange(period=(9, 9, 0), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))
    plan = alma.sweep_plan(sweep, x.size, kernel="fp32")
    d_x = alma.ctx.to_device(x)
    d_out = alma.ctx.alloc(x.size * 4)
    for _ in range(3):
        plan.launch(d_x.ptr, 0, d_out.ptr)
    alma.synchronize()
    ks = []
    for _ in range(steps):
        a, b = ev.create(), ev.create()
        ev.record(a, alma.ctx.stream)
        plan.launch(d_x.ptr, 0, d_out.ptr)
        ev.record(b, alma.ctx.stream)
        ks.append(ev.elapsed_ms(a, b) * 1e3)
    ks = np.array(ks)
    return {"workload": "alma_single_100k_p9", "kernel_us_p50": float(np.percentile(ks, 50)),
            "kernel_us_p99": float(np.percentile(ks, 99))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and error check")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--forms", default="mfma,mfma_f16x3,fp32")
    ap.add_argument("--settle", type=float, default=1.0,
                    help="seconds of untimed back-to-back launches before each timed "
                         "workload's warm-up (0: none)")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the MFMA sweep kernel variant on this workload before "
                         "timing (default: the built-in variant, which the autotuner's "
                         "shape sweep chose)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary workloads (many-series OHLC, rolling windows, "
                         "single ALMA)")
    ap.add_argument("--mode", default="sweep",
                    choices=["sweep", "calib", "ohlc", "rolling", "cross", "serve", "ma", "rec",
                             "graph", "tune", "pack", "sustain"],
                    help="calib: 256 MiB device-to-device copies, the known-size reference "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    global SETTLE_S, CPU_THREADS, CPU_EXTRA
    SETTLE_S = max(0.0, args.settle)
    CPU_THREADS = max(0, args.cpu_threads)
    CPU_EXTRA = not args.no_cpu
    if args.gpus != 1:
        print(f"bench.py: vectorta_amd runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding the sweep over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    if args.mode == "calib":
        calibrate_copy()
        return
    if args.mode == "sustain":
        print(json.dumps(bench_sustained()))
        return
    if args.mode == "ohlc":
        r = bench_many_series_ohlc(args.steps, args.warmup, HipEvents())
        print(json.dumps(r))
        return
    if args.mode == "serve":
        print(json.dumps(bench_serving()))
        return
    if args.mode == "cross":
        r = bench_double_alma(args.steps, args.warmup, HipEvents())
        print(json.dumps(r))
        return
    if args.mode == "rec":
        r = bench_recurrence_batches(args.steps, args.warmup, HipEvents())
        print(json.dumps(r))
        return
    if args.mode == "pack":
        r = bench_packed_batch(args.steps, args.warmup, HipEvents())
        print(json.dumps(r))
        return
    if args.mode == "tune":
        r = bench_tune_shapes(5)
        print(json.dumps(r))
        return
    if args.mode == "graph":
        r = bench_captured_batch(args.steps, args.warmup, HipEvents())
        print(json.dumps(r))
        return
    if args.mode == "ma":
        r = bench_moving_averages(args.steps, args.warmup, HipEvents())
        print(json.dumps(r))
        return
    if args.mode == "rolling":
        r = bench_rolling(args.steps, args.warmup, HipEvents())
        print(json.dumps(r))
        return

    x = datagen.alma_sweep_series()
    sweep = vta.AlmaBatchRange(**datagen.ALMA_SWEEP_RANGE)
    combos = vta.expand_grid(sweep)
    T, C = x.size, len(combos)
    alg_bytes, alg_flops = algorithmic_work(combos, T)

    alma = vta.CudaAlma(0)
    d_x = alma.ctx.to_device(x)
    d_out = alma.ctx.alloc(T * C * 4)
    ev = HipEvents()

    forms = {}
    outputs = {}
    for form in args.forms.split(","):
        plan = alma.sweep_plan(sweep, T, kernel=form)
        tuned = None
        if form == "mfma" and args.tune:
            # kernel variant picked by the autotuner on this workload (untimed)
            from vectorta_amd import autotune
            tuner = autotune.Autotuner(reps=5)
            best = tuner.tune_alma_sweep(plan, d_x.ptr, 0, d_out.ptr)
            tuned = {"knob": "alma_mfma_variant", "value": best,
                     "candidate_ms": next(iter(tuner.table.values()))["ms"]}
            print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
        ms_step, k_mean, ks = time_form(alma, plan, d_x, d_out, args.steps, args.warmup, ev)
        forms[form] = {
            "ms_per_step": ms_step,
            "kernel_ms_mean": k_mean,
            "kernel_ms_p50": float(np.percentile(ks, 50)),
            "kernel_ms_p99": float(np.percentile(ks, 99)),
            "outputs_per_s": C * T / (ms_step * 1e-3),
            "achieved_GBs": alg_bytes / (k_mean * 1e-3) / 1e9,
            "achieved_TFLOPs": alg_flops / (k_mean * 1e-3) / 1e12,
        }
        if tuned:
            forms[form]["autotune"] = tuned
        if not args.no_cpu:
            outputs[form] = d_out.copy_to_host().reshape(C, T)
        plan.close()
        print(f"[bench] {form}: kernel {forms[form]['kernel_ms_mean']:.4f} ms "
              f"(p50 {forms[form]['kernel_ms_p50']:.4f}, p99 {forms[form]['kernel_ms_p99']:.4f}), "
              f"step {ms_step:.4f} ms", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        threads = _cpu_threads()
        ref, dt = cpu_baseline(x, combos, threads)
        cpu = {"value": C * T / dt, "unit": "outputs/s", "cores": threads, "kind": "port",
               "sample": f"full {C}-combo x {T}-bar sweep, f64 C restatement of "
                         f"alma_row_scalar, OpenMP over rows, median of 10 runs {dt:.2f} s"}
        dtv = cpu_baseline_avx2(x, combos, threads)
        if dtv is not None:
            cpu["avx2"] = {"value": C * T / dtv, "unit": "outputs/s", "cores": threads,
                           "sample": f"full sweep, AVX2+FMA restatement of alma_row_avx2 "
                                     f"(the reference's nightly-avx build), median of 10 runs "
                                     f"{dtv:.2f} s"}
        dt1, n1 = cpu_baseline_one_core(x, combos)
        cpu["one_core"] = {"value": n1 * T / dt1, "unit": "outputs/s", "cores": 1,
                           "sample": f"every 10th combo ({n1}) x {T} bars, median of 10 runs "
                                     f"{dt1:.2f} s"}
        for form, g in outputs.items():
            err, nan_ok = max_rel_err(g, ref)
            forms[form]["max_rel_err_vs_f64"] = err
            forms[form]["nan_prefix_exact"] = nan_ok
            print(f"[bench] {form}: max rel err {err:.3e}, NaN prefixes exact: {nan_ok}",
                  file=sys.stderr, flush=True)
        del ref

    def ok(f):
        return f in forms and forms[f].get("max_rel_err_vs_f64", 0.0) <= 1e-5 and \
            forms[f].get("nan_prefix_exact", True)

    head = "mfma" if ok("mfma") else "fp32"
    hf = forms[head]
    variant = (forms.get("mfma") or {}).get("autotune", {}).get(
        "value", builtin_mfma_variant(T, C))
    out_kernel = form_kernel(head, variant)
    for form in forms:
        forms[form]["kernel"] = form_kernel(form, variant if form == "mfma" else 0)
    mf = forms.get("mfma")
    ff = forms.get("fp32")
    out = {
        "metric": METRIC,
        "value": hf["outputs_per_s"],
        "unit": "outputs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_s": SETTLE_S,
        "ms_per_step": hf["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-fp32acc" if head == "mfma" else "fp32",
        "data": "synthetic",
        "config": {"workload": "alma_sweep_200k_x_1250", "series_len": T, "combos": C,
                   "outputs": C * T, "periods": "10..=250 step 10",
                   "offsets": "0.05..=0.95 step 0.1", "sigmas": "2..=10 step 2",
                   "kernel": out_kernel, "parallelism": "single-gpu"},
        "latency_ms": {"p50": hf["kernel_ms_p50"], "p99": hf["kernel_ms_p99"]},
        "roofline": {"bound": "hbm", "achieved": hf["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": hf["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": profiled_traffic("alma_sweep_200k_x1250_traffic.txt",
                                                 out_kernel),
                     "algorithmic_bytes": alg_bytes, "kernel": out_kernel},
        "forms": forms,
        "mfma_speedup_vs_fp32": (ff["kernel_ms_mean"] / mf["kernel_ms_mean"])
        if (mf and ff) else None,
        "cpu_baseline": cpu,
    }
    d_out.free()
    if not args.no_extra:
        extra = {}
        for name, fn in (("many_series_ohlc", lambda: bench_many_series_ohlc(10, 2, ev)),
                         ("rolling_windows", lambda: bench_rolling(5, 1, ev)),
                         ("moving_averages", lambda: bench_moving_averages(5, 1, ev)),
                         ("recurrence_batches", lambda: bench_recurrence_batches(5, 1, ev)),
                         ("alma_host_boundary", lambda: bench_alma_host_boundary()),
                         ("double_alma_crossover", lambda: bench_double_alma(5, 1, ev)),
                         ("serving", lambda: bench_serving()),
                         ("captured_batch", lambda: bench_captured_batch(20, 3, ev)),
                         ("packed_batch", lambda: bench_packed_batch(10, 2, ev)),
                         ("alma_single", lambda: bench_alma_single(20, ev))):
            extra[name] = fn()
            print(f"[bench] {name}: {extra[name]}", file=sys.stderr, flush=True)
        out["extra_workloads"] = extra
    print(json.dumps(out))


if __name__ == "__main__":
    main()
