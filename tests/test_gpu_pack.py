"""Packed one-series batches (vta_rec_pack_*): many series of uneven lengths,
each with its own period grid, in one launch per scan pass.  Every job's
outputs must equal, bit for bit, its own vta_<ind>_batch_device call on the
same context (same chunking, same kernels) -- the "same alone and in the mix"
property -- for all six recurrence engines, with jobs that exercise one-chunk
series (<= 64 bars), a late first_valid, a non-finite bar after first_valid
(the exact path, run once in chunk 0 of the output pass) and grids from one
period to more than one 64-combo group.  One job per indicator is also
checked against the oracle, and a captured replay of a pack against its
eager launch."""
import ctypes

import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import datagen, lib
from vectorta_amd._lib import check
from tests.conftest import require_gpu
from tests.helpers import max_rel_err, nan_mask_equal

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p
KINDS = ("ema", "rsi", "atr", "adx", "smma", "wilders")
HLC = ("atr", "adx")

# (series_len, leading NaN, period grid start/stop/step, interior non-finite bar)
SHAPES = [(50, 0, (2, 10, 4), None),
          (3_000, 0, (2, 60, 1), None),
          (20_000, 1_500, (5, 400, 5), None),
          (61_000, 3, (1, 1_300, 10), None),
          (9_000, 2, (3, 40, 3), 4_000),
          (700, 0, (14, 14, 0), None)]


def _direct(kind, ctx, d, job):
    L, h = lib(), ctx.handle
    T, fv, C, mp = job["series_len"], job["first_valid"], job["n_combos"], job["max_period"]
    out = P(d["direct"].ptr)
    if kind == "ema":
        check(L.vta_ema_batch_device(h, P(d["c"].ptr), P(d["p"].ptr), None, T, fv, C, out))
    elif kind == "rsi":
        check(L.vta_rsi_batch_device(h, P(d["c"].ptr), P(d["p"].ptr), mp, T, fv, C, out))
    elif kind == "smma":
        check(L.vta_smma_batch_device(h, P(d["c"].ptr), P(d["p"].ptr), P(d["w"].ptr), fv, T, C,
                                      out))
    elif kind == "wilders":
        # warm = fv + p - 1 as the reference's host code sets it; no alphas
        # buffer, so the smoothing factor is the f64 1/p the pack uses
        check(L.vta_wilders_batch_device(h, P(d["c"].ptr), P(d["p"].ptr), None, P(d["w"].ptr),
                                         T, fv, C, out))
    else:
        fn = getattr(L, f"vta_{kind}_batch_device")
        check(fn(h, P(d["h"].ptr), P(d["l"].ptr), P(d["c"].ptr), P(d["p"].ptr), mp, T, fv, C,
                 out))


def _oracle(kind, c, h, lo, periods):
    f = [a.astype(np.float64) for a in (h, lo, c)]
    if kind == "ema":
        return O.ema_batch(f[2], periods)
    if kind == "rsi":
        return O.rsi_batch(f[2], periods)
    if kind == "atr":
        return O.atr_batch(*f, periods)
    if kind == "adx":
        return O.adx_batch(*f, periods)
    fn = O.smma if kind == "smma" else O.wilders
    return np.stack([fn(f[2], p) for p in periods])


@pytest.mark.parametrize("kind", KINDS)
def test_pack_matches_direct_calls(kind):
    require_gpu()
    ctx = vta.HipContext(0)
    rng = np.random.Generator(np.random.PCG64(90 + KINDS.index(kind)))
    jobs, bufs, host = [], [], []
    for i, (T, lead, (p0, p1, ps), bad) in enumerate(SHAPES):
        c = datagen.gbm_close(T, 0.01, seed=500 + i)
        _, hi, lo, c = datagen.ohlc_from_close(c, rng)
        for a in (hi, lo, c):
            a[:lead] = np.nan
        if bad is not None:
            c[bad] = np.nan
        periods = np.arange(p0, p1 + 1, ps if ps else 1, dtype=np.int32)
        if kind == "adx":
            periods = np.maximum(periods, 2)
        C = int(periods.size)
        d = {"c": ctx.to_device(c), "h": ctx.to_device(hi), "l": ctx.to_device(lo),
             "p": ctx.to_device(periods), "out": ctx.alloc(C * T * 4),
             "direct": ctx.alloc(C * T * 4),
             "w": ctx.to_device((lead + periods - 1).astype(np.int32))}
        job = {"close": d["c"], "periods": d["p"], "out": d["out"], "series_len": T,
               "first_valid": lead, "n_combos": C, "max_period": int(periods.max())}
        if kind in HLC:
            job["high"], job["low"] = d["h"], d["l"]
        jobs.append(job)
        bufs.append(d)
        host.append((c, hi, lo, [int(p) for p in periods]))
    pack = vta.RecPack(ctx, kind, jobs)
    pack.launch()
    for job, d in zip(jobs, bufs):
        _direct(kind, ctx, d, job)
    ctx.synchronize()
    got = []
    for i, (job, d) in enumerate(zip(jobs, bufs)):
        C, T = job["n_combos"], job["series_len"]
        g = d["out"].copy_to_host(np.float32).reshape(C, T)
        r = d["direct"].copy_to_host(np.float32).reshape(C, T)
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32)), (kind, i, SHAPES[i])
        got.append(g)
    # the oracle on the job with a late first_valid and on the unclean one
    for i in (2, 4):
        c, hi, lo, periods = host[i]
        ref = _oracle(kind, c, hi, lo, periods)
        nan_mask_equal(got[i], ref)
        err = max_rel_err(got[i], ref, floor=1e-3 if kind in ("rsi", "adx") else 0.0)
        print(f"{kind} job {i}: max rel err {err:.3e}")
        assert err <= 1e-5, (kind, i, err)
    # a captured replay of the pack equals its eager launch
    for d in bufs:
        check(lib().vta_memcpy_htod(ctx.handle, P(d["out"].ptr),
                                    np.full(d["out"].nbytes // 4, -3.0, np.float32).ctypes.data_as(P),
                                    d["out"].nbytes))
    ctx.begin_capture()
    pack.launch()
    graph = ctx.end_capture()
    graph.launch()
    ctx.synchronize()
    for i, (job, d) in enumerate(zip(jobs, bufs)):
        g = d["out"].copy_to_host(np.float32).reshape(job["n_combos"], job["series_len"])
        assert np.array_equal(g.view(np.uint32), got[i].view(np.uint32)), ("graph", kind, i)
    graph.destroy()
    pack.close()
    for d in bufs:
        for b in d.values():
            b.free()


def test_pack_rejects_bad_jobs():
    require_gpu()
    ctx = vta.HipContext(0)
    x = ctx.to_device(np.ones(100, np.float32))
    p = ctx.to_device(np.array([5], np.int32))
    o = ctx.alloc(400)
    good = {"close": x, "periods": p, "out": o, "series_len": 100, "first_valid": 0,
            "n_combos": 1, "max_period": 5}
    for bad in ({"first_valid": 100}, {"n_combos": 0}, {"series_len": 0}):
        with pytest.raises(vta.VtaError):
            vta.RecPack(ctx, "ema", [dict(good, **bad)])
    with pytest.raises(vta.VtaError):
        vta.RecPack(ctx, "atr", [good])          # ATR needs high and low
    with pytest.raises(vta.VtaError):
        vta.RecPack(ctx, "macd", [good])
    vta.RecPack(ctx, "ema", [good]).close()


def test_bb_pack_matches_direct_calls():
    """Packed Bollinger Bands batches: every job's three outputs equal its own
    vta_bollinger_bands_batch_device call bit for bit (uneven lengths, a late
    first_valid, a non-finite bar and its poison rule, grids of 1 to 70
    combos); one job also against the oracle."""
    require_gpu()
    ctx = vta.HipContext(0)
    L = lib()
    shapes = [(300, 0, (5, 20, 5), None), (5_000, 0, (2, 60, 1), None),
              (40_000, 900, (10, 200, 10), None), (12_000, 3, (5, 50, 5), 6_000)]
    jobs, bufs, host = [], [], []
    for i, (T, lead, (p0, p1, ps), bad) in enumerate(shapes):
        x = datagen.gbm_close(T, 0.01, seed=700 + i)
        x[:lead] = np.nan
        if bad is not None:
            x[bad] = np.nan
        per = np.arange(p0, p1 + 1, ps, dtype=np.int32)
        C = int(per.size)
        up = np.linspace(1.5, 2.5, C)
        dn = np.full(C, 2.0)
        fnf = int(np.argmax(~np.isfinite(x[lead:]))) + lead if not np.isfinite(x[lead:]).all() \
            else T
        d = {"x": ctx.to_device(x), "p": ctx.to_device(per), "u": ctx.to_device(up),
             "d": ctx.to_device(dn)}
        outs = [ctx.alloc(C * T * 4) for _ in range(6)]
        jobs.append({"prices": d["x"], "periods": d["p"], "devup": d["u"], "devdn": d["d"],
                     "upper": outs[0], "middle": outs[1], "lower": outs[2], "series_len": T,
                     "first_valid": lead, "first_nonfinite": fnf, "n_combos": C,
                     "max_period": int(per.max())})
        bufs.append((d, outs))
        host.append((x, per, up, dn))
    pack = vta.BbPack(ctx, jobs)
    pack.launch()
    for j, (d, outs) in zip(jobs, bufs):
        check(L.vta_bollinger_bands_batch_device(
            ctx.handle, P(d["x"].ptr), j["series_len"], j["first_valid"], j["first_nonfinite"],
            P(d["p"].ptr), P(d["u"].ptr), P(d["d"].ptr), j["n_combos"], j["max_period"],
            P(outs[3].ptr), P(outs[4].ptr), P(outs[5].ptr)))
    ctx.synchronize()
    for i, (j, (d, outs)) in enumerate(zip(jobs, bufs)):
        C, T = j["n_combos"], j["series_len"]
        for o in range(3):
            g = outs[o].copy_to_host(np.float32).reshape(C, T)
            r = outs[3 + o].copy_to_host(np.float32).reshape(C, T)
            assert np.array_equal(g.view(np.uint32), r.view(np.uint32)), (i, o, shapes[i])
    x, per, up, dn = host[2]
    xd = x.astype(np.float64)
    for o, name in enumerate(("upper", "middle", "lower")):
        ref = np.stack([O.bollinger_bands(xd, int(p), float(a), float(b))[o]
                        for p, a, b in zip(per, up, dn)])
        g = bufs[2][1][o].copy_to_host(np.float32).reshape(len(per), -1)
        nan_mask_equal(g, ref)
        assert max_rel_err(g, ref) <= 1e-5, name
    pack.close()
    for d, outs in bufs:
        for b in list(d.values()) + outs:
            b.free()


# (series_len, leading NaN, ALMA grid, interior non-finite bar): uneven lengths
# (partial last tiles, T % 4 != 0), one combo to several 128-combo groups
ALMA_JOBS = [(1_003, 0, dict(period=(9, 9, 0), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0)),
              None),
             (20_000, 37, dict(period=(10, 250, 10), offset=(0.05, 0.95, 0.3),
                               sigma=(2.0, 10.0, 4.0)), None),
             (6_002, 3, dict(period=(5, 40, 5), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0)),
              None),
             (50_000, 0, dict(period=(20, 180, 20), offset=(0.1, 0.9, 0.2),
                              sigma=(3.0, 9.0, 3.0)), 31_000),
             (64, 0, dict(period=(10, 60, 25), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0)),
              None)]


@pytest.mark.parametrize("kernel", ["fp32", "mfma", "mfma_f16x3", "auto"])
def test_alma_pack_matches_own_sweeps(kernel):
    """vta_alma_pack_*: five one-series sweeps (uneven lengths, own grids, a
    late first_valid, an interior NaN) in one launch.  Each job's outputs equal,
    bit for bit, a sweep plan of its own grid on the same kernel (the split
    forms at variant 9, the diagonal form the pack runs), and two jobs are
    checked against the oracle; AUTO takes FP32 here (a job has fewer than 16
    combos)."""
    require_gpu()
    from vectorta_amd import autotune as at
    ctx = vta.HipContext(0)
    alma = vta.CudaAlma(0)
    bufs, jobs, ref_rows = [], [], []
    for i, (T, lead, grid, bad) in enumerate(ALMA_JOBS):
        x = datagen.gbm_close(T, 0.015, seed=80 + i)
        x[:lead] = np.nan
        if bad is not None:
            x[bad] = np.nan
        sw = vta.AlmaBatchRange(**grid)
        C = len(vta.expand_grid(sw))
        d = {"x": x, "sweep": sw, "C": C, "T": T, "fv": lead,
             "prices": ctx.to_device(x), "out": ctx.alloc(C * T * 4),
             "direct": alma.ctx.alloc(C * T * 4)}
        d["prices_a"] = alma.ctx.to_device(x)
        bufs.append(d)
        jobs.append({"prices": d["prices"], "out": d["out"], "sweep": sw, "series_len": T,
                     "first_valid": lead})
    pack = vta.AlmaPack(ctx, jobs, kernel=kernel)
    expect = "fp32" if kernel in ("fp32", "auto") else \
        ("mfma_bf16x3" if kernel == "mfma" else "mfma_f16x3")
    assert pack.kernel == expect
    assert pack.combos == [d["C"] for d in bufs]
    pack.launch()
    ctx.synchronize()
    at.set_tuning(alma.ctx, at.ALMA_MFMA_VARIANT, 9)
    plan_kernel = {"fp32": "fp32", "mfma_bf16x3": "mfma", "mfma_f16x3": "mfma_f16x3"}[expect]
    for i, d in enumerate(bufs):
        plan = alma.sweep_plan(d["sweep"], d["T"], kernel=plan_kernel)
        plan.launch(d["prices_a"].ptr, d["fv"], d["direct"].ptr)
        alma.synchronize()
        g = d["out"].copy_to_host().reshape(d["C"], d["T"])
        r = d["direct"].copy_to_host().reshape(d["C"], d["T"])
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32)), (kernel, i)
        if i in (1, 3):
            combos = [(c.period, c.offset, c.sigma) for c in vta.expand_grid(d["sweep"])]
            ref = O.alma_batch(d["x"].astype(np.float64), combos)
            nan_mask_equal(g, ref)
            err = max_rel_err(g, ref)
            print(f"alma pack {kernel} job {i}: max rel err {err:.3e}")
            assert err <= 1e-5
        plan.close()
    at.set_tuning(alma.ctx, at.ALMA_MFMA_VARIANT, 0)
    pack.close()
    for d in bufs:
        for k in ("prices", "out", "direct", "prices_a"):
            d[k].free()


def test_alma_pack_rejects_bad_jobs():
    """Pack creation validates every job like alma_batch_from_device_prices:
    null buffers, first_valid past the series, a period longer than the valid
    data; and the explicit MFMA forms refuse periods above 256."""
    require_gpu()
    ctx = vta.HipContext(0)
    x = datagen.gbm_close(1_000, 0.015, seed=90)
    dx, do = ctx.to_device(x), ctx.alloc(1_000 * 4 * 4)
    ok = {"prices": dx, "out": do, "series_len": 1_000, "first_valid": 0,
          "sweep": vta.AlmaBatchRange(period=(9, 12, 1))}
    vta.AlmaPack(ctx, [ok]).close()
    for bad in ({"first_valid": 1_000}, {"sweep": vta.AlmaBatchRange(period=(2_000, 2_000, 0))}):
        with pytest.raises(vta.VtaError):
            vta.AlmaPack(ctx, [dict(ok, **bad)])
    long = dict(ok, sweep=vta.AlmaBatchRange(period=(300, 300, 0)))
    vta.AlmaPack(ctx, [long], kernel="fp32").close()
    with pytest.raises(vta.VtaError):
        vta.AlmaPack(ctx, [long], kernel="mfma_f16x3")
    dx.free()
    do.free()
