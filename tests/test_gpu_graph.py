"""Captured batches: enqueue-only calls recorded with vta_ctx_begin_capture /
vta_ctx_end_capture and replayed as one hipGraph launch give the same bits as
the same calls launched eagerly, on every replay.

The batch mixes the launch shapes of the library: the chunked recurrence scan
(EMA, RSI; ADX in two rounds), whose passes share the context's scratch and
carry buffers in stream order, the FIR prefix kernel (SMA), Bollinger Bands and
the split-BF16 ALMA sweep plan.  Replays are separated by eager work on the
same context and by host synchronisations, and their outputs are cleared to a
sentinel before each replay, so a replay that skipped or repeated a node would
show.  The capture contains kernel nodes only (the enqueue-only forms clear
their flags with a fill kernel, not a memset).
"""
import ctypes

import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import datagen, lib
from vectorta_amd._lib import check
from tests.conftest import require_gpu

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def _batch(ctx, alma_plan, d, outs, T, fv):
    """Enqueues the whole batch on ctx's stream (no sync, no allocation)."""
    L = lib()
    h = ctx.handle
    check(L.vta_ema_batch_device(h, P(d["c"].ptr), P(d["p"].ptr), None, T, fv, d["C"],
                                 P(outs["ema"].ptr)))
    check(L.vta_rsi_batch_device(h, P(d["c"].ptr), P(d["p"].ptr), d["maxp"], T, fv, d["C"],
                                 P(outs["rsi"].ptr)))
    check(L.vta_adx_batch_device(h, P(d["h"].ptr), P(d["l"].ptr), P(d["c"].ptr), P(d["p"].ptr),
                                 d["maxp"], T, fv, d["C"], P(outs["adx"].ptr)))
    check(L.vta_sma_batch_device(h, P(d["c"].ptr), P(d["p"].ptr), T, d["C"], fv, d["maxp"],
                                 P(outs["sma"].ptr)))
    check(L.vta_bollinger_bands_batch_device(
        h, P(d["c"].ptr), T, fv, T, P(d["bp"].ptr), P(d["bu"].ptr), P(d["bd"].ptr), d["BC"],
        d["bmaxp"], P(outs["bb_up"].ptr), P(outs["bb_mid"].ptr), P(outs["bb_lo"].ptr)))
    alma_plan.launch(d["c"].ptr, fv, outs["alma"].ptr)


def test_graph_replays_match_eager():
    require_gpu()
    T, fv = 40_000, 5
    rng = np.random.Generator(np.random.PCG64(41))
    c = datagen.gbm_close(T, 0.01, seed=41)
    _, hi, lo, c = datagen.ohlc_from_close(c, rng)
    for a in (hi, lo, c):
        a[:fv] = np.nan
    periods = np.arange(2, 130, 2, dtype=np.int32)          # 64 combos
    bper = np.array([10, 20, 40], dtype=np.int32)
    alma = vta.CudaAlma(0)
    ctx = alma.ctx
    sweep = vta.AlmaBatchRange(period=(10, 90, 20), offset=(0.25, 0.85, 0.3),
                               sigma=(3.0, 6.0, 3.0))
    plan = vta.AlmaSweepPlan(alma, sweep, T, kernel="mfma")
    d = {"c": ctx.to_device(c), "h": ctx.to_device(hi), "l": ctx.to_device(lo),
         "p": ctx.to_device(periods), "maxp": int(periods.max()), "C": int(periods.size),
         "bp": ctx.to_device(bper), "bu": ctx.to_device(np.full(3, 2.0)),
         "bd": ctx.to_device(np.full(3, 1.5)), "BC": 3, "bmaxp": int(bper.max())}
    rows = {"ema": d["C"], "rsi": d["C"], "adx": d["C"], "sma": d["C"], "bb_up": 3,
            "bb_mid": 3, "bb_lo": 3, "alma": plan.n_combos}
    eager = {k: ctx.alloc(r * T * 4) for k, r in rows.items()}
    graph_out = {k: ctx.alloc(r * T * 4) for k, r in rows.items()}

    _batch(ctx, plan, d, eager, T, fv)
    ctx.synchronize()
    ref = {k: b.copy_to_host(np.float32).reshape(rows[k], T) for k, b in eager.items()}
    # the eager batch itself is right (one indicator against the oracle)
    ema_ref = O.ema_batch(c.astype(np.float64), [int(p) for p in periods])
    m = np.isfinite(ema_ref)
    assert np.array_equal(np.isnan(ref["ema"]), ~m)
    assert np.max(np.abs(ref["ema"][m] - ema_ref[m]) / np.abs(ema_ref[m])) < 1e-5

    ctx.begin_capture()
    _batch(ctx, plan, d, graph_out, T, fv)
    g = ctx.end_capture()
    print(f"graph nodes: {g.num_nodes}")
    assert g.num_nodes >= 10

    sentinel = np.full(max(rows.values()) * T, -7.0, dtype=np.float32)
    for rep in range(3):
        for k, b in graph_out.items():
            check(lib().vta_memcpy_htod(ctx.handle, P(b.ptr),
                                        sentinel.ctypes.data_as(P), rows[k] * T * 4))
        g.launch()
        ctx.synchronize()
        for k, b in graph_out.items():
            got = b.copy_to_host(np.float32).reshape(rows[k], T)
            assert np.array_equal(got.view(np.uint32), ref[k].view(np.uint32)), (rep, k)
        # eager work on the same context between replays
        _batch(ctx, plan, d, eager, T, fv)
        ctx.synchronize()
        for k, b in eager.items():
            got = b.copy_to_host(np.float32).reshape(rows[k], T)
            assert np.array_equal(got.view(np.uint32), ref[k].view(np.uint32)), ("eager", rep, k)
    g.destroy()
    plan.close()
    for b in list(eager.values()) + list(graph_out.values()):
        b.free()


def test_capture_state_errors():
    require_gpu()
    ctx = vta.HipContext(0)
    with pytest.raises(vta.VtaError):
        ctx.end_capture()                 # not capturing
    ctx.begin_capture()
    with pytest.raises(vta.VtaError):
        ctx.begin_capture()               # already capturing
    g = ctx.end_capture()                 # empty capture is a valid (empty) graph
    assert g.num_nodes == 0
    g.launch()
    ctx.synchronize()
    g.destroy()
