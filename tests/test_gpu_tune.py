"""Launch-shape knobs and the autotuner on the GPU: every knob value gives the
same results (bit-identical where the knob only repartitions work, within the
parity bound of the f64 oracle where it changes the summation), the knob API
rejects out-of-range values, and the autotuner leaves the context at one of
its candidates with outputs still right."""
import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import autotune as at
from vectorta_amd import datagen
from tests.conftest import require_gpu
from tests.helpers import max_rel_err, nan_mask_equal

pytestmark = pytest.mark.gpu


def test_knob_range_checks():
    require_gpu()
    ctx = vta.HipContext(0)
    assert at.get_tuning(ctx, at.SCAN_WAVES) == 0
    for knob, bad in ((at.ALMA_MFMA_VARIANT, 12), (at.SCAN_WAVES, 32), (at.FIR_GROUP, 257),
                      (at.MS_CHUNK, 63), (at.ROLL_FORM, 3), (9, 1)):
        with pytest.raises(vta.VtaError):
            at.set_tuning(ctx, knob, bad)
    at.set_tuning(ctx, at.FIR_GROUP, 32)
    assert at.get_tuning(ctx, at.FIR_GROUP) == 32
    at.set_tuning(ctx, at.FIR_GROUP, 0)


def test_alma_mfma_variants_and_autotune():
    """All eleven split-BF16 kernel variants meet 1e-5 against the oracle with
    exact NaN prefixes; the autotuner picks one of them."""
    require_gpu()
    T = 20_000
    x = datagen.gbm_close(T, 0.015, seed=61)
    x[:3] = np.nan
    sweep = vta.AlmaBatchRange(period=(10, 250, 40), offset=(0.05, 0.95, 0.3),
                               sigma=(2.0, 10.0, 4.0))
    combos = [(c.period, c.offset, c.sigma) for c in vta.expand_grid(sweep)]
    ref = O.alma_batch(x.astype(np.float64), combos)
    alma = vta.CudaAlma(0)
    ctx = alma.ctx
    plan = vta.AlmaSweepPlan(alma, sweep, T, kernel="mfma")
    d_x = ctx.to_device(x)
    d_o = ctx.alloc(plan.n_combos * T * 4)
    for v in range(1, 12):
        at.set_tuning(ctx, at.ALMA_MFMA_VARIANT, v)
        plan.launch(d_x.ptr, 3, d_o.ptr)
        ctx.synchronize()
        g = d_o.copy_to_host(np.float32).reshape(plan.n_combos, T)
        nan_mask_equal(g, ref)
        err = max_rel_err(g, ref)
        print(f"variant {v}: max rel err {err:.3e}")
        assert err <= 1e-5, (v, err)
    tuner = at.Autotuner(reps=3)
    best = tuner.tune_alma_sweep(plan, d_x.ptr, 3, d_o.ptr)
    print(f"autotuned variant {best}: {tuner.table}")
    assert best in at.CANDIDATES[at.ALMA_MFMA_VARIANT]
    assert at.get_tuning(ctx, at.ALMA_MFMA_VARIANT) == best
    # a second tune of the same shape is a cache hit (no measurement)
    at.set_tuning(ctx, at.ALMA_MFMA_VARIANT, 0)
    assert tuner.tune_alma_sweep(plan, d_x.ptr, 3, d_o.ptr) == best
    assert at.get_tuning(ctx, at.ALMA_MFMA_VARIANT) == best
    plan.close()
    d_x.free()
    d_o.free()


@pytest.mark.parametrize("kernel", ["mfma", "mfma_f16x3"])
def test_sweep_diagonal_variants_ragged(kernel):
    """Diagonal-form variants 8-11 on ragged series (a partial last tile, T % 4
    == 0 and != 0, 37 leading NaNs, one interior NaN): each within 1e-5 of the
    oracle with exact NaN prefixes; 9 (direct stores) and 11 (deferred stores)
    share 8's arithmetic and are bit-identical to it (11 runs as 8 when
    T % 4 != 0)."""
    require_gpu()
    sweep = vta.AlmaBatchRange(period=(10, 250, 30), offset=(0.05, 0.95, 0.3),
                               sigma=(2.0, 10.0, 4.0))
    combos = [(c.period, c.offset, c.sigma) for c in vta.expand_grid(sweep)]
    alma = vta.CudaAlma(0)
    ctx = alma.ctx
    for T in (10_004, 10_003):
        x = datagen.gbm_close(T, 0.015, seed=62)
        x[:37] = np.nan
        x[5000] = np.nan
        ref = O.alma_batch(x.astype(np.float64), combos)
        plan = vta.AlmaSweepPlan(alma, sweep, T, kernel=kernel)
        d_x = ctx.to_device(x)
        d_o = ctx.alloc(plan.n_combos * T * 4)
        outs = {}
        for v in (8, 9, 10, 11):
            at.set_tuning(ctx, at.ALMA_MFMA_VARIANT, v)
            plan.launch(d_x.ptr, 37, d_o.ptr)
            ctx.synchronize()
            g = d_o.copy_to_host(np.float32).reshape(plan.n_combos, T)
            nan_mask_equal(g, ref)
            err = max_rel_err(g, ref)
            print(f"{kernel} T={T} variant {v}: max rel err {err:.3e}")
            assert err <= 1e-5, (kernel, T, v, err)
            outs[v] = g
        for v in (9, 11):
            assert np.array_equal(outs[8].view(np.uint32), outs[v].view(np.uint32)), (T, v)
        at.set_tuning(ctx, at.ALMA_MFMA_VARIANT, 0)
        plan.close()
        d_x.free()
        d_o.free()


def test_scan_waves_and_ms_chunk_autotune():
    """Recurrence knobs: the chunked-scan wave target (one-series batch) and the
    many-series chunk length, each autotuned, outputs within 1e-5 of the
    oracle afterwards."""
    require_gpu()
    ema = vta.CudaEma(0)
    ctx = ema.ctx
    T = 60_000
    x = datagen.gbm_close(T, 0.01, seed=62)
    periods = np.arange(2, 402, 4, dtype=np.int32)
    d_x, d_p = ctx.to_device(x), ctx.to_device(periods)
    d_o = ctx.alloc(periods.size * T * 4)
    run = lambda: ema.ema_batch_device(d_x, d_p, None, T, 0, int(periods.size), d_o)
    best, times = at.measure(ctx, at.SCAN_WAVES, at.CANDIDATES[at.SCAN_WAVES], run, reps=3)
    print(f"scan waves {best}: {times}")
    assert best in at.CANDIDATES[at.SCAN_WAVES] and all(v is not None for v in times.values())
    run()
    ctx.synchronize()
    g = d_o.copy_to_host(np.float32).reshape(periods.size, T)
    ref = O.ema_batch(x.astype(np.float64), [int(p) for p in periods])
    nan_mask_equal(g, ref)
    assert max_rel_err(g, ref) <= 1e-5

    rows, cols = 8_000, 256
    rng = np.random.Generator(np.random.PCG64(63))
    tm = np.stack([datagen.gbm_close(rows, 0.01, seed=int(s)) for s in rng.integers(1 << 30,
                                                                                     size=cols)],
                  axis=1).astype(np.float32)
    tm[: rows // 3, ::5] = np.nan
    ref = np.stack([O.ema(tm[:, j].astype(np.float64), 20) for j in range(cols)], axis=1)
    for chunk in at.CANDIDATES[at.MS_CHUNK]:
        at.set_tuning(ctx, at.MS_CHUNK, chunk)
        g = ema.ema_many_series_one_param_time_major_dev(tm, cols, rows, 20).to_host()
        nan_mask_equal(g, ref)
        assert max_rel_err(g, ref) <= 1e-5, chunk
    for b in (d_x, d_p, d_o):
        b.free()


def test_fir_group_bit_identical():
    """Combos per workgroup only repartitions the FIR batch: every group size
    gives the same bits."""
    require_gpu()
    sma = vta.CudaSma(0)
    x = datagen.gbm_close(30_000, 0.01, seed=64)
    outs = []
    for grp in at.CANDIDATES[at.FIR_GROUP]:
        at.set_tuning(sma.ctx, at.FIR_GROUP, grp)
        outs.append(sma.sma_batch_dev(x, (2, 300, 3)).to_host())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
