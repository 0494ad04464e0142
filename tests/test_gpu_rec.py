"""EMA / RSI / ATR / ADX parity on the GPU (batch, many-series, fused).

Bar: NaN warm-up prefixes and poison rules bit-exact; values within 1e-5
relative of the f64 oracle on the same f32 inputs, for every indicator.  RSI
and ADX live in [0, 100] and can be exactly 0 (a run of losses, no
directional movement), where a relative error is undefined: there the
denominator is floored at 1e-3, so a value below 1e-3 must lie within 1e-8 of
the oracle.  Long series exercise the chunked burn-in path; series with
interior NaN/Inf exercise the exact path.
"""
import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import datagen
from tests.conftest import require_gpu
from tests.helpers import max_rel_err, nan_mask_equal

pytestmark = pytest.mark.gpu

FLOOR = {"ema": 0.0, "rsi": 1e-3, "atr": 0.0, "adx": 1e-3}


@pytest.fixture(scope="module")
def ctxs():
    require_gpu()
    return {"ema": vta.CudaEma(0), "rsi": vta.CudaRsi(0), "atr": vta.CudaAtr(0),
            "adx": vta.CudaAdx(0)}


def _hlc(n, seed, lead=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    c = datagen.gbm_close(n, 0.01, seed)
    _, h, lo, c = datagen.ohlc_from_close(c, rng)
    for a in (h, lo, c):
        a[:lead] = np.nan
    return h, lo, c


def _check(g, ref, name, bound=1e-5):
    nan_mask_equal(g, ref)
    err = max_rel_err(g, ref, floor=FLOOR[name])
    print(f"{name}: max rel err {err:.3e}")
    assert err <= bound, (name, err)


@pytest.mark.parametrize("n,period", [(3_000, (5, 60, 11)), (300_000, (9, 250, 60))])
def test_ema_batch(ctxs, n, period):
    x = datagen.gbm_close(n, 0.01, seed=31)
    x[:7] = np.nan
    ps = ctxs["ema"].expand_grid(period)
    g = ctxs["ema"].ema_batch_dev(x, period).to_host()
    _check(g, O.ema_batch(x.astype(np.float64), ps), "ema")


@pytest.mark.parametrize("n,period", [(3_000, (2, 60, 11)), (300_000, (14, 250, 59))])
def test_rsi_batch(ctxs, n, period):
    x = datagen.gbm_close(n, 0.01, seed=32)
    x[:4] = np.nan
    ps = ctxs["rsi"].expand_grid(period)
    g = ctxs["rsi"].rsi_batch_dev(x, period).to_host()
    _check(g, O.rsi_batch(x.astype(np.float64), ps), "rsi")


@pytest.mark.parametrize("bad", [None, np.nan])
def test_adx_batch_close_periods(ctxs, bad):
    """Waves of 64 consecutive periods (300..427): every lane of a wave is in
    its DX seed (p < i < 2p-1) for whole 64-bar blocks, and past its
    first-stage seed early in chunk 0, so the DX-seed and state-only blocks
    run; with an interior NaN (poison) the exact path replays every chunk."""
    n = 50_000 if bad is None else 12_000
    h, lo, c = _hlc(n, 35, lead=5)
    if bad is not None:
        h[9_000] = bad
    period = (300, 427, 1)
    ps = ctxs["adx"].expand_grid(period)
    assert len(ps) == 128
    g = ctxs["adx"].adx_batch_dev(h, lo, c, period).to_host()
    f = [a.astype(np.float64) for a in (h, lo, c)]
    _check(g, O.adx_batch(*f, ps), "adx")


@pytest.mark.parametrize("name", ["ema", "rsi", "atr", "adx", "smma", "wilders"])
def test_scan_short_chunks_long_lead(ctxs, monkeypatch, name):
    """Chunked scan with 256-bar chunks over a series whose first 3,000 bars
    are NaN: eleven whole chunks end before first_valid (their chunk-end
    states are undefined and must not reach the fold), and the warm-ups of
    periods up to 1,200 (ADX: 2,400 bars) span several chunks, so carried
    chunks start inside the seed sums, at the step that takes their mean and
    inside ADX's DX seed."""
    monkeypatch.setenv("VTA_SCAN_CHUNK", "256")
    h, lo, c = _hlc(40_000, 36, lead=3_000)
    period = (1, 1_201, 100) if name in ("atr", "smma", "wilders") else (2, 1_202, 100)
    if name in ("ema", "rsi", "atr", "adx"):
        op = ctxs[name]
        ps = op.expand_grid(period)
        if name in ("atr", "adx"):
            g = getattr(op, f"{name}_batch_dev")(h, lo, c, period).to_host()
            ref = getattr(O, f"{name}_batch")(*[a.astype(np.float64) for a in (h, lo, c)], ps)
        else:
            g = getattr(op, f"{name}_batch_dev")(c, period).to_host()
            ref = getattr(O, f"{name}_batch")(c.astype(np.float64), ps)
        _check(g, ref, name)
        return
    op = vta.CudaSmma(0) if name == "smma" else vta.CudaWilders(0)
    ps = op.expand_grid(period)
    g = getattr(op, f"{name}_batch_dev")(c, period).to_host()
    f = O.smma if name == "smma" else O.wilders
    ref = np.stack([f(c.astype(np.float64), p) for p in ps])
    nan_mask_equal(g, ref)
    err = max_rel_err(g, ref, floor=0.0)
    print(f"{name}: max rel err {err:.3e}")
    assert err <= 1e-5, (name, err)


@pytest.mark.parametrize("n,period", [(3_000, (1, 60, 11)), (300_000, (14, 250, 59))])
def test_atr_batch(ctxs, n, period):
    h, lo, c = _hlc(n, 33, lead=3)
    ps = ctxs["atr"].expand_grid(period)
    g = ctxs["atr"].atr_batch_dev(h, lo, c, period).to_host()
    f = [a.astype(np.float64) for a in (h, lo, c)]
    _check(g, O.atr_batch(*f, ps), "atr")


@pytest.mark.parametrize("n,period", [(3_000, (5, 50, 9)), (200_000, (14, 200, 62))])
def test_adx_batch(ctxs, n, period):
    h, lo, c = _hlc(n, 34, lead=2)
    ps = ctxs["adx"].expand_grid(period)
    g = ctxs["adx"].adx_batch_dev(h, lo, c, period).to_host()
    f = [a.astype(np.float64) for a in (h, lo, c)]
    _check(g, O.adx_batch(*f, ps), "adx")


def test_batch_unclean_series_exact_path(ctxs):
    """Interior NaN/Inf: EMA skips them, RSI poisons only from the seed, ATR
    poisons forever -- the series runs exactly from first_valid."""
    n = 20_000
    x = datagen.gbm_close(n, 0.01, seed=35)
    x[:3] = np.nan
    x[9_000] = np.nan
    x[15_000] = np.inf
    ps = (5, 40, 7)
    g = ctxs["ema"].ema_batch_dev(x, ps).to_host()
    _check(g, O.ema_batch(x.astype(np.float64), ctxs["ema"].expand_grid(ps)), "ema")
    g = ctxs["rsi"].rsi_batch_dev(x, ps).to_host()
    ref = O.rsi_batch(x.astype(np.float64), ctxs["rsi"].expand_grid(ps))
    nan_mask_equal(g, ref)
    y = x.copy()
    y[6] = np.nan                                  # inside every seed window
    g = ctxs["rsi"].rsi_batch_dev(y, ps).to_host()
    assert np.isnan(g).all()
    h, lo, c = _hlc(n, 36, lead=2)
    h[12_000] = np.nan
    g = ctxs["atr"].atr_batch_dev(h, lo, c, ps).to_host()
    f = [a.astype(np.float64) for a in (h, lo, c)]
    ref = O.atr_batch(*f, ctxs["atr"].expand_grid(ps))
    nan_mask_equal(g, ref)
    assert np.isnan(g[:, 12_000:]).all()


def _ms_data(cols, rows, seed):
    o, h, lo, c, lens = datagen.many_series_ohlc(cols, rows, min_len=max(64, rows // 20), seed=seed)
    return h, lo, c


@pytest.mark.parametrize("rows", [2_000, 30_000])
def test_many_series_ema_rsi(ctxs, rows):
    h, lo, c = _ms_data(130, rows, 40)
    cols = c.shape[1]
    for name, p in (("ema", 20), ("rsi", 14)):
        fn = getattr(ctxs[name], f"{name}_many_series_one_param_time_major_dev")
        g = fn(c, cols, rows, p)
        assert (g.rows, g.cols) == (rows, cols)
        ref, st = O.many_series(name, c.astype(np.float64), p)
        assert (st == 0).all()
        _check(g.to_host(), ref, name)


@pytest.mark.parametrize("rows", [2_000, 30_000])
def test_many_series_atr_adx(ctxs, rows):
    h, lo, c = _ms_data(70, rows, 41)
    cols = c.shape[1]
    f = [a.astype(np.float64) for a in (h, lo, c)]
    for name, p in (("atr", 14), ("adx", 14)):
        fn = getattr(ctxs[name], f"{name}_many_series_one_param_time_major_dev")
        g = fn(h, lo, c, cols, rows, p).to_host()
        ref, st = O.many_series_hlc(name, *f, p)
        assert (st == 0).all()
        _check(g, ref, name)


def test_fused_ema_rsi_atr_matches_separate_oracles(ctxs):
    rows, cols = 20_000, 200
    h, lo, c = _ms_data(cols, rows, 42)
    # a few unclean series: interior NaN close, interior NaN high
    c[15_000, 3] = np.nan
    h[17_000, 5] = np.nan
    ctx = ctxs["ema"].ctx
    e, r, a = (d.to_host() for d in vta.ema_rsi_atr_many_series_dev(ctx, c, h, lo, cols, rows,
                                                                    20, 14, 14))
    f = [x.astype(np.float64) for x in (h, lo, c)]
    _check(e, O.many_series("ema", f[2], 20)[0], "ema")
    _check(r, O.many_series("rsi", f[2], 14)[0], "rsi")
    _check(a, O.many_series_hlc("atr", *f, 14)[0], "atr")
    assert np.isnan(a[17_000:, 5]).all()


def test_fused_device_plan_config3_shape(ctxs):
    """The 4,096-ticker x 50,000-bar workload shape: enqueue-only plan launch;
    every series' NaN prefix checked, 64 series against the oracle."""
    ctx = ctxs["ema"].ctx
    o, h, lo, c, lens = datagen.many_series_ohlc(4096, 50_000, seed=3)
    plan = vta.ManySeriesOhlcPlan(ctx, c, h, lo)
    plan.launch(20, 14, 14)
    ctx.synchronize()
    e, r, a = plan.outputs()
    rows = c.shape[0]
    fv = rows - lens
    for out, warm in ((e, fv), (r, fv + 14), (a, fv + 13)):
        lead = np.argmax(~np.isnan(out), axis=0)
        assert np.array_equal(lead, warm)
    sel = np.linspace(0, 4095, 64).astype(int)
    f = [x[:, sel].astype(np.float64) for x in (h, lo, c)]
    _check(e[:, sel], O.many_series("ema", f[2], 20)[0], "ema")
    _check(r[:, sel], O.many_series("rsi", f[2], 14)[0], "rsi")
    _check(a[:, sel], O.many_series_hlc("atr", *f, 14)[0], "atr")


def test_device_entry_points(ctxs):
    x = datagen.gbm_close(50_000, 0.01, seed=43)
    x[:10] = np.nan
    ctx = ctxs["ema"].ctx
    ps = np.arange(5, 200, 7, dtype=np.int32)
    d_x, d_p = ctx.to_device(x), ctx.to_device(ps)
    d_out = ctx.alloc(ps.size * x.size * 4)
    ctxs["ema"].ema_batch_device(d_x, d_p, int(ps.max()), x.size, 10, ps.size, d_out)
    ctx.synchronize()
    g = d_out.copy_to_host().reshape(ps.size, x.size)
    _check(g, O.ema_batch(x.astype(np.float64), ps.tolist()), "ema")
    # many-series device form with and without clean flags
    h, lo, c = _ms_data(96, 10_000, 44)
    rows, cols = c.shape
    d_c = ctx.to_device(c)
    fv = np.argmax(~np.isnan(c), axis=0).astype(np.int32)
    d_fv = ctx.to_device(fv)
    d_o = ctx.alloc(rows * cols * 4)
    ref = O.many_series("ema", c.astype(np.float64), 20)[0]
    ctxs["ema"].ema_many_series_one_param_device(d_c, d_fv, 20, cols, rows, d_o, None)
    ctx.synchronize()
    _check(d_o.copy_to_host().reshape(rows, cols), ref, "ema")
    d_cl = ctx.to_device(np.ones(cols, dtype=np.int32))
    ctxs["ema"].ema_many_series_one_param_device(d_c, d_fv, 20, cols, rows, d_o, d_cl)
    ctx.synchronize()
    _check(d_o.copy_to_host().reshape(rows, cols), ref, "ema")


def test_errors(ctxs):
    with pytest.raises(vta.VtaError):
        ctxs["ema"].ema_batch_dev(np.full(10, np.nan, dtype=np.float32), (5, 5, 0))
    with pytest.raises(vta.VtaError):
        ctxs["rsi"].rsi_batch_dev(np.ones(10, dtype=np.float32), (20, 20, 0))
    with pytest.raises(vta.VtaError):
        ctxs["ema"].ema_batch_dev(np.ones(10, dtype=np.float32), (0, 0, 0))


def test_tr_series_and_atr_batch_with_tr(ctxs):
    """tr_from_hlc_device + atr_batch_device_with_tr (atr_wrapper.rs:566, 626):
    the TR series is f32(exact TR) bit for bit; ATR over it matches the ATR
    oracle of the same high/low/close within the ATR bound (the TR is rounded
    to f32 once, ~6e-8 relative, before the averaging)."""
    h, lo, c = _hlc(60_000, 41, lead=9)
    op = ctxs["atr"]
    fv = 9
    n = c.size
    d_h, d_l, d_c = (op.ctx.to_device(a) for a in (h, lo, c))
    d_tr = op.ctx.alloc(n * 4)
    op.tr_from_hlc_device(d_h, d_l, d_c, n, fv, d_tr)
    op.synchronize()
    tr = d_tr.copy_to_host(np.float32)
    hd, ld, cd = (a.astype(np.float64) for a in (h, lo, c))
    ref = np.zeros(n)
    ref[fv] = hd[fv] - ld[fv]
    pc = cd[fv:-1]
    ref[fv + 1:] = np.maximum.reduce([hd[fv + 1:] - ld[fv + 1:], np.abs(hd[fv + 1:] - pc),
                                      np.abs(ld[fv + 1:] - pc)])
    assert np.array_equal(tr.view(np.uint32), ref.astype(np.float32).view(np.uint32))
    ps = np.asarray(op.expand_grid((3, 200, 13)), dtype=np.int32)
    d_p = op.ctx.to_device(ps)
    d_o = op.ctx.alloc(ps.size * n * 4)
    op.atr_batch_device_with_tr(d_tr, d_p, int(ps.max()), n, fv, ps.size, d_o)
    op.synchronize()
    g = d_o.copy_to_host(np.float32).reshape(ps.size, n)
    _check(g, np.stack([O.atr(hd, ld, cd, int(p)) for p in ps]), "atr")
