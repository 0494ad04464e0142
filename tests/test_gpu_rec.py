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


def test_wilders_seeded_matches_replay(monkeypatch):
    """Wilders batches of up to 4,096 combos start each exact lane from a
    per-combo seed at its last warm-up bar; larger ones replay the warm-up
    from first_valid.  Both are the same arithmetic, so with one chunk length
    the 4,096 rows of a seeded batch equal bit for bit those of a 4,097-combo
    batch over the same periods, and a sample of rows matches the oracle."""
    require_gpu()
    monkeypatch.setenv("VTA_SCAN_CHUNK", "256")
    _, _, c = _hlc(9_000, 37, lead=700)
    op = vta.CudaWilders(0)
    seeded = op.wilders_batch_dev(c, (1, 4_096, 1)).to_host()
    replay = op.wilders_batch_dev(c, (1, 4_097, 1)).to_host()
    assert seeded.shape == (4_096, 9_000) and replay.shape == (4_097, 9_000)
    assert np.array_equal(seeded.view(np.uint32), replay[:4_096].view(np.uint32))
    rows = [0, 1, 2, 3, 4, 13, 255, 256, 1_000, 4_095]
    ref = np.stack([O.wilders(c.astype(np.float64), r + 1) for r in rows])
    nan_mask_equal(seeded[rows], ref)
    assert max_rel_err(seeded[rows], ref, floor=0.0) <= 1e-5


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
    ref = O.ema_batch(x.astype(np.float64), ps.tolist())
    # ema_wrapper.rs:256 argument order; no alphas: 2/(p+1) in f64
    ctxs["ema"].ema_batch_device(d_x, d_p, None, x.size, 10, ps.size, d_out)
    ctx.synchronize()
    g = d_out.copy_to_host().reshape(ps.size, x.size)
    _check(g, ref, "ema")
    assert np.array_equal(g.view(np.uint32),
                          ctxs["ema"].ema_batch_dev(x, (5, 199, 7)).to_host().view(np.uint32))
    # the reference host code's f32 alphas 2/(p+1): within the bound of the f64 ones
    d_a = ctx.to_device((2.0 / (ps.astype(np.float32) + 1.0)).astype(np.float32))
    ctxs["ema"].ema_batch_device(d_x, d_p, d_a, x.size, 10, ps.size, d_out)
    ctx.synchronize()
    _check(d_out.copy_to_host().reshape(ps.size, x.size), ref, "ema")
    # many-series device form: the caller's alpha; clean flags found on the device
    h, lo, c = _ms_data(96, 10_000, 44)
    c[7_000, 5] = np.nan                                   # one unclean series
    rows, cols = c.shape
    d_c = ctx.to_device(c)
    fv = np.argmax(~np.isnan(c), axis=0).astype(np.int32)
    d_fv = ctx.to_device(fv)
    d_o = ctx.alloc(rows * cols * 4)
    ref = O.many_series("ema", c.astype(np.float64), 20)[0]
    ctxs["ema"].ema_many_series_one_param_device(d_c, d_fv, 20, np.float32(2.0 / 21.0), cols,
                                                 rows, d_o)
    ctx.synchronize()
    _check(d_o.copy_to_host().reshape(rows, cols), ref, "ema")
    assert ctxs["ema"].selected_many_series_kernel() == ("chunked", 64)


def test_device_form_clean_word_per_call(ctxs):
    """The device forms' clean flag is a scratch word that a call stamps with
    its own epoch when it finds a non-finite bar, never cleared: a series with
    an interior NaN (EMA skips it, which only the exact path does) between two
    clean calls on one context must run exactly, and the clean calls after it
    must not inherit its mark -- all three against the oracle."""
    x = datagen.gbm_close(30_000, 0.01, seed=47)
    x[:5] = np.nan
    bad = x.copy()
    bad[12_345] = np.nan
    ctx = ctxs["ema"].ctx
    ps = np.arange(2, 400, 9, dtype=np.int32)
    d_p = ctx.to_device(ps)
    d_o = ctx.alloc(ps.size * x.size * 4)
    for series in (x, bad, x, bad):
        d_x = ctx.to_device(series)
        ctxs["ema"].ema_batch_device(d_x, d_p, None, series.size, 5, ps.size, d_o)
        ctx.synchronize()
        g = d_o.copy_to_host().reshape(ps.size, series.size)
        _check(g, O.ema_batch(series.astype(np.float64), ps.tolist()), "ema")
        d_x.free()


def _ema_custom_alpha(x, p, alpha):
    """EMA with the reference kernels' structure and a given smoothing factor:
    running mean over the first p valid bars (outputs from first_valid), then
    prev = (1 - alpha) prev + alpha x, skipping non-finite bars."""
    out = np.full(x.size, np.nan)
    fv = int(np.argmax(~np.isnan(x)))
    prev, vc = 0.0, 0
    for t in range(fv, x.size):
        v = float(x[t])
        if t < fv + p:
            if np.isfinite(v):
                vc += 1
                prev = ((vc - 1) * prev + v) / vc
        elif np.isfinite(v):
            prev = (1.0 - alpha) * prev + alpha * v
        out[t] = prev
    return out


def test_device_forms_use_caller_alphas(ctxs):
    """ema_batch_device / ema_many_series_one_param_device / the Wilders device
    forms run with the caller's smoothing factors, not 2/(p+1) or 1/p."""
    x = datagen.gbm_close(3_000, 0.01, seed=45).astype(np.float64)
    x[:6] = np.nan
    x = x.astype(np.float32)
    ctx = ctxs["ema"].ctx
    ps = np.array([5, 9, 30], dtype=np.int32)
    al = np.array([0.05, 0.5, 0.9], dtype=np.float32)
    d_x, d_p, d_a = ctx.to_device(x), ctx.to_device(ps), ctx.to_device(al)
    d_o = ctx.alloc(ps.size * x.size * 4)
    ctxs["ema"].ema_batch_device(d_x, d_p, d_a, x.size, 6, ps.size, d_o)
    ctx.synchronize()
    g = d_o.copy_to_host().reshape(ps.size, x.size)
    ref = np.stack([_ema_custom_alpha(x.astype(np.float64), int(p), float(a))
                    for p, a in zip(ps, al)])
    nan_mask_equal(g, ref)
    assert max_rel_err(g, ref) <= 1e-5
    # many-series, one alpha for all series
    tm = np.stack([x, np.roll(x, 0) * 1.1], axis=1).astype(np.float32)
    d_tm = ctx.to_device(tm)
    d_fv = ctx.to_device(np.array([6, 6], dtype=np.int32))
    d_o2 = ctx.alloc(tm.size * 4)
    ctxs["ema"].ema_many_series_one_param_device(d_tm, d_fv, 9, np.float32(0.25), 2, x.size,
                                                 d_o2)
    ctx.synchronize()
    g2 = d_o2.copy_to_host().reshape(x.size, 2)
    for j in range(2):
        r = _ema_custom_alpha(tm[:, j].astype(np.float64), 9, 0.25)
        nan_mask_equal(g2[:, j], r)
        assert max_rel_err(g2[:, j], r) <= 1e-5


def test_policies_and_selected_kernels(ctxs):
    """CudaEmaPolicy / CudaRsiPolicy / CudaAdxPolicy: the plain form (one lane
    per combo over the whole series) gives the same values within the bound,
    and the selected kernel reports which form ran."""
    x = datagen.gbm_close(40_000, 0.01, seed=46)
    x[:3] = np.nan
    h, lo, c = _hlc(40_000, 46, lead=3)
    for name in ("ema", "rsi", "adx"):
        op = ctxs[name]
        assert op.selected_batch_kernel() is None or op.selected_batch_kernel()[1] == 64
        args = (h, lo, c) if name == "adx" else (x,)
        f = [a.astype(np.float64) for a in args]
        ps = op.expand_grid((5, 60, 11))
        ref = getattr(O, f"{name}_batch")(*f, ps)
        op.set_policy(vta.KernelPolicy(batch="plain", batch_block_x=64))
        assert op.policy() == vta.KernelPolicy("plain", 64, "auto", 0)
        g = getattr(op, f"{name}_batch_dev")(*args, (5, 60, 11)).to_host()
        assert op.selected_batch_kernel() == ("plain", 64)
        _check(g, ref, name)
        op.set_policy(vta.KernelPolicy())
        g = getattr(op, f"{name}_batch_dev")(*args, (5, 60, 11)).to_host()
        assert op.selected_batch_kernel() == ("chunked", 64)
        _check(g, ref, name)
        with pytest.raises(vta.VtaError) as e:
            op.set_policy(vta.KernelPolicy(batch="plain", batch_block_x=256))
        assert e.value.kind == "InvalidPolicy"
    # many-series plain form
    hm, lm, cm = _ms_data(70, 8_000, 47)
    cols, rows = cm.shape[1], cm.shape[0]
    op = vta.CudaRsi(0, policy=vta.KernelPolicy(many_series="plain"))
    g = op.rsi_many_series_one_param_time_major_dev(cm, cols, rows, 14).to_host()
    assert op.selected_many_series_kernel() == ("plain", 64)
    _check(g, O.many_series("rsi", cm.astype(np.float64), 14)[0], "rsi")


def test_host_copy_forms(ctxs):
    """ema_batch_into_host_f32 / ema_many_series_..._into_host_f32 and the ADX
    pair (ema_wrapper.rs:309, 422; adx_wrapper.rs:253, 417): the device
    results copied into caller-owned host slices."""
    x = datagen.gbm_close(20_000, 0.01, seed=48)
    x[:2] = np.nan
    h, lo, c = _hlc(20_000, 48, lead=2)
    ps = ctxs["ema"].expand_grid((4, 90, 9))
    out = np.full(len(ps) * x.size, -1.0, np.float32)
    ctxs["ema"].ema_batch_into_host_f32(x, (4, 90, 9), out)
    _check(out.reshape(len(ps), -1), O.ema_batch(x.astype(np.float64), ps), "ema")
    with pytest.raises(vta.VtaError):
        ctxs["ema"].ema_batch_into_host_f32(x, (4, 90, 9), out[:-1])
    aps = ctxs["adx"].expand_grid((5, 40, 7))
    out = np.empty(len(aps) * c.size, np.float32)
    rows, cols, got = ctxs["adx"].adx_batch_into_host_f32(h, lo, c, (5, 40, 7), out)
    assert (rows, cols, got) == (len(aps), c.size, aps)
    f = [a.astype(np.float64) for a in (h, lo, c)]
    _check(out.reshape(rows, cols), O.adx_batch(*f, aps), "adx")
    hm, lm, cm = _ms_data(33, 5_000, 49)
    R, N = cm.shape
    o = np.empty(R * N, np.float32)
    ctxs["ema"].ema_many_series_one_param_time_major_into_host_f32(cm, N, R, 20, o)
    _check(o.reshape(R, N), O.many_series("ema", cm.astype(np.float64), 20)[0], "ema")
    ctxs["adx"].adx_many_series_one_param_time_major_into_host_f32(hm, lm, cm, N, R, 14, o)
    fm = [a.astype(np.float64) for a in (hm, lm, cm)]
    _check(o.reshape(R, N), O.many_series_hlc("adx", *fm, 14)[0], "adx")
    with pytest.raises(vta.VtaError):
        ctxs["adx"].adx_many_series_one_param_time_major_into_host_f32(hm, lm, cm, N, R, 14,
                                                                       o[:-2])


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
