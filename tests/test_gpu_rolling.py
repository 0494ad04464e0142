"""Bollinger Bands / percentile_nearest_rank / rolling MAD parity on the GPU.

Bar: NaN warm-up prefixes and poison rules bit-exact everywhere.  PNR selects
an input value, and MAD's median/deviation arithmetic is the oracle's f64
sequence on the same f32 inputs, so both must equal f32(oracle) bit for bit.
BB re-anchors its f64 window sums at every chunk start: values within 1e-6
relative of the f64 oracle (bands of prices ~100; f32 output rounding is 6e-8),
+-Inf positions identical.
"""
import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import datagen
from tests.conftest import require_gpu
from tests.helpers import max_rel_err, nan_mask_equal

pytestmark = pytest.mark.gpu

BB_TOL = 1e-6


@pytest.fixture(scope="module")
def ops():
    require_gpu()
    return {"bb": vta.CudaBollingerBands(0), "pnr": vta.CudaPercentileNearestRank(0),
            "mad": vta.CudaMediumAd(0)}


def _exact(g, ref, what):
    r = np.asarray(ref, dtype=np.float64).astype(np.float32)
    g = np.asarray(g, dtype=np.float32)
    nan_mask_equal(g, r)
    bad = np.flatnonzero(~((g == r) | (np.isnan(g) & np.isnan(r))))
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first {bad[:5]}: " \
                          f"gpu={g.ravel()[bad[:5]]} ref={r.ravel()[bad[:5]]}"


def _bb_check(gpu3, ref3, what):
    for g, r, name in zip(gpu3, ref3, ("upper", "middle", "lower")):
        g = np.asarray(g, dtype=np.float64)
        nan_mask_equal(g, r)
        assert np.array_equal(np.isposinf(g), np.isposinf(r)), (what, name, "+inf")
        assert np.array_equal(np.isneginf(g), np.isneginf(r)), (what, name, "-inf")
        err = max_rel_err(g, r)
        print(f"bb {what} {name}: max rel err {err:.3e}")
        assert err <= BB_TOL, (what, name, err)


# ---------------------------------------------------------------- Bollinger Bands

@pytest.mark.parametrize("n,sd", [(3_000, 0.01), (300_000, 0.002)])
def test_bb_batch(ops, n, sd):
    x = datagen.gbm_close(n, sd, seed=51)
    x[:6] = np.nan
    sweep = vta.BollingerBandsBatchRange(period=(5, 60, 11), devup=(1.0, 2.5, 0.5),
                                         devdn=(2.0, 3.0, 1.0))
    combos = ops["bb"].expand_grid(sweep)
    u, m, lo = (d.to_host() for d in ops["bb"].bollinger_bands_batch_dev(x, sweep))
    assert u.shape == (len(combos), n)
    xd = x.astype(np.float64)
    refs = [O.bollinger_bands(xd, p, a, b) for p, a, b in combos]
    _bb_check((u, m, lo), [np.stack([r[i] for r in refs]) for i in range(3)], f"batch n={n}")


def test_bb_batch_drifting_series_matches_two_pass(ops):
    """A 300,000-bar walk that drifts from ~900 down to ~0.001.  The oracle's
    running sums (never re-anchored, as bollinger_bands.rs:640-641) carry the
    rounding of the large-price era into the small-price windows: 3% off the
    exact bands there.  The GPU re-anchors in f64 at every chunk start, so it
    is checked against an exact two-pass mean / population sd instead."""
    from numpy.lib.stride_tricks import sliding_window_view
    n, lead = 300_000, 6
    x = datagen.gbm_close(n, 0.01, seed=51)
    x[:lead] = np.nan
    sweep = vta.BollingerBandsBatchRange(period=(5, 60, 11), devup=(1.0, 2.5, 0.5),
                                         devdn=(2.0, 3.0, 1.0))
    combos = ops["bb"].expand_grid(sweep)
    u, m, lo = (d.to_host() for d in ops["bb"].bollinger_bands_batch_dev(x, sweep))
    xd = x.astype(np.float64)
    worst = 0.0
    for r, (p, a, b) in enumerate(combos):
        w = sliding_window_view(xd[lead:], p)
        mu, sd = w.mean(axis=1), w.std(axis=1)
        for g, e in ((u[r], mu + a * sd), (m[r], mu), (lo[r], mu - b * sd)):
            assert np.isnan(g[:lead + p - 1]).all()
            gv = g[lead + p - 1:].astype(np.float64)
            worst = max(worst, float(np.max(np.abs(gv - e) / np.maximum(np.abs(e), mu * 1e-3))))
    print(f"bb drifting series: max rel err vs two-pass {worst:.3e}")
    assert worst <= BB_TOL


def test_bb_batch_poison_rules(ops):
    """A NaN poisons every later output; an Inf makes the middle band +-Inf
    (bands NaN or Inf) until it leaves the window, then NaN for good."""
    n = 40_000
    sweep = vta.BollingerBandsBatchRange(period=(5, 45, 20), devup=(2.0, 2.0, 0.0),
                                         devdn=(2.0, 2.0, 0.0))
    combos = ops["bb"].expand_grid(sweep)
    for bad in (np.nan, np.inf, -np.inf):
        x = datagen.gbm_close(n, 0.01, seed=52)
        x[:3] = np.nan
        x[21_000] = bad
        out = [d.to_host() for d in ops["bb"].bollinger_bands_batch_dev(x, sweep)]
        xd = x.astype(np.float64)
        refs = [O.bollinger_bands(xd, p, a, b) for p, a, b in combos]
        _bb_check(out, [np.stack([r[i] for r in refs]) for i in range(3)], f"poison {bad}")
        assert np.isnan(out[1][:, 21_050:]).all()


def test_bb_many_series_jumps_shape(ops):
    """1,024 series x 50,000 bars with jumps, period 20, k = 2: NaN prefix of
    every series, 48 sampled series against the oracle."""
    x = datagen.many_series_with_jumps(1024, 50_000, seed=4)
    rows, cols = x.shape
    lead = (np.arange(cols) * 7) % 300
    for s in range(cols):
        x[:lead[s], s] = np.nan
    out = [d.to_host() for d in
           ops["bb"].bollinger_bands_many_series_one_param_time_major_dev(x, cols, rows, 20,
                                                                           2.0, 2.0)]
    first = np.argmax(~np.isnan(out[1]), axis=0)
    assert np.array_equal(first, lead + 19)
    sel = np.linspace(0, cols - 1, 48).astype(int)
    xd = x[:, sel].astype(np.float64)
    refs = [O.bollinger_bands(xd[:, i], 20, 2.0, 2.0) for i in range(sel.size)]
    _bb_check([o[:, sel] for o in out], [np.stack([r[i] for r in refs], axis=1)
                                        for i in range(3)], "many-series")


@pytest.mark.parametrize("chunk", [None, "64"])
def test_bb_many_series_periods_ring(ops, monkeypatch, chunk):
    """Many-series BB over periods either side of the 8-bar load block (the
    outgoing bars of a block then come from the block itself) and of the
    256-bar LDS ring (longer windows read the outgoing bars from HBM), on a
    ragged panel with NaN runs and +-Inf bars (poison rules), chunked and as
    one chunk per series: every column against the oracle."""
    if chunk:
        monkeypatch.setenv("VTA_ROLL_CHUNK", chunk)
    rows, cols = 3_000, 70
    x = _ragged_panel(rows, cols, seed=94, short=400)
    xd = x.astype(np.float64)
    for p in (2, 3, 7, 8, 9, 20, 255, 256, 257):
        refs = [O.bollinger_bands(xd[:, s], p, 2.0, 1.5) for s in range(cols)]
        ok = np.array([not np.isnan(r[0]).all() for r in refs])
        ref3 = [np.stack([r[i] for r in refs], axis=1) for i in range(3)]
        out = [d.to_host() for d in
               ops["bb"].bollinger_bands_many_series_one_param_time_major_dev(x, cols, rows, p,
                                                                               2.0, 1.5)]
        _bb_check([o[:, ok] for o in out], [r[:, ok] for r in ref3], f"ring period {p}")


def test_bb_many_series_unclean_and_device_form(ops):
    rows, cols = 30_000, 130
    x = datagen.many_series_with_jumps(cols, rows, seed=53)
    x[:50, 1] = np.nan
    x[10_000, 3] = np.nan
    x[20_000, 4] = np.inf
    x[20_003, 4] = -np.inf
    x[25_000, 5] = np.inf
    xd = x.astype(np.float64)
    refs = [O.bollinger_bands(xd[:, s], 30, 1.5, 2.5) for s in range(cols)]
    ref3 = [np.stack([r[i] for r in refs], axis=1) for i in range(3)]
    out = [d.to_host() for d in
           ops["bb"].bollinger_bands_many_series_one_param_time_major_dev(x, cols, rows, 30,
                                                                           1.5, 2.5)]
    _bb_check(out, ref3, "ms unclean")
    # device form, with host first_valid / first non-finite bar, and without
    ctx = ops["bb"].ctx
    fv = np.argmax(~np.isnan(x), axis=0).astype(np.int32)
    z = np.full(cols, rows, dtype=np.int32)
    for s in range(cols):
        nf = np.flatnonzero(~np.isfinite(x[fv[s]:, s]))
        if nf.size:
            z[s] = fv[s] + nf[0]
    d_x, d_fv, d_z = ctx.to_device(x), ctx.to_device(fv), ctx.to_device(z)
    d_o = [ctx.alloc(rows * cols * 4) for _ in range(3)]
    for d_poison in (d_z, None):
        ops["bb"].bollinger_bands_many_series_one_param_device(d_x, d_fv, d_poison, cols, rows,
                                                              30, 1.5, 2.5, *d_o)
        ctx.synchronize()
        _bb_check([b.copy_to_host().reshape(rows, cols) for b in d_o], ref3,
                  f"device poison={'table' if d_poison is not None else 'none'}")


# ---------------------------------------------------------------- percentile nearest rank

def test_pnr_tiny_vectors(ops):
    # tests/python/test_utils.py:165-179 shapes: 1..10 ramp, window 5
    x = np.arange(1.0, 11.0, dtype=np.float32)
    for pct, want in ((50.0, 3.0), (25.0, 1.0), (75.0, 4.0), (100.0, 5.0)):
        g, _ = ops["pnr"].pnr_batch_dev(x, vta.PercentileNearestRankBatchRange((5, 5, 0),
                                                                                (pct, pct, 0.0)))
        g = g.to_host()[0]
        assert np.isnan(g[:4]).all() and g[4] == want, (pct, g)
        _exact(g, O.pnr(x.astype(np.float64), 5, pct), f"tiny {pct}")


@pytest.mark.parametrize("n", [5_000, 60_000])
def test_pnr_batch(ops, n):
    x = datagen.gbm_close(n, 0.01, seed=54)
    x[:9] = np.nan
    x[1_000:1_003] = np.nan                    # interior NaNs are skipped
    x[2_000:2_200] = np.nan                    # a gap longer than every window
    x[3_000] = np.inf
    sweep = vta.PercentileNearestRankBatchRange(length=(5, 120, 23), percentage=(10.0, 90.0, 20.0))
    d, combos = ops["pnr"].pnr_batch_dev(x, sweep)
    g = d.to_host()
    assert g.shape == (len(combos), n)
    xd = x.astype(np.float64)
    ref = np.stack([O.pnr(xd, l, p) for l, p in combos])
    _exact(g, ref, f"pnr batch n={n}")


def test_pnr_many_series_jumps_shape(ops):
    """window 100, percentile 70 over 1,024 x 50,000 with jumps; every value of
    48 sampled series bit-exact; NaN prefixes of all series."""
    x = datagen.many_series_with_jumps(1024, 50_000, seed=4)
    rows, cols = x.shape
    lead = (np.arange(cols) * 5) % 200
    for s in range(cols):
        x[:lead[s], s] = np.nan
    x[7_000:7_010, 11] = np.nan
    g = ops["pnr"].pnr_many_series_one_param_time_major_dev(x, cols, rows, 100, 70.0).to_host()
    assert np.array_equal(np.argmax(~np.isnan(g), axis=0), lead + 99)
    sel = np.unique(np.concatenate([np.linspace(0, cols - 1, 47).astype(int), [11]]))
    ref, st = O.many_series("pnr", x[:, sel].astype(np.float64), 100, 70.0)
    assert (st == 0).all()
    _exact(g[:, sel], ref, "pnr many-series")


# ---------------------------------------------------------------- rolling MAD

@pytest.mark.parametrize("n", [4_000, 40_000])
def test_mad_batch(ops, n):
    x = datagen.gbm_close(n, 0.01, seed=55)
    x[:4] = np.nan
    x[1_500] = np.nan                          # any NaN in the window -> NaN
    periods = ops["mad"].expand_grid((1, 70, 6))
    g = ops["mad"].medium_ad_batch_dev(x, (1, 70, 6)).to_host()
    xd = x.astype(np.float64)
    ref = np.stack([O.medium_ad(xd, p) for p in periods])
    _exact(g, ref, f"mad batch n={n}")
    g2 = ops["mad"].medium_ad_batch_dev(x, (64, 61, 1)).to_host()   # even and odd
    ref2 = np.stack([O.medium_ad(xd, p) for p in (64, 63, 62, 61)])
    _exact(g2, ref2, "mad even/odd")


def test_mad_ties(ops):
    """Many equal values: ties at the median and zero deviations."""
    rng = np.random.Generator(np.random.PCG64(56))
    x = (100.0 + rng.integers(-3, 4, size=20_000) * 0.25).astype(np.float32)
    g = ops["mad"].medium_ad_batch_dev(x, (2, 66, 8)).to_host()
    xd = x.astype(np.float64)
    ref = np.stack([O.medium_ad(xd, p) for p in ops["mad"].expand_grid((2, 66, 8))])
    _exact(g, ref, "mad ties")
    g, _ = ops["pnr"].pnr_batch_dev(x, vta.PercentileNearestRankBatchRange((3, 99, 16),
                                                                            (0.0, 100.0, 12.5)))
    combos = ops["pnr"].expand_grid(vta.PercentileNearestRankBatchRange((3, 99, 16),
                                                                         (0.0, 100.0, 12.5)))
    _exact(g.to_host(), np.stack([O.pnr(xd, l, p) for l, p in combos]), "pnr ties")


def test_mad_many_series_jumps_shape(ops):
    x = datagen.many_series_with_jumps(1024, 50_000, seed=4)
    rows, cols = x.shape
    x[:33, 2] = np.nan
    x[9_000, 17] = np.nan
    g = ops["mad"].medium_ad_many_series_one_param_time_major_dev(x, cols, rows, 64).to_host()
    lead = np.zeros(cols, dtype=int)
    lead[2] = 33
    assert np.array_equal(np.argmax(~np.isnan(g), axis=0), lead + 63)
    sel = np.unique(np.concatenate([np.linspace(0, cols - 1, 30).astype(int), [2, 17]]))
    ref, st = O.many_series("medium_ad", x[:, sel].astype(np.float64), 64)
    assert (st == 0).all()
    _exact(g[:, sel], ref, "mad many-series")


def test_errors(ops):
    x = datagen.gbm_close(100, 0.01, seed=57)
    with pytest.raises(vta.VtaError):
        ops["pnr"].pnr_batch_dev(x, vta.PercentileNearestRankBatchRange((5, 5, 0),
                                                                        (101.0, 101.0, 0.0)))
    with pytest.raises(vta.VtaError):
        ops["mad"].medium_ad_batch_dev(x, (200, 200, 0))
    with pytest.raises(vta.VtaError):
        ops["bb"].bollinger_bands_batch_dev(np.full(50, np.nan, np.float32),
                                            vta.BollingerBandsBatchRange())
    with pytest.raises(vta.VtaError):
        ops["bb"].bollinger_bands_many_series_one_param_time_major_dev(
            np.ones((10, 3), np.float32), 3, 10, 11, 2.0, 2.0)


def test_rolling_plan_enqueue_only(ops):
    """RollingManySeriesPlan: device-resident data, three enqueue-only launches
    back to back on one stream, each checked on sampled series."""
    rows, cols = 20_000, 256
    x = datagen.many_series_with_jumps(cols, rows, seed=58)
    x[:40, 7] = np.nan
    x[12_345, 9] = np.nan
    ctx = ops["bb"].ctx
    plan = vta.RollingManySeriesPlan(ctx, x)
    sel = [0, 7, 9, 100, 255]
    xd = x[:, sel].astype(np.float64)
    plan.launch_bb(20, 2.0, 2.0)
    ctx.synchronize()
    refs = [O.bollinger_bands(xd[:, i], 20, 2.0, 2.0) for i in range(len(sel))]
    _bb_check([o[:, sel] for o in plan.outputs(3)],
              [np.stack([r[i] for r in refs], axis=1) for i in range(3)], "plan")
    plan.launch_pnr(100, 70.0)
    ctx.synchronize()
    _exact(plan.outputs()[0][:, sel], O.many_series("pnr", xd, 100, 70.0)[0], "plan pnr")
    plan.launch_mad(64)
    ctx.synchronize()
    _exact(plan.outputs()[0][:, sel], O.many_series("medium_ad", xd, 64)[0], "plan mad")


# ---------------------------------------------------------------- window forms

def _ragged_panel(rows, cols, seed, quant=None, short=200):
    """Time-major panel: GBM columns with ragged NaN leads, interior NaN runs,
    a few +-Inf bars and (quant) values rounded to a grid so windows hold ties."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.stack([datagen.gbm_close(rows, 0.01, seed=int(s)) for s in rng.integers(1 << 30, size=cols)],
                 axis=1).astype(np.float32)
    if quant:
        x = (np.round(x / quant) * quant).astype(np.float32)
    lead = rng.integers(0, rows // 4, size=cols)
    lead[:3] = (0, 1, rows - short)             # one column with few valid bars
    for s in range(cols):
        x[:lead[s], s] = np.nan
    for s in rng.integers(0, cols, size=cols // 4):
        t = int(rng.integers(rows // 4, rows - 10))
        x[t:t + int(rng.integers(1, 6)), s] = np.nan
    x[rows // 2, 5] = np.inf
    x[rows // 3, 6] = -np.inf
    x[rows // 2 + 7, 8] = 1e-40                 # denormal keys
    x[rows // 2 + 9, 8] = 3e-41
    return x


def _same_bits(a, b, what):
    """Bit-identical values with identical NaN positions (NaN payloads aside:
    a NaN computed from infinities may carry another sign or payload than the
    canonical quiet NaN)."""
    nan_mask_equal(a, b)
    m = ~np.isnan(a)
    assert np.array_equal(a.view(np.uint32)[m], b.view(np.uint32)[m]), what


def _with_form(ctx, form, fn):
    from vectorta_amd import autotune as at
    at.set_tuning(ctx, at.ROLL_FORM, form)
    try:
        return fn()
    finally:
        at.set_tuning(ctx, at.ROLL_FORM, 0)


@pytest.mark.parametrize("chunk", [None, "160"])
def test_pnr_window_forms_many_series(ops, monkeypatch, chunk):
    """Many-series PNR in the register form (every capacity 16..128, lengths at
    and either side of each 16-slot boundary, percentages 0..100) and in the
    LDS form: ragged leads, NaN runs, +-Inf, denormals and tied values, every
    column bit-exact against the oracle, and the two forms bit-identical.
    Lengths above 128 run the LDS form under either setting."""
    if chunk:
        monkeypatch.setenv("VTA_ROLL_CHUNK", chunk)
    rows, cols = 2_500, 70
    x = _ragged_panel(rows, cols, seed=91, quant=0.05)
    xd = x.astype(np.float64)
    pnr = ops["pnr"]
    cases = [(1, 50.0), (2, 0.0), (15, 70.0), (16, 100.0), (17, 33.3), (31, 50.0), (47, 99.9),
             (64, 12.5), (65, 70.0), (100, 70.0), (112, 1.0), (127, 50.0), (128, 87.5),
             (129, 70.0)]
    for length, pct in cases:
        ref, st = O.many_series("pnr", xd, length, pct)
        outs = {}
        for form in (1, 2):
            g = _with_form(pnr.ctx, form, lambda: pnr.pnr_many_series_one_param_time_major_dev(
                x, cols, rows, length, pct).to_host())
            ok = st == 0
            _exact(g[:, ok], ref[:, ok], f"pnr form {form} len {length} pct {pct}")
            assert np.isnan(g[:, ~ok]).all()
            outs[form] = g
        _same_bits(outs[1], outs[2], f"pnr forms len {length}")


@pytest.mark.parametrize("chunk", [None, "160"])
def test_mad_window_forms_many_series(ops, monkeypatch, chunk):
    """Many-series MAD in the register form (capacities 16..64, even and odd
    periods either side of each boundary, period 1) and in the LDS form: the
    same ragged panel, every column bit-exact against the oracle, the two
    forms bit-identical; periods above 64 run the LDS form."""
    if chunk:
        monkeypatch.setenv("VTA_ROLL_CHUNK", chunk)
    rows, cols = 2_000, 70
    x = _ragged_panel(rows, cols, seed=92, quant=0.05)
    xd = x.astype(np.float64)
    mad = ops["mad"]
    for p in (1, 2, 3, 15, 16, 17, 31, 32, 33, 48, 49, 63, 64, 65):
        ref, st = O.many_series("medium_ad", xd, p)
        outs = {}
        for form in (1, 2):
            g = _with_form(mad.ctx, form, lambda: mad.medium_ad_many_series_one_param_time_major_dev(
                x, cols, rows, p).to_host())
            ok = st == 0
            _exact(g[:, ok], ref[:, ok], f"mad form {form} period {p}")
            assert np.isnan(g[:, ~ok]).all()
            outs[form] = g
        _same_bits(outs[1], outs[2], f"mad forms period {p}")


def test_window_forms_signed_zeros(ops, monkeypatch):
    """Windows holding -0.0 and +0.0 together.  The oracle erases the first key
    equal to the outgoing bar, so which zero sign its window keeps depends on
    the whole history.  One chunk per series replays that history: PNR bit-exact
    in both forms (the register form takes its compare path whenever a -0.0
    key is present).  With chunks, every PNR value still equals the oracle's
    numerically (only a zero's sign may differ).  MAD's value does not depend
    on zero signs: bit-exact either way."""
    rows, cols = 1_500, 64
    rng = np.random.Generator(np.random.PCG64(93))
    x = np.round(rng.normal(0.0, 1.0, size=(rows, cols))).astype(np.float32)
    x[rng.random((rows, cols)) < 0.2] = -0.0
    x[:3, :] = np.nan
    xd = x.astype(np.float64)
    pnr, mad = ops["pnr"], ops["mad"]
    for chunk in (str(rows), "200"):
        monkeypatch.setenv("VTA_ROLL_CHUNK", chunk)
        for length, pct in ((9, 50.0), (40, 50.0), (100, 30.0)):
            ref, _ = O.many_series("pnr", xd, length, pct)
            for form in (1, 2):
                g = _with_form(pnr.ctx, form, lambda: pnr.pnr_many_series_one_param_time_major_dev(
                    x, cols, rows, length, pct).to_host())
                r32 = ref.astype(np.float32)
                if chunk == str(rows):
                    _exact(g, ref, f"pnr zeros form {form} len {length}")
                else:
                    nan_mask_equal(g, r32)
                    m = ~np.isnan(r32)
                    assert np.array_equal(g[m], r32[m]), (form, length)
        for p in (8, 33, 64):
            ref, _ = O.many_series("medium_ad", xd, p)
            for form in (1, 2):
                g = _with_form(mad.ctx, form, lambda: mad.medium_ad_many_series_one_param_time_major_dev(
                    x, cols, rows, p).to_host())
                _exact(g, ref, f"mad zeros form {form} period {p} chunk {chunk}")
