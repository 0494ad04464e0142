"""Pins the CPU oracle before it is trusted as the checker.

Three kinds of pin, chosen so a plausible restatement mistake (dropped term,
wrong sign/index, off-by-one warm-up) fails at least one:
  * the reference's own known-answer values on the Bitfinex 4h fixture
    (file:line of each test cited below), at the reference's tolerance;
  * closed forms / brute force computed a different way (numpy windows,
    explicit sums), on small inputs;
  * invariants stated by the reference's property tests (alma.rs:2251-2326).
"""
import numpy as np
import pytest

import oracle as O


def last5(a):
    return np.asarray(a)[-5:]


# ------------------------------------------------------------ known answers

def test_alma_known_answer(candles):
    # alma.rs:1835-1856 (tol 1e-8)
    exp = [59286.72216704, 59273.53428138, 59204.37290721, 59155.93381742, 59026.92526112]
    out = O.alma(candles["close"], 9, 0.85, 6.0)
    assert np.all(np.abs(last5(out) - exp) < 1e-8)


def test_alma_batch_default_row(candles):
    # alma.rs:2410-2440: default grid periods 9..=258, row (9, 0.85, 6) (tol 1e-8)
    exp = [59286.72216704, 59273.53428138, 59204.37290721, 59155.93381742, 59026.92526112]
    out = O.alma_batch(candles["close"], [(9, 0.85, 6.0), (10, 0.85, 6.0)])
    assert np.all(np.abs(out[0, -5:] - exp) < 1e-8)


def test_alma_reinput(candles):
    # tests/python/test_utils.py:120-126
    exp = [59140.73195170, 59211.58090986, 59238.16030697, 59222.63528822, 59165.14427332]
    a = O.alma(candles["close"])
    assert np.all(np.abs(last5(O.alma(a)) - exp) < 1e-8)


def test_ema_known_answer(candles):
    # ema.rs:971-983 (tol 1e-1)
    exp = [59302.2, 59277.9, 59230.2, 59215.1, 59103.1]
    assert np.all(np.abs(last5(O.ema(candles["close"], 9)) - exp) < 1e-1)


def test_rsi_known_answer(candles):
    # rsi.rs:1142-1156 (tol 1e-2)
    exp = [43.42, 42.68, 41.62, 42.86, 39.01]
    assert np.all(np.abs(last5(O.rsi(candles["close"], 14)) - exp) < 1e-2)


def test_atr_known_answer(candles):
    # atr.rs:1422-1436 (tol 1e-2); the batch-prefix row agrees to the same tol
    exp = [916.89, 874.33, 838.45, 801.92, 811.57]
    h, l, c = candles["high"], candles["low"], candles["close"]
    assert np.all(np.abs(last5(O.atr(h, l, c, 14)) - exp) < 1e-2)
    assert np.all(np.abs(last5(O.atr_batch_row(h, l, c, 14)) - exp) < 1e-2)


def test_adx_known_answer(candles):
    # adx.rs:2163-2175 (tol 1e-1)
    exp = [36.14, 36.52, 37.01, 37.46, 38.47]
    out = O.adx(candles["high"], candles["low"], candles["close"], 14)
    assert np.all(np.abs(last5(out) - exp) < 1e-1)


def test_bollinger_known_answer(candles):
    # bollinger_bands.rs:1705-1745 (tol 1e-4)
    em = [59403.199999999975, 59423.24999999998, 59370.49999999998, 59371.39999999998,
          59351.299999999974]
    el = [58299.51497247008, 58351.47038179873, 58332.65135978715, 58334.33194052157,
          58275.767369163135]
    eu = [60506.88502752987, 60495.029618201224, 60408.348640212804, 60408.468059478386,
          60426.83263083681]
    u, m, lo = O.bollinger_bands(candles["close"], 20, 2.0, 2.0)
    assert np.all(np.abs(last5(m) - em) < 1e-4)
    assert np.all(np.abs(last5(lo) - el) < 1e-4)
    assert np.all(np.abs(last5(u) - eu) < 1e-4)


def test_pnr_known_answer(candles):
    # percentile_nearest_rank.rs:1395-1425 (tol 1e-6), NaN warm-up 14
    out = O.pnr(candles["close"], 15, 50.0)
    assert np.isnan(out[:14]).all() and not np.isnan(out[14])
    assert np.all(np.abs(last5(out) - [59419.0, 59419.0, 59300.0, 59285.0, 59273.0]) < 1e-6)


def test_pnr_tiny_vectors():
    # tests/python/test_utils.py:165-179
    x = np.arange(1.0, 11.0)
    assert O.pnr(x, 5, 50.0)[4] == 3.0 and O.pnr(x, 5, 50.0)[5] == 4.0
    assert O.pnr(x, 5, 25.0)[4] == 1.0
    assert O.pnr(x, 5, 75.0)[4] == 4.0
    assert O.pnr(x, 5, 100.0)[4] == 5.0


def test_medium_ad_known_answer(candles):
    # medium_ad.rs:1478-1497 on hl2 = (high + low) / 2 (tol 1e-1)
    hl2 = (candles["high"] + candles["low"]) / 2.0
    exp = [220.0, 78.5, 126.5, 48.0, 28.5]
    assert np.all(np.abs(last5(O.medium_ad(hl2, 5)) - exp) < 1e-1)


# ------------------------------------------------------------ NaN prefixes

@pytest.mark.parametrize("lead", [0, 3, 17])
def test_warmup_prefix_lengths(lead):
    rng = np.random.default_rng(5)
    n = 400
    c = 100 + np.cumsum(rng.normal(0, 1, n))
    c[:lead] = np.nan
    h = c + 1.0
    lo = c - 1.0
    p = 14
    checks = {
        "alma": (O.alma(c, p, 0.85, 6.0), lead + p - 1),
        "ema": (O.ema(c, p), lead),
        "rsi": (O.rsi(c, p), lead + p),
        "atr": (O.atr(h, lo, c, p), lead + p - 1),
        "bb": (O.bollinger_bands(c, p, 2.0, 2.0)[1], lead + p - 1),
        "pnr": (O.pnr(c, p, 70.0), lead + p - 1),
        "mad": (O.medium_ad(c, p), lead + p - 1),
        "adx": (O.adx(h, lo, c, p), lead + 2 * p - 1),
    }
    for name, (out, warm) in checks.items():
        assert np.isnan(out[:warm]).all(), name
        assert not np.isnan(out[warm:]).any(), name


# ------------------------------------------------------------ closed forms, brute force

def _windows(x, w):
    return np.lib.stride_tricks.sliding_window_view(x, w)


def test_alma_matches_explicit_weighted_mean():
    rng = np.random.default_rng(7)
    x = 50 + np.cumsum(rng.normal(0, 0.5, 300))
    for p, off, sig in [(9, 0.85, 6.0), (20, 0.05, 2.0), (33, 0.5, 10.0), (1, 0.85, 6.0)]:
        i = np.arange(p, dtype=np.float64)
        m = off * (p - 1)
        s = p / sig
        w = np.exp(-((i - m) ** 2) / (2 * s * s))
        ref = _windows(x, p) @ (w / w.sum())
        out = O.alma(x, p, off, sig)
        assert np.allclose(out[p - 1:], ref, rtol=1e-13, atol=0)
        outb = O.alma_batch(x, [(p, off, sig)])[0]
        assert np.allclose(outb[p - 1:], ref, rtol=1e-13, atol=0)


def test_alma_invariants():
    rng = np.random.default_rng(8)
    x = 10 + rng.random(200)
    out = O.alma(x, 12, 0.85, 6.0)
    win = _windows(x, 12)
    assert np.all(out[11:] <= win.max(axis=1) + 1e-12)
    assert np.all(out[11:] >= win.min(axis=1) - 1e-12)
    assert np.array_equal(O.alma(x, 1, 0.85, 6.0), x)          # period 1 is the identity
    const = np.full(100, 3.25)
    assert np.allclose(O.alma(const, 10)[9:], 3.25, rtol=1e-15)
    # sigma -> 0+ (Gaussian width p/sigma -> infinity) is the simple moving average
    sma = _windows(x, 12).mean(axis=1)
    assert np.allclose(O.alma(x, 12, 0.85, 1e-12)[11:], sma, rtol=1e-12)
    w, inv = O.alma_weights_batch(12, 0.85, 6.0)
    assert abs(w.sum() * inv - 1.0) < 1e-15


def test_ema_closed_form():
    rng = np.random.default_rng(9)
    x = 100 + np.cumsum(rng.normal(0, 1, 200))
    p = 10
    a = 2.0 / (p + 1)
    out = O.ema(x, p)
    # warm-up: running mean of x[0..i]
    assert np.allclose(out[:p], np.cumsum(x[:p]) / np.arange(1, p + 1), rtol=1e-14)
    # afterwards the non-recursive sum beta^k weights
    seed = out[p - 1]
    for t in [p, p + 5, 150, 199]:
        k = t - (p - 1)
        js = np.arange(p, t + 1)
        val = (1 - a) ** k * seed + np.sum(a * (1 - a) ** (t - js) * x[js])
        assert abs(out[t] - val) < 1e-10 * abs(val)
    assert np.allclose(O.ema(np.full(50, 7.5), 9), 7.5, rtol=1e-15)
    # non-finite values are skipped (ema.rs:441-467)
    y = x.copy()
    y[50] = np.nan
    oy = O.ema(y, p)
    assert oy[50] == oy[49] and np.isfinite(oy[51:]).all()


def test_rsi_closed_forms():
    rng = np.random.default_rng(10)
    x = 100 + np.cumsum(rng.normal(0, 1, 300))
    p = 14
    out = O.rsi(x, p)
    d = np.diff(x)
    g0 = np.clip(d[:p], 0, None).mean()
    l0 = np.clip(-d[:p], 0, None).mean()
    assert abs(out[p] - 100 * g0 / (g0 + l0)) < 1e-12
    assert np.all((out[p:] >= 0) & (out[p:] <= 100))
    assert np.all(O.rsi(np.full(60, 5.0), p)[p:] == 50.0)
    assert np.all(O.rsi(np.arange(60.0), p)[p:] == 100.0)
    assert np.all(O.rsi(-np.arange(60.0), p)[p:] == 0.0)
    # a NaN inside the seed poisons everything after (rsi.rs:340-360)
    y = x.copy()
    y[5] = np.nan
    assert np.isnan(O.rsi(y, p)[p:]).all()


def test_atr_brute_force():
    rng = np.random.default_rng(11)
    c = 100 + np.cumsum(rng.normal(0, 1, 200))
    h = c + rng.random(200)
    lo = c - rng.random(200)
    p = 14
    tr = np.empty(200)
    tr[0] = h[0] - lo[0]
    pc = c[:-1]
    tr[1:] = np.maximum.reduce([h[1:] - lo[1:], np.abs(h[1:] - pc), np.abs(lo[1:] - pc)])
    out = O.atr(h, lo, c, p)
    assert abs(out[p - 1] - tr[:p].mean()) < 1e-12
    a = 1.0 / p
    for t in [p, 60, 199]:
        k = t - (p - 1)
        js = np.arange(p, t + 1)
        val = (1 - a) ** k * tr[:p].mean() + np.sum(a * (1 - a) ** (t - js) * tr[js])
        assert abs(out[t] - val) < 1e-10 * val
    assert np.all(O.atr(np.full(40, 2.0), np.full(40, 2.0), np.full(40, 2.0), 5)[4:] == 0.0)


def test_bollinger_brute_force():
    rng = np.random.default_rng(12)
    x = 100 + np.cumsum(rng.normal(0, 1, 500))
    p, k = 20, 2.0
    u, m, lo = O.bollinger_bands(x, p, k, k)
    win = _windows(x, p)
    mean = win.mean(axis=1)
    sd = win.std(axis=1)  # population
    assert np.allclose(m[p - 1:], mean, rtol=1e-12)
    assert np.allclose(u[p - 1:], mean + k * sd, rtol=1e-9)
    assert np.allclose(lo[p - 1:], mean - k * sd, rtol=1e-9)
    y = x.copy()
    y[100] = np.nan
    u2, m2, l2 = O.bollinger_bands(y, p, k, k)
    assert np.isnan(m2[100:]).all() and np.isfinite(m2[p - 1:100]).all()


def test_pnr_brute_force():
    rng = np.random.default_rng(13)
    x = np.round(rng.normal(0, 3, 400), 1)   # many ties
    x[[30, 31, 200]] = np.nan
    for length, pct in [(10, 70.0), (7, 50.0), (25, 0.0), (25, 100.0)]:
        out = O.pnr(x, length, pct)
        for i in range(length - 1, x.size):
            w = x[i - length + 1:i + 1]
            w = np.sort(w[~np.isnan(w)])
            if w.size == 0:
                assert np.isnan(out[i])
                continue
            kk = int(np.floor(pct / 100.0 * w.size + 0.5)) - 1
            kk = min(max(kk, 0), w.size - 1)
            assert out[i] == w[kk]


def test_medium_ad_brute_force():
    rng = np.random.default_rng(14)
    x = 100 + np.cumsum(rng.normal(0, 1, 300))
    x[150] = np.nan
    for p in [1, 2, 5, 64]:
        out = O.medium_ad(x, p)
        for i in range(p - 1, x.size):
            w = x[i - p + 1:i + 1]
            if np.isnan(w).any():
                assert np.isnan(out[i])
                continue
            med = np.median(w)
            assert out[i] == np.median(np.abs(w - med))


def test_adx_pure_trend_is_100():
    n, p = 120, 14
    c = np.arange(n, dtype=np.float64) + 100.0
    h = c + 0.5
    lo = c - 0.5
    out = O.adx(h, lo, c, p)
    assert np.allclose(out[2 * p - 1:], 100.0, rtol=1e-12)
    assert np.all((out[2 * p - 1:] >= 0) & (out[2 * p - 1:] <= 100.0 + 1e-9))


def test_oracle_error_codes():
    with pytest.raises(O.OracleError):
        O.alma(np.array([]), 9)
    with pytest.raises(O.OracleError):
        O.alma(np.full(10, np.nan), 3)
    with pytest.raises(O.OracleError):
        O.alma(np.arange(5.0), 9)
    with pytest.raises(O.OracleError):
        O.alma(np.arange(50.0), 9, 1.5, 6.0)
    with pytest.raises(O.OracleError):
        O.pnr(np.arange(50.0), 5, 101.0)


def test_cross_pairs_vectorised_and_antisymmetric():
    """or_cross_pairs against an independent vectorised numpy formulation of the
    same statistics, plus the identities fast<->slow swap (long<->short,
    checksum and pnl negated), n_long + n_short <= bars, and a pair of identical
    rows giving all zeros."""
    from vectorta_amd import datagen
    x = datagen.gbm_close(4_000, 0.015, seed=61).astype(np.float64)
    combos = [(p, 0.85, 6.0) for p in (5, 9, 20, 33, 50)] + [(20, 0.5, 3.0)]
    rows = O.alma_batch(x, combos)
    n = len(combos)
    pf = [i for i in range(n) for j in range(n)]
    ps = [j for i in range(n) for j in range(n)]
    st, pnl = O.cross_pairs(rows, x, pf, ps)
    for k, (i, j) in enumerate(zip(pf, ps)):
        d = rows[i] - rows[j]
        ok = ~(np.isnan(rows[i]) | np.isnan(rows[j]))
        w = int(np.argmax(ok))
        pos = np.sign(d[w:]).astype(np.int64)
        t = np.arange(w, x.size)
        assert st[k, 0] == np.count_nonzero(np.diff(pos))
        assert st[k, 1] == np.count_nonzero(pos > 0)
        assert st[k, 2] == np.count_nonzero(pos < 0)
        assert st[k, 3] == int(np.sum(pos * (t + 1)))
        assert pnl[k] == pytest.approx(float(np.sum(pos[:-1] * np.diff(x[w:]))), rel=1e-12,
                                       abs=1e-9)
        assert st[k, 1] + st[k, 2] <= x.size - w
        kk = j * n + i
        assert st[kk, 1] == st[k, 2] and st[kk, 2] == st[k, 1] and st[kk, 0] == st[k, 0]
        assert st[kk, 3] == -st[k, 3] and pnl[kk] == -pnl[k]
        if i == j:
            assert (st[k] == 0).all() and pnl[k] == 0.0


# ------------------------------------------------------------ SMA / WMA / SWMA / SMMA / Wilders

def test_sma_known_answer(candles):
    # sma.rs:1625-1641 (SMA(9) of close, tol 1e-1)
    exp = [59180.8, 59175.0, 59129.4, 59085.4, 59133.7]
    for f in (O.sma, O.sma_batch_row):
        assert np.all(np.abs(last5(f(candles["close"], 9)) - exp) < 1e-1)


def test_wma_known_answer(candles):
    # wma.rs:855-883 (WMA(30) of close, tol 1e-6)
    exp = [59638.52903225806, 59563.7376344086, 59489.4064516129, 59432.02580645162,
           59350.58279569892]
    for f in (O.wma, O.wma_batch_row):
        assert np.all(np.abs(last5(f(candles["close"], 30)) - exp) < 1e-6)


def test_swma_known_answer(candles):
    # swma.rs:1299-1318 (SWMA(5) of close, tol 1e-8)
    exp = [59288.22222222222, 59301.99999999999, 59247.33333333333, 59179.88888888889,
           59080.99999999999]
    assert np.all(np.abs(last5(O.swma(candles["close"], 5)) - exp) < 1e-8)


def test_smma_known_answer(candles):
    # smma.rs:915-931 (SMMA(7) of close, tol 1e-1)
    exp = [59434.4, 59398.2, 59346.9, 59319.4, 59224.5]
    assert np.all(np.abs(last5(O.smma(candles["close"], 7)) - exp) < 1e-1)


def test_wilders_known_answer(candles):
    # wilders.rs:1226-1247 (Wilders(5) of close, tol 1e-8)
    exp = [59302.18156619092, 59277.94525295273, 59230.15620236219, 59215.12496188975,
           59103.0999695118]
    assert np.all(np.abs(last5(O.wilders(candles["close"], 5)) - exp) < 1e-8)


def _windows(x, p):
    return np.lib.stride_tricks.sliding_window_view(x, p)


@pytest.mark.parametrize("p", [1, 2, 3, 8, 31])
def test_fir_averages_brute_force(p):
    """Windowed dot products with the textbook weights: box (SMA), linear
    1..p (WMA), symmetric triangle 1..m..1 (SWMA)."""
    rng = np.random.default_rng(90 + p)
    x = 50 + np.cumsum(rng.normal(0, 1, 400))
    x[:6] = np.nan
    w = _windows(x[6:], p)
    tri = np.minimum(np.arange(1, p + 1), np.arange(p, 0, -1)).astype(float)
    forms = {"sma": (O.sma, np.ones(p)), "sma_b": (O.sma_batch_row, np.ones(p)),
             "swma": (O.swma, tri)}
    if p >= 2:
        forms["wma"] = (O.wma, np.arange(1, p + 1, dtype=float))
        forms["wma_b"] = (O.wma_batch_row, np.arange(1, p + 1, dtype=float))
    for name, (f, wt) in forms.items():
        out = f(x, p)
        assert np.isnan(out[:6 + p - 1]).all(), name
        ref = w @ (wt / wt.sum())
        # the batch prefix forms difference global prefix sums (sum i*x grows as
        # len^2), which costs a few more digits than the running sums
        rtol = 1e-9 if name.endswith("_b") else 1e-12
        assert np.allclose(out[6 + p - 1:], ref, rtol=rtol, atol=0), name


def test_swma_triangle_is_two_boxes():
    """Closed form: SWMA_p = SMA_b(SMA_a(x)) with a + b = p + 1."""
    rng = np.random.default_rng(95)
    x = 10 + rng.random(300)
    for p in (3, 4, 9, 10):
        a, b = ((p + 1) // 2, (p + 1) // 2) if p % 2 else (p // 2, p // 2 + 1)
        ref = O.sma(O.sma(x, a), b)
        assert np.allclose(O.swma(x, p)[p - 1:], ref[p - 1:], rtol=1e-12), p


def test_smma_wilders_closed_form():
    """SMMA and Wilders are both y_t = y_{t-1} + (x_t - y_{t-1}) / p seeded by the
    mean of the first p bars: check both against the unrolled sum
    y_t = b^k y_w + a sum_j b^(k-j) x_(w+j), computed with numpy powers."""
    rng = np.random.default_rng(96)
    x = 100 + np.cumsum(rng.normal(0, 1, 200))
    x[:3] = np.nan
    for p in (1, 2, 5, 14):
        w = 3 + p - 1
        a, b = 1.0 / p, 1.0 - 1.0 / p
        seed = np.mean(x[3:3 + p])
        k = np.arange(1, x.size - w)
        ref = np.empty(x.size - w)
        ref[0] = seed
        for i, kk in enumerate(k, start=1):
            j = np.arange(1, kk + 1)
            ref[i] = b ** kk * seed + a * np.sum(b ** (kk - j) * x[w + j])
        for f in (O.smma, O.wilders):
            out = f(x, p)
            assert np.isnan(out[:w]).all()
            assert np.allclose(out[w:], ref, rtol=1e-11), (f.__name__, p)


def test_fir_nonfinite_rules():
    """A NaN after first_valid poisons every later output of the running-sum
    and prefix forms, but only the window of the p <= 2 SWMA forms."""
    x = 10 + np.arange(60, dtype=float)
    x[30] = np.nan
    for f, p in ((O.sma, 4), (O.sma_batch_row, 4), (O.wma, 4), (O.wma_batch_row, 4),
                 (O.swma, 5), (O.smma, 4)):
        out = f(x, p)
        assert np.isfinite(out[p - 1:30]).all() and np.isnan(out[30:]).all(), f.__name__
    assert np.isnan(O.swma(x, 2)[30:32]).all() and np.isfinite(O.swma(x, 2)[32:]).all()
    assert np.isnan(O.swma(x, 1)[30]) and np.isfinite(O.swma(x, 1)[31:]).all()
    # Wilders rejects a non-finite bar inside the warm window (wilders.rs:223-230)
    with pytest.raises(O.OracleError):
        O.wilders(x[25:], 8)


def test_alma_avx2_baseline_matches_scalar_rows():
    """The AVX2 restatement of alma_row_avx2 (a CPU timing baseline) sums in
    another order than alma_row_scalar: same NaN prefixes, values within 1e-12
    relative, for periods with every tail length (p mod 4 = 0..3)."""
    rng = np.random.Generator(np.random.PCG64(44))
    x = 100.0 * np.exp(np.cumsum(rng.normal(0.0, 0.015, 3_000)))
    x[:5] = np.nan
    combos = [(p, o, s) for p in (4, 5, 6, 7, 9, 64, 250) for o in (0.1, 0.85) for s in (2.0, 6.0)]
    a = O.alma_batch_avx2(x, combos, nthreads=2)
    if a is None:
        pytest.skip("CPU without AVX2/FMA")
    b = O.alma_batch(x, combos, nthreads=2)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = np.isfinite(b)
    assert np.max(np.abs(a[m] - b[m]) / np.abs(b[m])) < 1e-12


def test_series_major_driver_matches_single_series_calls():
    """The series-major OpenMP driver (the CPU baselines' layout) returns, row
    for row, what the single-series oracle returns, with NaN prefixes and
    per-series status, for every kind it runs."""
    rng = np.random.default_rng(11)
    n, length = 5, 700
    c = 100.0 + np.cumsum(rng.normal(0.0, 0.5, (n, length)), axis=1)
    h, lo = c + rng.uniform(0.0, 1.0, c.shape), c - rng.uniform(0.0, 1.0, c.shape)
    for arr in (c, h, lo):
        arr[1, :37] = np.nan
    c[3, :] = np.nan
    h[3, :] = np.nan
    lo[3, :] = np.nan
    cases = [("ema", (20,), lambda x: O.ema(x, 20)), ("rsi", (14,), lambda x: O.rsi(x, 14)),
             ("pnr", (30, 70.0), lambda x: O.pnr(x, 30, 70.0)),
             ("medium_ad", (16,), lambda x: O.medium_ad(x, 16))]
    for kind, params, one in cases:
        got, st = O.series_major(kind, c, *params, nthreads=3)
        for j in range(n):
            if j == 3:
                assert st[j] != 0
                continue
            assert st[j] == 0
            np.testing.assert_array_equal(got[j], one(c[j]))
    (up, mid, dn), st = O.series_major("bb", c, 20, 2.0, 1.5, nthreads=2)
    for j in (0, 1, 2, 4):
        u, m, d = O.bollinger_bands(c[j], 20, 2.0, 1.5)
        np.testing.assert_array_equal(up[j], u)
        np.testing.assert_array_equal(mid[j], m)
        np.testing.assert_array_equal(dn[j], d)
    got, st = O.series_major("atr", h, 14, low=lo, close=c, nthreads=2)
    for j in (0, 1, 2, 4):
        np.testing.assert_array_equal(got[j], O.atr(h[j], lo[j], c[j], 14))


def test_batch_periods_driver_equals_single_series_rows():
    """The recurrence batches' CPU-baseline driver (OpenMP over periods) gives
    row for row what the single-series oracle gives."""
    from vectorta_amd import datagen
    rng = np.random.Generator(np.random.PCG64(12))
    c = datagen.gbm_close(3_000, 0.01, 3)
    _, h, lo, c = datagen.ohlc_from_close(c, rng)
    f = [a.astype(np.float64) for a in (h, lo, c)]
    ps = [2, 5, 14, 50, 300]
    refs = {"ema": O.ema_batch(f[2], ps), "rsi": O.rsi_batch(f[2], ps),
            "atr": O.atr_batch(*f, ps), "adx": O.adx_batch(*f, ps)}
    for kind, ref in refs.items():
        out, st = O.batch_periods(kind, f[2], ps, h=f[0], l=f[1], nthreads=2)
        assert (st == 0).all(), kind
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), kind
