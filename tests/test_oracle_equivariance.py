"""Pins of the oracle by how each indicator must transform under a change of
price scale and origin, x -> a*x + b.  These follow from the indicators'
published definitions, not from the oracle's code, and a dropped term, a
wrong sign, a shifted index or an unnormalised weight breaks them:

- normalised linear filters (ALMA, EMA, SMA, WMA, SWMA, SMMA, Wilders):
  f(a*x + b) = a*f(x) + b, for a < 0 too;
- RSI (ratio of smoothed gains to losses): invariant for a > 0, and
  RSI(-x) = 100 - RSI(x) (gains and losses swap);
- ATR: scales by a for a > 0 (true range is a difference of prices);
- ADX: invariant for a > 0 (directional movement and TR scale together);
- Bollinger Bands: every band maps to a*band + b for a > 0 (the deviation
  scales by a);
- percentile nearest rank: selects an element, so it commutes exactly with
  any strictly increasing map (affine a > 0, exp);
- rolling median absolute deviation: scales by |a|.
Also: the time-major (one series per column) restatements that check the
many-series kernels equal the single-series functions column by column.
CPU only."""
import numpy as np
import pytest

import oracle as O

N = 3_000
LEAD = 5


def _walk(seed):
    rng = np.random.default_rng(seed)
    x = 100.0 * np.exp(np.cumsum(rng.normal(0.0, 0.01, N)))
    x[:LEAD] = np.nan
    return x


def _hlc(seed):
    rng = np.random.default_rng(seed)
    c = _walk(seed)
    spread = np.abs(rng.normal(0.0, 0.005, N))
    return c * (1 + spread), c * (1 - spread), c


def _close(got, want, scale):
    assert np.array_equal(np.isnan(got), np.isnan(want))
    m = ~np.isnan(want)
    err = np.max(np.abs(got[m] - want[m])) if m.any() else 0.0
    assert err <= 1e-10 * scale, err


LINEAR = {
    "alma": lambda x: O.alma(x, 20, 0.85, 6.0),
    "ema": lambda x: O.ema(x, 20),
    "sma": lambda x: O.sma(x, 20),
    "wma": lambda x: O.wma(x, 20),
    "swma": lambda x: O.swma(x, 9),
    "smma": lambda x: O.smma(x, 20),
    "wilders": lambda x: O.wilders(x, 20),
}


@pytest.mark.parametrize("name", sorted(LINEAR))
@pytest.mark.parametrize("a,b", [(3.5, -40.0), (-0.25, 7.0), (1.0, 1e3)])
def test_linear_filters_are_affine_equivariant(name, a, b):
    f = LINEAR[name]
    x = _walk(11)
    _close(f(a * x + b), a * f(x) + b, abs(a) * 200.0 + abs(b))


def test_rsi_scale_invariant_and_reflected():
    x = _walk(12)
    r = O.rsi(x, 14)
    _close(O.rsi(2.5 * x + 30.0, 14), r, 100.0)
    _close(O.rsi(-x, 14), 100.0 - r, 100.0)


def test_atr_scales_adx_invariant():
    h, lo, c = _hlc(13)
    a, b = 4.0, -150.0
    _close(O.atr(a * h + b, a * lo + b, a * c + b, 14), a * O.atr(h, lo, c, 14), a * 10.0)
    _close(O.adx(a * h + b, a * lo + b, a * c + b, 14), O.adx(h, lo, c, 14), 100.0)


def test_bollinger_bands_affine():
    x = _walk(14)
    a, b = 0.5, 12.0
    got = O.bollinger_bands(a * x + b, 20, 2.0, 1.5)
    want = O.bollinger_bands(x, 20, 2.0, 1.5)
    for g, w in zip(got, want):
        _close(g, a * w + b, a * 200.0 + b)


@pytest.mark.parametrize("pct", [0.0, 10.0, 50.0, 70.0, 100.0])
def test_pnr_commutes_with_increasing_maps(pct):
    x = _walk(15)
    r = O.pnr(x, 25, pct)
    for g in (lambda v: 3.0 * v + 5.0, np.exp):
        got = O.pnr(g(x), 25, pct)
        want = g(r)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        m = ~np.isnan(want)
        assert np.array_equal(got[m], want[m])


def test_medium_ad_scales_by_abs_a():
    x = _walk(16)
    base = O.medium_ad(x, 15)
    for a, b in ((2.0, 9.0), (-3.0, 1.0)):
        _close(O.medium_ad(a * x + b, 15), abs(a) * base, abs(a) * 200.0 + abs(b))


def _ragged_tm(seed, rows=400, cols=7):
    rng = np.random.default_rng(seed)
    tm = 50.0 * np.exp(np.cumsum(rng.normal(0.0, 0.01, (rows, cols)), axis=0))
    for s, lead in enumerate((0, 3, 17, 50, 1, 120, 9)[:cols]):
        tm[:lead, s] = np.nan
    return tm


def test_time_major_oracles_equal_their_columns():
    """The time-major restatements (one series per column, ragged leading NaNs,
    two threads) equal the single-series functions column by column, bit for
    bit: a stride or column-offset mistake in the gathering fails here."""
    tm = _ragged_tm(17)
    single = {
        "alma": (lambda x: O.alma(x, 12, 0.85, 6.0), (12, 0.85, 6.0)),
        "ema": (lambda x: O.ema(x, 12), (12,)),
        "rsi": (lambda x: O.rsi(x, 12), (12,)),
        "pnr": (lambda x: O.pnr(x, 12, 70.0), (12, 70.0)),
        "medium_ad": (lambda x: O.medium_ad(x, 12), (12,)),
    }
    for kind, (f, params) in single.items():
        got, st = O.many_series(kind, tm, *params, nthreads=2)
        assert not st.any(), (kind, st)
        for s in range(tm.shape[1]):
            np.testing.assert_array_equal(got[:, s], f(tm[:, s]), err_msg=f"{kind} col {s}")
    for kind in ("sma", "wma", "swma", "smma", "wilders"):
        got, st = O.ma_many(kind, tm, 9, nthreads=2)
        assert not st.any(), (kind, st)
        for s in range(tm.shape[1]):
            np.testing.assert_array_equal(got[:, s], getattr(O, kind)(tm[:, s], 9),
                                          err_msg=f"{kind} col {s}")
    rng = np.random.default_rng(18)
    spread = np.abs(rng.normal(0.0, 0.005, tm.shape))
    h, lo = tm * (1 + spread), tm * (1 - spread)
    for kind, f in (("atr", O.atr), ("adx", O.adx)):
        got, st = O.many_series_hlc(kind, h, lo, tm, 10, nthreads=2)
        assert not st.any(), (kind, st)
        for s in range(tm.shape[1]):
            np.testing.assert_array_equal(got[:, s], f(h[:, s], lo[:, s], tm[:, s], 10),
                                          err_msg=f"{kind} col {s}")
