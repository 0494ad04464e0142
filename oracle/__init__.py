"""CPU f64 oracle for the HIP batch-indicator path -- TEST INFRASTRUCTURE ONLY.

Thin ctypes binding over ``liboracle.so`` (built from ``vta_oracle.c`` by the
Makefile beside it).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline).  The ``vectorta_amd`` product path never
imports it.

Each function follows the reference's scalar f64 kernel named in the C source
(file:line citations there).  Inputs are float64 arrays (the f32 series the GPU
sees, widened), outputs are float64 arrays with the reference's NaN prefixes.

Parity pins (see tests/test_oracle.py): the reference's own known-answer values
on the Bitfinex 4h fixture (alma.rs:1835-1856, ema.rs:971-983, rsi.rs:1142-1156,
atr.rs:1422-1436, adx.rs:2163-2175, bollinger_bands.rs:1705-1745,
percentile_nearest_rank.rs:1395-1425, medium_ad.rs:1478-1497, sma.rs:1625-1641,
wma.rs:855-883, swma.rs:1299-1318, smma.rs:915-931, wilders.rs:1226-1247), the tiny PNR
vectors (tests/python/test_utils.py:165-179), the ALMA re-input values
(test_utils.py:120-126), and closed-form / brute-force invariants.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

STATUS = {0: "ok", 1: "empty input", 2: "all values NaN", 3: "invalid period",
          4: "not enough valid data", 5: "invalid parameter"}


class OracleError(ValueError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {STATUS.get(code, code)}")
        self.code = code


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        sz = ctypes.c_size_t
        L.or_first_valid.argtypes = [dp, sz]
        L.or_first_valid.restype = sz
        L.or_alma_weights_single.argtypes = [sz, ctypes.c_double, ctypes.c_double, dp]
        L.or_alma_weights_single.restype = ctypes.c_double
        L.or_alma_weights_batch.argtypes = [sz, ctypes.c_double, ctypes.c_double, dp]
        L.or_alma_weights_batch.restype = ctypes.c_double
        L.or_alma_single.argtypes = [dp, sz, sz, ctypes.c_double, ctypes.c_double, dp]
        L.or_alma_batch.argtypes = [dp, sz, ctypes.POINTER(sz), dp, dp, sz, dp, ctypes.c_int]
        L.or_alma_batch_avx2.argtypes = [dp, sz, ctypes.POINTER(sz), dp, dp, sz, dp,
                                         ctypes.c_int]
        L.or_ema.argtypes = [dp, sz, sz, dp]
        L.or_rsi.argtypes = [dp, sz, sz, dp]
        L.or_atr.argtypes = [dp, dp, dp, sz, sz, dp]
        L.or_atr_batch_row.argtypes = [dp, dp, dp, sz, sz, dp]
        L.or_bb.argtypes = [dp, sz, sz, ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.or_pnr.argtypes = [dp, sz, sz, ctypes.c_double, dp]
        L.or_mad.argtypes = [dp, sz, sz, dp]
        L.or_adx.argtypes = [dp, dp, dp, sz, sz, dp]
        L.or_first_valid_hlc.argtypes = [dp, dp, dp, sz]
        L.or_first_valid_hlc.restype = sz
        L.or_many_series_1in.argtypes = [ctypes.c_int, dp, sz, sz, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, dp,
                                         ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.or_many_series_1in.restype = None
        L.or_many_series_3in.argtypes = [ctypes.c_int, dp, dp, dp, sz, sz, sz, dp,
                                         ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.or_many_series_3in.restype = None
        L.or_series_major.argtypes = [ctypes.c_int, dp, dp, dp, sz, sz, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, dp, dp, dp,
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.or_series_major.restype = None
        L.or_batch_periods.argtypes = [ctypes.c_int, dp, dp, dp, sz,
                                       ctypes.POINTER(ctypes.c_int), sz, dp,
                                       ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.or_batch_periods.restype = None
        L.or_cross_pairs.argtypes = [dp, sz, sz, dp, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), sz,
                                     ctypes.POINTER(ctypes.c_longlong), dp, ctypes.c_int]
        L.or_cross_pairs.restype = None
        for name in ("or_sma", "or_sma_batch_row", "or_wma", "or_wma_batch_row", "or_swma",
                     "or_smma", "or_wilders"):
            getattr(L, name).argtypes = [dp, sz, sz, dp]
        L.or_ma_many.argtypes = [ctypes.c_int, dp, sz, sz, sz, dp, ctypes.POINTER(ctypes.c_int),
                                 ctypes.c_int]
        L.or_ma_many.restype = None
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(code, what):
    if code != 0:
        raise OracleError(code, what)


def alma(x, period=9, offset=0.85, sigma=6.0):
    x, xp = _d(x)
    out = np.full(x.size, np.nan)
    _check(lib().or_alma_single(xp, x.size, period, offset, sigma, _d(out)[1]), "alma")
    return out


def alma_weights_batch(period, offset, sigma):
    w = np.zeros(period)
    inv = lib().or_alma_weights_batch(period, offset, sigma, _d(w)[1])
    return w, inv


def alma_batch_avx2(x, combos, nthreads=1):
    """alma_row_avx2 (alma.rs:1529-1573) over the same combos: a CPU timing
    baseline (different summation order).  None when the CPU lacks AVX2/FMA."""
    x, xp = _d(x)
    n = len(combos)
    per = (ctypes.c_size_t * n)(*[int(c[0]) for c in combos])
    off = np.array([c[1] for c in combos], dtype=np.float64)
    sig = np.array([c[2] for c in combos], dtype=np.float64)
    out = np.empty((n, x.size))
    st = lib().or_alma_batch_avx2(xp, x.size, per, _d(off)[1], _d(sig)[1], n, _d(out)[1],
                                  int(nthreads))
    if st == -1:
        return None
    _check(st, "alma_batch_avx2")
    return out


def alma_batch(x, combos, nthreads=1):
    """combos: sequence of (period, offset, sigma).  Returns [len(combos), len(x)]."""
    x, xp = _d(x)
    n = len(combos)
    per = (ctypes.c_size_t * n)(*[int(c[0]) for c in combos])
    off = np.array([c[1] for c in combos], dtype=np.float64)
    sig = np.array([c[2] for c in combos], dtype=np.float64)
    out = np.empty((n, x.size))
    _check(lib().or_alma_batch(xp, x.size, per, _d(off)[1], _d(sig)[1], n, _d(out)[1],
                               int(nthreads)), "alma_batch")
    return out


def ema(x, period=9):
    x, xp = _d(x)
    out = np.full(x.size, np.nan)
    _check(lib().or_ema(xp, x.size, period, _d(out)[1]), "ema")
    return out


def rsi(x, period=14):
    x, xp = _d(x)
    out = np.full(x.size, np.nan)
    _check(lib().or_rsi(xp, x.size, period, _d(out)[1]), "rsi")
    return out


def atr(h, l, c, length=14):
    h, hp = _d(h)
    l, lp = _d(l)
    c, cp = _d(c)
    out = np.full(c.size, np.nan)
    _check(lib().or_atr(hp, lp, cp, c.size, length, _d(out)[1]), "atr")
    return out


def atr_batch_row(h, l, c, length=14):
    h, hp = _d(h)
    l, lp = _d(l)
    c, cp = _d(c)
    out = np.full(c.size, np.nan)
    _check(lib().or_atr_batch_row(hp, lp, cp, c.size, length, _d(out)[1]), "atr_batch_row")
    return out


def bollinger_bands(x, period=20, devup=2.0, devdn=2.0):
    """Returns (upper, middle, lower), the order of BollingerBandsOutput."""
    x, xp = _d(x)
    u = np.full(x.size, np.nan)
    m = np.full(x.size, np.nan)
    lo = np.full(x.size, np.nan)
    _check(lib().or_bb(xp, x.size, period, devup, devdn, _d(u)[1], _d(m)[1], _d(lo)[1]), "bb")
    return u, m, lo


def pnr(x, length=15, percentage=50.0):
    x, xp = _d(x)
    out = np.full(x.size, np.nan)
    _check(lib().or_pnr(xp, x.size, length, percentage, _d(out)[1]), "pnr")
    return out


def medium_ad(x, period=5):
    x, xp = _d(x)
    out = np.full(x.size, np.nan)
    _check(lib().or_mad(xp, x.size, period, _d(out)[1]), "medium_ad")
    return out


def adx(h, l, c, period=14):
    h, hp = _d(h)
    l, lp = _d(l)
    c, cp = _d(c)
    out = np.full(c.size, np.nan)
    _check(lib().or_adx(hp, lp, cp, c.size, period, _d(out)[1]), "adx")
    return out


def first_valid(x):
    x, xp = _d(x)
    return int(lib().or_first_valid(xp, x.size))


_KIND = {"alma": 0, "ema": 1, "rsi": 2, "pnr": 3, "medium_ad": 4}


def many_series(kind, tm, p0, p1=0.0, p2=0.0, nthreads=0):
    """Applies a single-series oracle to every column of a [rows, cols] matrix."""
    tm, tp = _d(tm)
    rows, cols = tm.shape
    out = np.full((rows, cols), np.nan)
    st = (ctypes.c_int * cols)()
    lib().or_many_series_1in(_KIND[kind], tp, cols, rows, float(p0), float(p1), float(p2),
                             _d(out)[1], st, int(nthreads))
    return out, np.frombuffer(st, dtype=np.int32).copy()


_KIND3 = {"atr": 0, "adx": 1, "atr_batch_row": 2}


def many_series_hlc(kind, h_tm, l_tm, c_tm, period, nthreads=0):
    """ATR / ADX of every column of time-major [rows, cols] high/low/close."""
    h, hp = _d(h_tm)
    l, lp = _d(l_tm)
    c, cp = _d(c_tm)
    rows, cols = c.shape
    out = np.full((rows, cols), np.nan)
    st = (ctypes.c_int * cols)()
    lib().or_many_series_3in(_KIND3[kind], hp, lp, cp, cols, rows, int(period), _d(out)[1], st,
                             int(nthreads))
    return out, np.frombuffer(st, dtype=np.int32).copy()


_SM_KIND = {"ema": 0, "rsi": 1, "pnr": 2, "medium_ad": 3, "bb": 4, "atr": 5}


def series_major(kind, sm, p0, p1=0.0, p2=0.0, low=None, close=None, nthreads=0):
    """A single-series oracle over each row of a series-major [n_series, len]
    f64 matrix (OpenMP over rows), the layout of the CPU baselines.  For "atr"
    sm is the high matrix and low/close the others; "bb" returns (upper,
    middle, lower).  Returns (outputs, per-series status)."""
    a, ap = _d(sm)
    n, length = a.shape
    b, bp = _d(low) if low is not None else (a, ap)
    c, cp = _d(close) if close is not None else (a, ap)
    outs = [np.empty((n, length)) for _ in range(3 if kind == "bb" else 1)]
    ptrs = [_d(o)[1] for o in outs] + [None] * (3 - len(outs))
    st = (ctypes.c_int * n)()
    lib().or_series_major(_SM_KIND[kind], ap, bp, cp, n, length, float(p0), float(p1),
                          float(p2), ptrs[0], ptrs[1], ptrs[2], st, int(nthreads))
    status = np.frombuffer(st, dtype=np.int32).copy()
    return (tuple(outs) if kind == "bb" else outs[0]), status


def ema_batch(x, periods):
    return np.stack([ema(x, p) for p in periods])


def rsi_batch(x, periods):
    return np.stack([rsi(x, p) for p in periods])


def atr_batch(h, l, c, periods):
    """The batch path's prefix-sum seed (atr.rs:1101-1213)."""
    return np.stack([atr_batch_row(h, l, c, p) for p in periods])


def adx_batch(h, l, c, periods):
    return np.stack([adx(h, l, c, p) for p in periods])


_BP_KIND = {"ema": 0, "rsi": 1, "atr": 2, "adx": 3}


def batch_periods(kind, c, periods, h=None, l=None, nthreads=0):
    """One series x many periods with the single-series oracle per row
    (OpenMP over rows; "atr" is the batch path's row).  Returns
    (outputs [n, len], per-row status)."""
    cc, cp = _d(c)
    hh, hp = _d(h) if h is not None else (cc, cp)
    ll, lp = _d(l) if l is not None else (cc, cp)
    ps = np.ascontiguousarray(periods, dtype=np.int32)
    out = np.empty((ps.size, cc.size))
    st = (ctypes.c_int * ps.size)()
    lib().or_batch_periods(_BP_KIND[kind], hp, lp, cp, cc.size,
                           ps.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ps.size, _d(out)[1],
                           st, int(nthreads))
    return out, np.frombuffer(st, dtype=np.int32).copy()


def cross_pairs(rows, x, pair_fast, pair_slow, nthreads=0):
    """Double-ALMA crossover statistics of row pairs of f64 ALMA rows [nrows, len].
    Returns (stats int64 [npairs, 4] = trades, n_long, n_short, checksum; pnl f64)."""
    rows, rp = _d(rows)
    nrows, n = rows.shape
    x, xp = _d(x)
    pf = np.ascontiguousarray(pair_fast, dtype=np.int32)
    ps = np.ascontiguousarray(pair_slow, dtype=np.int32)
    npairs = pf.size
    stats = np.zeros((npairs, 4), dtype=np.int64)
    pnl = np.zeros(npairs)
    ip = ctypes.POINTER(ctypes.c_int)
    lib().or_cross_pairs(rp, nrows, n, xp, pf.ctypes.data_as(ip), ps.ctypes.data_as(ip), npairs,
                         stats.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _d(pnl)[1],
                         int(nthreads))
    return stats, pnl


def _ma(name, x, period):
    x, xp = _d(x)
    out = np.full(x.size, np.nan)
    _check(getattr(lib(), name)(xp, x.size, int(period), _d(out)[1]), name[3:])
    return out


def sma(x, period=9):
    """sma.rs:302-330 (running sum)."""
    return _ma("or_sma", x, period)


def sma_batch_row(x, period=9):
    """sma.rs:909-1021 (prefix-sum batch row)."""
    return _ma("or_sma_batch_row", x, period)


def wma(x, period=30):
    """wma.rs:243-320 (running weighted sum)."""
    return _ma("or_wma", x, period)


def wma_batch_row(x, period=30):
    """wma.rs:656-760 (prefix-sum batch row)."""
    return _ma("or_wma_batch_row", x, period)


def swma(x, period=5):
    """swma.rs:406-476 (two running box sums; also the batch row)."""
    return _ma("or_swma", x, period)


def smma(x, period=7):
    """smma.rs:243-271."""
    return _ma("or_smma", x, period)


def wilders(x, period=5):
    """wilders.rs:195-320."""
    return _ma("or_wilders", x, period)


_MA_KIND = {"sma": 0, "wma": 1, "swma": 2, "smma": 3, "wilders": 4, "sma_batch_row": 5,
            "wma_batch_row": 6}


def ma_many(kind, tm, period, nthreads=0):
    """One-period moving average of every column of a time-major [rows, cols] matrix."""
    tm, tp = _d(tm)
    rows, cols = tm.shape
    out = np.full((rows, cols), np.nan)
    st = (ctypes.c_int * cols)()
    lib().or_ma_many(_MA_KIND[kind], tp, cols, rows, int(period), _d(out)[1], st, int(nthreads))
    return out, np.frombuffer(st, dtype=np.int32).copy()


def ma_batch(kind, x, periods):
    """Rows of a one-series batch: [len(periods), len(x)]."""
    f = globals()[kind]
    return np.stack([f(x, p) for p in periods])
