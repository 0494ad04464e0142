"""Seeded synthetic market data for the benchmark workloads and parity tests.

All series are f32 (what the GPU consumes); the oracle sees the same values
widened to f64.  Generators use numpy's PCG64 with explicit seeds so the CPU
and GPU legs of a run see identical bytes.
"""
from __future__ import annotations

import numpy as np


def gbm_close(n: int, sd: float, seed: int, start: float = 100.0) -> np.ndarray:
    """Geometric random walk: close[0] = start, log-returns i.i.d. N(0, sd^2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.normal(0.0, sd, size=n)
    r[0] = 0.0
    return (start * np.exp(np.cumsum(r))).astype(np.float32)


def alma_single_series(seed: int = 1) -> np.ndarray:
    """Plain ALMA workload: 100,000 closes, log-return sd 0.01."""
    return gbm_close(100_000, 0.01, seed)


def alma_sweep_series(n: int = 200_000, seed: int = 2) -> np.ndarray:
    """ALMA parameter-sweep workload: 200,000 bars, log-return sd 0.015."""
    return gbm_close(n, 0.015, seed)


# Grid of the 1,250-combo sweep: periods 10..=250 step 10, offsets 0.05..=0.95
# step 0.1, sigmas 2..=10 step 2.
ALMA_SWEEP_RANGE = dict(period=(10, 250, 10), offset=(0.05, 0.95, 0.1), sigma=(2.0, 10.0, 2.0))


def ohlc_from_close(close: np.ndarray, rng: np.random.Generator, sd: float = 0.005):
    """high/low = close * (1 +/- |N(0, sd)|), open = previous close."""
    c = close.astype(np.float64)
    up = np.abs(rng.normal(0.0, sd, size=c.shape))
    dn = np.abs(rng.normal(0.0, sd, size=c.shape))
    high = c * (1.0 + up)
    low = c * (1.0 - dn)
    open_ = np.empty_like(c)
    open_[0] = c[0]
    open_[1:] = c[:-1]
    return (open_.astype(np.float32), high.astype(np.float32), low.astype(np.float32),
            c.astype(np.float32))


def many_series_ohlc(n_series: int = 4096, rows: int = 50_000, min_len: int = 2_000,
                     seed: int = 3, sd: float = 0.01):
    """Time-major [rows, n_series] OHLC with uneven lengths.

    Lengths are log-uniform in [min_len, rows]; each ticker is right-aligned
    (its last bar is row rows-1) and padded with leading NaN.
    Returns (open, high, low, close, lengths).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    lengths = np.exp(rng.uniform(np.log(min_len), np.log(rows), size=n_series)).astype(np.int64)
    lengths = np.clip(lengths, min_len, rows)
    o = np.full((rows, n_series), np.nan, dtype=np.float32)
    h = np.full_like(o, np.nan)
    lo = np.full_like(o, np.nan)
    c = np.full_like(o, np.nan)
    r = rng.normal(0.0, sd, size=(rows, n_series))
    logp = np.cumsum(r, axis=0)
    for s in range(n_series):
        L = int(lengths[s])
        start = rows - L
        lp = logp[start:, s] - logp[start, s]
        close = (100.0 * np.exp(lp)).astype(np.float32)
        oo, hh, ll, cc = ohlc_from_close(close, rng)
        o[start:, s] = oo
        h[start:, s] = hh
        lo[start:, s] = ll
        c[start:, s] = cc
    return o, h, lo, c, lengths


def many_series_with_jumps(n_series: int = 1024, rows: int = 50_000, seed: int = 4,
                           sd: float = 0.01, jump_frac: float = 0.01) -> np.ndarray:
    """Time-major [rows, n_series] random walks with 1% of bars replaced by
    +/-5-sigma jumps (in price terms) to stress rank statistics."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.normal(0.0, sd, size=(rows, n_series))
    r[0] = 0.0
    p = 100.0 * np.exp(np.cumsum(r, axis=0))
    mask = rng.random(size=(rows, n_series)) < jump_frac
    sign = np.where(rng.random(size=(rows, n_series)) < 0.5, -1.0, 1.0)
    p = np.where(mask, p * (1.0 + sign * 5.0 * sd), p)
    return p.astype(np.float32)


# Double-ALMA crossover sweep: 220 fast x 265 slow ALMA periods = 58,300 pairs
# (offset 0.85, sigma 6 -- the ALMA defaults) over the 200,000-bar sweep series.
DOUBLE_ALMA_FAST = dict(period=(5, 224, 1), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))
DOUBLE_ALMA_SLOW = dict(period=(30, 294, 1), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))


# ------------------------------------------------------------------ served mix

SERVED_KINDS = ("alma", "ema", "rsi", "bollinger_bands", "adx")


def ohlc_pool(n_series: int = 8, length: int = 1_000_000, seed: int = 5):
    """Base OHLC random walks the served requests slice from: [n_series, length]
    arrays (close, high, low) of f32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = np.empty((n_series, length), dtype=np.float32)
    h = np.empty_like(c)
    lo = np.empty_like(c)
    for s in range(n_series):
        cs = gbm_close(length, 0.01, seed * 1000 + s)
        _, hh, ll, cc = ohlc_from_close(cs, rng)
        c[s], h[s], lo[s] = cc, hh, ll
    return c, h, lo


def _loguniform_int(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def served_requests(n: int, seed: int = 6, pool_series: int = 8, pool_len: int = 1_000_000,
                    len_range=(1_000, 1_000_000), grid_range=(1, 2_000), reuse: float = 0.3):
    """Request specs of the served mix: indicator uniform over SERVED_KINDS,
    series length and grid size log-uniform, a fraction `reuse` naming a
    series already seen (same key, same data).  Each spec is a dict with
    kind, series (pool index), start, len, key and the parameter grid."""
    rng = np.random.Generator(np.random.PCG64(seed))
    seen = {False: [], True: []}          # keyed series by "needs high/low"
    specs = []
    next_key = 1
    for _ in range(n):
        kind = SERVED_KINDS[rng.integers(len(SERVED_KINDS))]
        hlc = kind == "adx"
        if seen[hlc] and rng.random() < reuse:
            base = seen[hlc][rng.integers(len(seen[hlc]))]
            s, start, L, key = base["series"], base["start"], base["len"], base["key"]
        else:
            L = min(_loguniform_int(rng, *len_range), pool_len)
            s = int(rng.integers(pool_series))
            start = int(rng.integers(0, pool_len - L + 1))
            key = next_key
            next_key += 1
            seen[hlc].append(dict(series=s, start=start, len=L, key=key))
        g = _loguniform_int(rng, *grid_range)
        spec = dict(kind=kind, series=s, start=start, len=L, key=key)
        if kind == "alma":
            q = max(1, min(250, (g + 9) // 10, L // 2 - 5))
            spec["alma"] = dict(period=(5, 5 + q - 1, 1), offset=(0.05, 0.95, 0.1),
                                sigma=(6.0, 6.0, 0.0)) if g > 10 else \
                dict(period=(5, 5 + min(g, L // 2 - 5) - 1, 1), offset=(0.85, 0.85, 0.0),
                     sigma=(6.0, 6.0, 0.0))
        elif kind in ("ema", "rsi"):
            q = max(1, min(g, L // 2 - 2))
            spec["period"] = (2, 2 + q - 1, 1)
        elif kind == "adx":
            q = max(1, min(g, L // 4 - 2))
            spec["period"] = (2, 2 + q - 1, 1)
        else:
            q = max(1, min((g + 4) // 5, L // 2 - 5))
            spec["bb"] = dict(period=(5, 5 + q - 1, 1), devup=(1.0, 3.0, 0.5),
                              devdn=(2.0, 2.0, 0.0))
        specs.append(spec)
    return specs


def submit_spec(server, spec, pool):
    """Submits one spec from served_requests to an IndicatorServer."""
    from .alma import AlmaBatchRange
    from .rolling import BollingerBandsBatchRange
    c, h, lo = pool
    s, a, L = spec["series"], spec["start"], spec["len"]
    kw = dict(key=spec["key"], cache=True)
    if "alma" in spec:
        kw["alma"] = AlmaBatchRange(**spec["alma"])
    if "period" in spec:
        kw["period"] = spec["period"]
    if "bb" in spec:
        kw["bb"] = BollingerBandsBatchRange(**spec["bb"])
    if spec["kind"] == "adx":
        return server.submit("adx", c[s, a:a + L], h[s, a:a + L], lo[s, a:a + L], **kw)
    return server.submit(spec["kind"], c[s, a:a + L], **kw)
