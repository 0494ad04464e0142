"""Request server: results identical alone and in a concurrent mix, equal to the
direct entry points, HBM-budget admission and the series cache.

Bars: a result is bit-identical alone, in the mix and through the direct
entry point; served results are also checked against the f64 oracle (NaN
prefixes exact, values within 1e-5 relative; RSI/ADX denominators floored at
1e-3 on their 0-100 scale)."""
import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import datagen
from tests.conftest import require_gpu
from tests.helpers import max_rel_err, nan_mask_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    require_gpu()
    return datagen.ohlc_pool(n_series=4, length=120_000, seed=81)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _run_all(server, specs, pool, concurrent):
    """Returns {index: [host outputs]} and the served results."""
    outs, results = {}, {}
    if concurrent:
        ids = [datagen.submit_spec(server, s, pool) for s in specs]
        for i, rid in enumerate(ids):
            r = server.wait(rid)
            assert r.ok, (i, specs[i], r.status)
            outs[i] = [server.to_host(r, k) for k in range(r.n_outputs)]
            results[i] = r
            server.release(rid)
    else:
        for i, s in enumerate(specs):
            rid = datagen.submit_spec(server, s, pool)
            r = server.wait(rid)
            assert r.ok, (i, s, r.status)
            outs[i] = [server.to_host(r, k) for k in range(r.n_outputs)]
            results[i] = r
            server.release(rid)
    return outs, results


def test_mix_equals_alone(pool):
    specs = datagen.served_requests(40, seed=82, pool_series=4, pool_len=120_000,
                                    len_range=(2_000, 120_000), grid_range=(1, 60))
    alone = vta.IndicatorServer(0, n_slots=1, hbm_budget=8 << 30)
    a_out, _ = _run_all(alone, specs, pool, concurrent=False)
    alone.close()
    mix = vta.IndicatorServer(0, n_slots=4, hbm_budget=8 << 30)
    m_out, m_res = _run_all(mix, specs, pool, concurrent=True)
    st = mix.stats()
    mix.close()
    for i in range(len(specs)):
        assert len(a_out[i]) == len(m_out[i])
        for a, m in zip(a_out[i], m_out[i]):
            assert np.array_equal(_bits(a), _bits(m)), (i, specs[i]["kind"])
    assert st["peak_in_flight"] >= 2, st          # requests really overlapped
    assert st["queued"] == 0 and st["in_flight"] == 0


def test_served_equals_direct_entry_points(pool):
    c, h, lo = pool
    x, hh, ll = c[0, :50_000], h[0, :50_000], lo[0, :50_000]
    srv = vta.IndicatorServer(0, n_slots=2, hbm_budget=4 << 30)
    sweep = vta.AlmaBatchRange(period=(5, 60, 5), offset=(0.5, 0.85, 0.35), sigma=(6.0, 6.0, 0.0))
    bb = vta.BollingerBandsBatchRange(period=(5, 40, 7), devup=(1.5, 2.5, 0.5),
                                      devdn=(2.0, 2.0, 0.0))
    cases = [
        ("alma", dict(alma=sweep), lambda: [vta.CudaAlma(0).alma_batch_dev(x, sweep).to_host()]),
        ("ema", dict(period=(3, 40, 3)),
         lambda: [vta.CudaEma(0).ema_batch_dev(x, (3, 40, 3)).to_host()]),
        ("rsi", dict(period=(2, 30, 4)),
         lambda: [vta.CudaRsi(0).rsi_batch_dev(x, (2, 30, 4)).to_host()]),
        ("bollinger_bands", dict(bb=bb),
         lambda: [d.to_host() for d in vta.CudaBollingerBands(0).bollinger_bands_batch_dev(x, bb)]),
        ("adx", dict(period=(5, 30, 5)),
         lambda: [vta.CudaAdx(0).adx_batch_dev(hh, ll, x, (5, 30, 5)).to_host()]),
    ]
    for kind, kw, direct in cases:
        if kind == "adx":
            rid = srv.submit(kind, x, hh, ll, **kw)
        else:
            rid = srv.submit(kind, x, **kw)
        r = srv.wait(rid)
        assert r.ok, (kind, r.status)
        got = [srv.to_host(r, k) for k in range(r.n_outputs)]
        srv.release(rid)
        ref = direct()
        assert len(got) == len(ref)
        for g, d in zip(got, ref):
            assert np.array_equal(_bits(g), _bits(d)), kind
    srv.close()


def test_budget_admission_and_cache(pool):
    c, _, _ = pool
    x = c[1, :100_000]
    # each request: 20 EMA rows x 100k bars = 8 MB of outputs; budget 40 MB
    budget = 40 << 20
    srv = vta.IndicatorServer(0, n_slots=4, hbm_budget=budget)
    ids = [srv.submit("ema", x, key=7, cache=True, period=(2, 21, 1)) for _ in range(12)]
    first = None
    for rid in ids:
        r = srv.wait(rid)
        assert r.ok
        g = srv.to_host(r)
        if first is None:
            first = g
        assert np.array_equal(_bits(g), _bits(first))
        srv.release(rid)
    st = srv.stats()
    assert st["peak_in_use"] <= budget
    assert st["cached_series"] == 1
    # a request that can never fit fails with OutOfMemory instead of blocking
    rid = srv.submit("ema", x, period=(2, 111, 1))        # 44 MB of outputs > 40 MiB
    r = srv.wait(rid, timeout_ms=60_000)
    assert r.status == 2
    srv.release(rid)
    # an invalid request reports its error through the result
    rid = srv.submit("rsi", x[:50], period=(60, 60, 0))
    r = srv.wait(rid, timeout_ms=60_000)
    assert r.status == 4
    srv.release(rid)
    srv.close()


def test_small_requests_are_packed(pool):
    """Small EMA / RSI / ADX / Bollinger Bands requests queued behind others
    are served together in packed launches (one launch per scan pass for the
    group); every result equals its direct entry point bit for bit, packed or
    not."""
    c, h, lo = pool
    rng = np.random.Generator(np.random.PCG64(83))
    srv = vta.IndicatorServer(0, n_slots=1, hbm_budget=4 << 30)
    specs, ids = [], []
    for i in range(48):
        kind = ("ema", "rsi", "adx", "bollinger_bands")[i % 4]
        s = int(rng.integers(4))
        L = int(rng.integers(2_000, 20_000))
        a = int(rng.integers(0, 120_000 - L))
        p0 = int(rng.integers(2, 10))
        period = (p0, p0 + int(rng.integers(0, 40)), 1)
        x, hh, ll = c[s, a:a + L], h[s, a:a + L], lo[s, a:a + L]
        if kind == "adx":
            ids.append(srv.submit(kind, x, hh, ll, period=period))
        elif kind == "bollinger_bands":
            period = vta.BollingerBandsBatchRange(period=period, devup=(1.5, 2.5, 0.5),
                                                  devdn=(2.0, 2.0, 0.0))
            ids.append(srv.submit(kind, x, bb=period))
        else:
            ids.append(srv.submit(kind, x, period=period))
        specs.append((kind, x, hh, ll, period))
    direct = {"ema": vta.CudaEma(0), "rsi": vta.CudaRsi(0), "adx": vta.CudaAdx(0),
              "bollinger_bands": vta.CudaBollingerBands(0)}
    for rid, (kind, x, hh, ll, period) in zip(ids, specs):
        r = srv.wait(rid)
        assert r.ok, (kind, r.status)
        got = [srv.to_host(r, k) for k in range(r.n_outputs)]
        srv.release(rid)
        if kind == "adx":
            ref = [direct[kind].adx_batch_dev(hh, ll, x, period).to_host()]
        elif kind == "bollinger_bands":
            ref = [d.to_host() for d in direct[kind].bollinger_bands_batch_dev(x, period)]
        else:
            ref = [getattr(direct[kind], f"{kind}_batch_dev")(x, period).to_host()]
        assert len(got) == len(ref)
        for g_, r_ in zip(got, ref):
            assert np.array_equal(_bits(g_), _bits(r_)), (kind, len(x))
    st = srv.stats()
    srv.close()
    print(f"packs {st['packs']}, packed requests {st['packed_requests']}")
    assert st["packs"] >= 1 and st["packed_requests"] >= 2 * st["packs"], st


def test_small_alma_requests_are_packed(pool):
    """Small ALMA requests queued on one slot are served in packed launches
    (vta_alma_pack's kernels over the cached per-grid plans): grids of fewer
    than 16 combos run the FP32 form, larger ones the split-F16 form (AUTO),
    each group in one launch; every result equals alma_batch_dev bit for bit."""
    c, _, _ = pool
    rng = np.random.Generator(np.random.PCG64(84))
    srv = vta.IndicatorServer(0, n_slots=1, hbm_budget=4 << 30)
    specs, ids = [], []
    for i in range(40):
        s = int(rng.integers(4))
        L = int(rng.integers(1_000, 30_000))
        a = int(rng.integers(0, 120_000 - L))
        p0 = int(rng.integers(5, 30))
        if i % 2:
            sweep = vta.AlmaBatchRange(period=(p0, p0 + 5 * int(rng.integers(0, 3)), 5))
        else:
            sweep = vta.AlmaBatchRange(period=(p0, p0 + 120, 20), offset=(0.05, 0.95, 0.3),
                                       sigma=(3.0, 9.0, 6.0))
        x = c[s, a:a + L]
        ids.append(srv.submit("alma", x, alma=sweep))
        specs.append((x, sweep))
    direct = vta.CudaAlma(0)
    for rid, (x, sweep) in zip(ids, specs):
        r = srv.wait(rid)
        assert r.ok, r.status
        got = srv.to_host(r, 0)
        srv.release(rid)
        ref = direct.alma_batch_dev(x, sweep).to_host()
        assert np.array_equal(_bits(got), _bits(ref)), (len(x), sweep)
    st = srv.stats()
    srv.close()
    print(f"packs {st['packs']}, packed requests {st['packed_requests']}")
    assert st["packs"] >= 1 and st["packed_requests"] >= 2 * st["packs"], st


def _oracle_rows(spec_kind, x, hh, ll, combos):
    """Oracle outputs (list per output) of the given combos of one request."""
    xd = x.astype(np.float64)
    if spec_kind == "alma":
        return [O.alma_batch(xd, combos, nthreads=0)]
    if spec_kind == "ema":
        return [O.ema_batch(xd, combos)]
    if spec_kind == "rsi":
        return [O.rsi_batch(xd, combos)]
    if spec_kind == "adx":
        return [O.adx_batch(hh.astype(np.float64), ll.astype(np.float64), xd, combos)]
    bands = [O.bollinger_bands(xd, p, u, d) for p, u, d in combos]
    return [np.stack([b[k] for b in bands]) for k in range(3)]


def _grid(spec):
    if "alma" in spec:
        return [(c.period, c.offset, c.sigma)
                for c in vta.expand_grid(vta.AlmaBatchRange(**spec["alma"]))]
    if "bb" in spec:
        return vta.CudaBollingerBands.expand_grid(vta.BollingerBandsBatchRange(**spec["bb"]))
    return vta.CudaEma.expand_grid(spec["period"])


def _check_vs_oracle(spec, outs, pool, rows_to_check):
    c, h, lo = pool
    s, a, L = spec["series"], spec["start"], spec["len"]
    x, hh, ll = c[s, a:a + L], h[s, a:a + L], lo[s, a:a + L]
    grid = _grid(spec)
    sel = sorted(set(r for r in rows_to_check(len(grid))))
    ref = _oracle_rows(spec["kind"], x, hh, ll, [grid[r] for r in sel])
    floor = 1e-3 if spec["kind"] in ("rsi", "adx") else 0.0
    assert len(ref) == len(outs)
    for g, r in zip(outs, ref):
        assert g.shape == (len(grid), L)
        nan_mask_equal(g[sel], r)
        err = max_rel_err(g[sel], r, floor=floor)
        assert err <= 1e-5, (spec["kind"], L, len(grid), err)
    return len(sel)


def test_served_results_match_oracle(pool):
    """One served request of each kind, every combo against the oracle."""
    c, h, lo = pool
    srv = vta.IndicatorServer(0, n_slots=2, hbm_budget=4 << 30)
    specs = [
        dict(kind="alma", series=0, start=1_000, len=60_000, key=1,
             alma=dict(period=(5, 250, 15), offset=(0.05, 0.95, 0.3), sigma=(2.0, 10.0, 4.0))),
        dict(kind="ema", series=1, start=7, len=90_000, key=2, period=(2, 400, 21)),
        dict(kind="rsi", series=2, start=30_000, len=80_000, key=3, period=(2, 300, 17)),
        dict(kind="bollinger_bands", series=3, start=5, len=70_000, key=4,
             bb=dict(period=(5, 200, 13), devup=(1.0, 3.0, 0.5), devdn=(2.0, 2.0, 0.0))),
        dict(kind="adx", series=0, start=2_000, len=100_000, key=5, period=(2, 200, 11)),
    ]
    ids = [datagen.submit_spec(srv, sp, pool) for sp in specs]
    for sp, rid in zip(specs, ids):
        r = srv.wait(rid)
        assert r.ok, (sp["kind"], r.status)
        outs = [srv.to_host(r, k) for k in range(r.n_outputs)]
        srv.release(rid)
        n = _check_vs_oracle(sp, outs, pool, lambda C: range(C))
        print(f"served {sp['kind']}: {n} combos x {sp['len']} bars vs oracle ok")
    srv.close()


@pytest.fixture(scope="module")
def pool_1m():
    require_gpu()
    return datagen.ohlc_pool(n_series=4, length=1_000_000, seed=5)


def test_mix_at_served_workload_shape(pool_1m):
    """The served workload's real shape: 24 requests drawn from the served-mix
    generator (series 1k-1M bars, grids 1-2,000 combos, 30% cached-series
    reuse); this draw holds a 506,132-bar ALMA request and five requests of
    1,072-1,920 combos over all five kinds (2.3 GB of outputs).  Every result
    is bit-identical alone (one slot, one at a time) and in the concurrent mix
    (four slots), and three combos of every request (first, middle, last)
    match the oracle."""
    specs = datagen.served_requests(24, seed=102, pool_series=4, pool_len=1_000_000)
    grids = [len(_grid(sp)) for sp in specs]
    assert max(sp["len"] for sp in specs) >= 500_000
    assert max(grids) >= 1_000
    assert {sp["kind"] for sp in specs} == set(datagen.SERVED_KINDS)
    mix = vta.IndicatorServer(0, n_slots=4, hbm_budget=16 << 30)
    m_out, _ = _run_all(mix, specs, pool_1m, concurrent=True)
    st = mix.stats()
    mix.close()
    assert st["peak_in_flight"] >= 2, st
    alone = vta.IndicatorServer(0, n_slots=1, hbm_budget=16 << 30)
    checked = 0
    for i, sp in enumerate(specs):
        rid = datagen.submit_spec(alone, sp, pool_1m)
        r = alone.wait(rid)
        assert r.ok, (i, sp["kind"], r.status)
        outs = [alone.to_host(r, k) for k in range(r.n_outputs)]
        alone.release(rid)
        assert len(outs) == len(m_out[i])
        for a, m in zip(outs, m_out[i]):
            assert np.array_equal(_bits(a), _bits(m)), (i, sp["kind"], sp["len"], grids[i])
        checked += _check_vs_oracle(sp, outs, pool_1m, lambda C: (0, C // 2, C - 1))
        m_out[i] = None
    alone.close()
    print(f"served mix: {len(specs)} requests alone == mix, {checked} combos vs oracle")
