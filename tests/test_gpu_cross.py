"""Double-ALMA crossover parity on the GPU.

Bar: every statistic bit-exact against the f64 oracle (or_cross_pairs on the
f64 batch ALMA rows of the same f32 prices): trades, bars long, bars short and
the sign checksum sum(pos_t * (t + 1)) pin every position; the pnl is a sum of
exact f64 price differences and must be equal too.
"""
import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import datagen
from tests.conftest import require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cx():
    require_gpu()
    return vta.DoubleAlmaCrossover(0)


def _oracle(x, fast, slow):
    """f64 rows of the distinct combos, then the pair statistics."""
    fc = [(c.period, c.offset, c.sigma) for c in vta.expand_grid(fast)]
    sc = [(c.period, c.offset, c.sigma) for c in vta.expand_grid(slow)]
    keys = sorted(set(fc) | set(sc))
    idx = {k: i for i, k in enumerate(keys)}
    xd = np.asarray(x, dtype=np.float64)
    rows = O.alma_batch(xd, keys, nthreads=0)
    pf = [idx[a] for a in fc for b in sc]
    ps = [idx[b] for a in fc for b in sc]
    return rows, idx, fc, sc, O.cross_pairs(rows, xd, pf, ps)


def _check(got, ref, what):
    st, pnl = got
    rst, rpnl = ref
    bad = np.flatnonzero((st != rst).any(axis=1) | (pnl != rpnl))
    assert bad.size == 0, (f"{what}: {bad.size} pairs differ, first {bad[:4]}: "
                           f"gpu {st[bad[:2]]} {pnl[bad[:2]]} ref {rst[bad[:2]]} {rpnl[bad[:2]]}")


def test_small_grid_bit_exact(cx):
    x = datagen.gbm_close(20_000, 0.015, seed=71)
    x[:5] = np.nan
    fast = vta.AlmaBatchRange(period=(5, 40, 5), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))
    slow = vta.AlmaBatchRange(period=(20, 120, 10), offset=(0.5, 0.85, 0.35), sigma=(6.0, 6.0, 0.0))
    got = cx.run(x, fast, slow)
    *_, ref = _oracle(x, fast, slow)
    assert got[0].shape == (8 * 22, 4)
    _check(got, ref, "small grid")
    # pairs with identical parameter sets are flat: all zeros
    assert (got[0][:, 0] >= 0).all()


def test_exact_and_near_ties_bit_exact(cx):
    """Piecewise-constant prices: inside flat stretches both ALMAs equal the
    level up to f64 rounding, so the sign is decided by the last bits of the
    f64 dot products (and is often an exact tie) -- the device f64 rows must
    follow the oracle's operation order exactly."""
    rng = np.random.Generator(np.random.PCG64(72))
    levels = 100.0 + np.cumsum(rng.normal(0, 0.5, size=60))
    x = np.repeat(levels, 300).astype(np.float32)[:17_000]
    fast = vta.AlmaBatchRange(period=(5, 60, 11), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))
    slow = vta.AlmaBatchRange(period=(8, 96, 8), offset=(0.3, 0.9, 0.6), sigma=(4.0, 6.0, 2.0))
    got = cx.run(x, fast, slow)
    *_, ref = _oracle(x, fast, slow)
    _check(got, ref, "near ties")
    # the flat stretches really produce ties or near ties in f64
    assert (ref[0][:, 1] + ref[0][:, 2] < 17_000 - 96).any()


def test_full_58300_pair_sweep(cx):
    """220 fast x 265 slow periods over the 200,000-bar sweep series: 1.166e10
    positions.  Exact identities on every pair; 400 sampled pairs against the
    oracle."""
    x = datagen.alma_sweep_series()
    fast = vta.AlmaBatchRange(**datagen.DOUBLE_ALMA_FAST)
    slow = vta.AlmaBatchRange(**datagen.DOUBLE_ALMA_SLOW)
    plan = vta.DoubleAlmaPlan(cx.ctx, x, fast, slow)
    assert plan.pairs == 58_300
    plan.launch()
    st, pnl = plan.results()
    plan.close()
    fp = np.arange(5, 225)
    sp = np.arange(30, 295)
    pf = np.repeat(fp, sp.size)
    ps = np.tile(sp, fp.size)
    warm = np.maximum(pf, ps) - 1
    assert (st[:, 1] + st[:, 2] <= x.size - warm).all()
    same = pf == ps
    assert (st[same] == 0).all() and (pnl[same] == 0).all()
    # swapped pairs (both periods in 30..224) mirror each other exactly
    i = (pf >= 30) & (ps <= 224) & (pf < ps)
    a = np.flatnonzero(i)
    b = (ps[a] - 5) * sp.size + (pf[a] - 30)
    assert (st[a, 0] == st[b, 0]).all() and (st[a, 1] == st[b, 2]).all()
    assert (st[a, 3] == -st[b, 3]).all() and (pnl[a] == -pnl[b]).all()
    # sampled pairs against the f64 oracle
    rng = np.random.Generator(np.random.PCG64(73))
    sel = np.sort(rng.choice(58_300, size=400, replace=False))
    periods = np.unique(np.concatenate([pf[sel], ps[sel]]))
    xd = x.astype(np.float64)
    rows = O.alma_batch(xd, [(int(p), 0.85, 6.0) for p in periods], nthreads=0)
    ix = {int(p): k for k, p in enumerate(periods)}
    rst, rpnl = O.cross_pairs(rows, xd, [ix[int(p)] for p in pf[sel]],
                              [ix[int(p)] for p in ps[sel]])
    _check((st[sel], pnl[sel]), (rst, rpnl), "58,300 sample")


def test_errors(cx):
    x = datagen.gbm_close(5_000, 0.015, seed=74)
    x[3_000] = np.nan
    r = vta.AlmaBatchRange(period=(5, 10, 5), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))
    with pytest.raises(vta.VtaError):
        cx.run(x, r, r)
    with pytest.raises(vta.VtaError):
        cx.run(x[:3_000], vta.AlmaBatchRange(period=(300, 300, 0)), r)
