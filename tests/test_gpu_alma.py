"""ALMA parity on the GPU: HIP kernels through the C ABI vs the f64 oracle.

Bars: NaN warm-up prefixes bit-exact; values within 1e-5 relative of the f64
oracle on the same f32 inputs (widened) for the price-like series, and within
the reference's own atol+rtol 5e-6 form (tests/alma_cuda.rs:19-27, 79) for its
zero-crossing sine series.
"""
import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import datagen
from tests.conftest import require_gpu
from tests.helpers import assert_close, max_rel_err, nan_mask_equal

pytestmark = pytest.mark.gpu

POLICIES = ["fp32", "mfma", "mfma_f16x3"]
SELECTED = {"fp32": "fp32", "mfma": "mfma_bf16x3", "mfma_f16x3": "mfma_f16x3"}


@pytest.fixture(scope="module")
def alma():
    require_gpu()
    return vta.CudaAlma(0)


def _combos(sweep):
    return [(c.period, c.offset, c.sigma) for c in vta.expand_grid(sweep)]


def _ref_series(n=2048):
    # tests/alma_cuda.rs:46-51: sin-based series with a 3-NaN lead
    x = np.full(n, np.nan)
    i = np.arange(3, n, dtype=np.float64)
    x[3:] = np.sin(i * 0.001) + 0.0001 * i
    return x.astype(np.float32)


@pytest.mark.parametrize("policy", POLICIES)
def test_batch_matches_oracle_reference_case(alma, policy):
    x = _ref_series()
    sweep = vta.AlmaBatchRange(period=(9, 32, 1), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))
    alma.set_policy(policy)
    arr = alma.alma_batch_dev(x, sweep)
    assert alma.selected_batch_kernel() == SELECTED[policy]
    g = arr.to_host()
    ref = O.alma_batch(x.astype(np.float64), _combos(sweep))
    assert g.shape == ref.shape == (24, 2048)
    assert_close(g, ref, rtol=5e-6, atol=5e-6, what=f"alma batch {policy}")


@pytest.mark.parametrize("policy", POLICIES)
def test_batch_wide_grid_price_series(alma, policy):
    x = datagen.gbm_close(20_000, 0.015, seed=21)
    x[:5] = np.nan
    sweep = vta.AlmaBatchRange(period=(5, 250, 7), offset=(0.05, 0.95, 0.3), sigma=(2.0, 10.0, 4.0))
    alma.set_policy(policy)
    g = alma.alma_batch_dev(x, sweep).to_host()
    ref = O.alma_batch(x.astype(np.float64), _combos(sweep))
    nan_mask_equal(g, ref)
    err = max_rel_err(g, ref)
    print(f"alma {policy} wide grid max rel err {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("policy", POLICIES)
def test_sweep_plan_full_size_sampled_rows(alma, policy):
    """200,000 bars x 1,250 combos (the benchmark workload) via the enqueue-only
    plan; every row's NaN prefix is checked, 26 rows against the oracle."""
    x = datagen.alma_sweep_series()
    sweep = vta.AlmaBatchRange(**datagen.ALMA_SWEEP_RANGE)
    combos = _combos(sweep)
    T, C = x.size, len(combos)
    plan = alma.sweep_plan(sweep, T, kernel=policy)
    assert plan.n_combos == C == 1250
    d_x = alma.ctx.to_device(x)
    d_out = alma.ctx.alloc(T * C * 4)
    plan.launch(d_x.ptr, 0, d_out.ptr)
    alma.synchronize()
    g = d_out.copy_to_host().reshape(C, T)
    warm = np.array([p - 1 for p, _, _ in combos])
    lead = np.argmax(~np.isnan(g), axis=1)
    assert np.array_equal(lead, warm)
    assert not np.isnan(g[np.arange(T)[None, :] >= warm[:, None]]).any()
    rows = list(range(0, C, 50)) + [C - 1]
    ref = O.alma_batch(x.astype(np.float64), [combos[r] for r in rows], nthreads=0)
    err = max_rel_err(g[rows], ref)
    print(f"alma {policy} 200k x 1250 sampled max rel err {err:.3e}")
    assert err <= 1e-5
    d_out.free()


@pytest.mark.parametrize("time_major", [False, True])
def test_batch_device_precomputed(alma, time_major):
    """alma_batch_device(_tm): caller-owned buffers, pre-normalised f32 weights."""
    x = datagen.gbm_close(5_000, 0.01, seed=22)
    x[:2] = np.nan
    combos = [(p, 0.85, 6.0) for p in range(9, 41)]
    C, T, maxp = len(combos), x.size, 40
    w = np.zeros((C, maxp), dtype=np.float32)
    for i, (p, o, s) in enumerate(combos):
        wd, inv = O.alma_weights_batch(p, o, s)
        w[i, :p] = (wd * inv).astype(np.float32)
    ctx = alma.ctx
    d_x, d_w = ctx.to_device(x), ctx.to_device(w)
    d_p = ctx.to_device(np.array([c[0] for c in combos], dtype=np.int32))
    d_inv = ctx.to_device(np.ones(C, dtype=np.float32))
    d_out = ctx.alloc(C * T * 4)
    alma.alma_batch_device(d_x, d_w, d_p, d_inv, maxp, T, C, 2, d_out, time_major=time_major)
    alma.synchronize()
    g = d_out.copy_to_host()
    g = g.reshape(T, C).T if time_major else g.reshape(C, T)
    ref = O.alma_batch(x.astype(np.float64), combos)
    nan_mask_equal(g, ref)
    assert max_rel_err(g, ref) <= 1e-5


@pytest.mark.parametrize("policy", POLICIES)
def test_interior_nan_and_inf(alma, policy):
    x = datagen.gbm_close(6_000, 0.01, seed=23).astype(np.float64)
    x[:4] = np.nan
    x[1500] = np.nan
    x[4000] = np.inf
    x = x.astype(np.float32)
    sweep = vta.AlmaBatchRange(period=(9, 64, 11), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))
    alma.set_policy(policy)
    g = alma.alma_batch_dev(x, sweep).to_host()
    ref = O.alma_batch(x.astype(np.float64), _combos(sweep))
    nan_mask_equal(g, ref)
    assert np.array_equal(np.isinf(g), np.isinf(ref))
    assert max_rel_err(g, ref) <= 1e-5


def test_many_series_one_param(alma):
    # tests/alma_cuda.rs:96-157 shape: 4 series, staggered NaN leads, period 14
    cols, rows = 4, 1024
    tm = np.full((rows, cols), np.nan)
    for j in range(cols):
        t = np.arange(j, rows, dtype=np.float64)
        xx = t + j * 0.1
        tm[j:, j] = np.cos(xx * 0.003) + 0.001 * xx
    tm = tm.astype(np.float32)
    p = vta.AlmaParams(14, 0.85, 6.0)
    g = alma.alma_multi_series_one_param_time_major_dev(tm, cols, rows, p)
    assert (g.rows, g.cols) == (rows, cols)
    ref, st = O.many_series("alma", tm.astype(np.float64), 14, 0.85, 6.0)
    assert (st == 0).all()
    assert_close(g.to_host(), ref, rtol=1e-5, atol=1e-5, what="alma many-series")


def test_many_series_long_period_price_like(alma):
    cols, rows = 70, 3000
    tm = np.stack([datagen.gbm_close(rows, 0.01, seed=100 + j) for j in range(cols)], axis=1)
    for j in range(cols):
        tm[: j * 7, j] = np.nan
    p = vta.AlmaParams(200, 0.3, 4.0)
    g = alma.alma_multi_series_one_param_time_major_dev(tm, cols, rows, p).to_host()
    ref, st = O.many_series("alma", tm.astype(np.float64), 200, 0.3, 4.0)
    nan_mask_equal(g, ref)
    assert max_rel_err(g, ref) <= 1e-5


def test_errors(alma):
    with pytest.raises(vta.VtaError) as e:
        alma.alma_batch_dev(np.array([], dtype=np.float32), vta.AlmaBatchRange())
    assert e.value.kind == "InvalidInput"
    with pytest.raises(vta.VtaError):
        alma.alma_batch_dev(np.full(100, np.nan, dtype=np.float32), vta.AlmaBatchRange())
    with pytest.raises(vta.VtaError):
        alma.alma_batch_dev(np.ones(50, dtype=np.float32), vta.AlmaBatchRange(period=(9, 60, 1)))
    with pytest.raises(vta.VtaError):
        alma.alma_batch_dev(np.ones(500, dtype=np.float32),
                            vta.AlmaBatchRange(period=(9, 9, 0), offset=(1.5, 1.5, 0.0)))
    # periods beyond the MFMA limit: forced MFMA is an invalid policy, auto falls back
    x = datagen.gbm_close(2_000, 0.01, seed=24)
    sweep = vta.AlmaBatchRange(period=(250, 300, 25))
    with pytest.raises(vta.VtaError) as e:
        alma.sweep_plan(sweep, x.size, kernel="mfma")
    assert e.value.kind == "InvalidPolicy"
    alma.set_policy("auto")
    g = alma.alma_batch_dev(x, sweep).to_host()
    assert alma.selected_batch_kernel() == "fp32"
    ref = O.alma_batch(x.astype(np.float64), _combos(sweep))
    nan_mask_equal(g, ref)
    assert max_rel_err(g, ref) <= 1e-5


@pytest.mark.parametrize("policy", ["fp32", "auto"])
def test_single_series_gbm_100k_vs_alma_scalar(alma, policy):
    """Plain ALMA on one series: 100,000 closes from the seeded geometric walk
    (log-return sd 0.01), period 9, offset 0.85, sigma 6, one combo through
    the batch entry point, against the single-series scalar path
    (alma_scalar via alma_with_kernel, alma.rs:342-350, 426-463, whose weight
    rounding w = exp((-d*d)*inv_s2) differs from the batch table's).  The
    8-element NaN prefix is exact and every value within 1e-5 relative."""
    x = datagen.alma_single_series()
    assert x.size == 100_000 and not np.isnan(x).any()
    sweep = vta.AlmaBatchRange(period=(9, 9, 0), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))
    alma.set_policy(policy)
    g = alma.alma_batch_dev(x, sweep).to_host()
    assert g.shape == (1, 100_000)
    assert alma.selected_batch_kernel() == "fp32"
    ref = O.alma(x.astype(np.float64), 9, 0.85, 6.0)
    assert np.isnan(ref[:8]).all() and not np.isnan(ref[8:]).any()
    nan_mask_equal(g[0], ref)
    err = max_rel_err(g[0], ref)
    print(f"alma single 100k p=9 ({policy}) max rel err {err:.3e}")
    assert err <= 1e-5
    # the same series through the many-series form (one column)
    m = alma.alma_multi_series_one_param_time_major_dev(x.reshape(-1, 1), 1, x.size,
                                                        vta.AlmaParams(9, 0.85, 6.0)).to_host()
    nan_mask_equal(m.ravel(), ref)
    assert max_rel_err(m.ravel(), ref) <= 1e-5


@pytest.mark.parametrize("policy", POLICIES)
def test_batch_from_device_prices(alma, policy):
    """alma_batch_from_device_prices (alma_wrapper.rs:671-681): prices already
    on the device, caller-supplied first_valid; device-resident output."""
    x = datagen.gbm_close(30_000, 0.015, seed=25)
    x[:6] = np.nan
    sweep = vta.AlmaBatchRange(period=(10, 250, 30), offset=(0.05, 0.95, 0.3),
                               sigma=(2.0, 10.0, 4.0))
    alma.set_policy(policy)
    d_x = alma.ctx.to_device(x)
    arr = alma.alma_batch_from_device_prices(d_x, x.size, 6, sweep)
    combos = _combos(sweep)
    assert (arr.rows, arr.cols) == (len(combos), x.size)
    g = arr.to_host()
    ref = O.alma_batch(x.astype(np.float64), combos)
    nan_mask_equal(g, ref)
    err = max_rel_err(g, ref)
    print(f"alma from_device_prices {policy} max rel err {err:.3e}")
    assert err <= 1e-5
    with pytest.raises(vta.VtaError):
        alma.alma_batch_from_device_prices(d_x, x.size, x.size - 5, sweep)


@pytest.mark.parametrize("policy", POLICIES)
def test_batch_into_host_f32(alma, policy):
    """alma_batch_into_host_f32 (alma_wrapper.rs:783-817): the same grid copied
    into a caller-owned host slice, row-major [combo][t]."""
    x = datagen.gbm_close(25_000, 0.015, seed=26)
    x[:2] = np.nan
    sweep = vta.AlmaBatchRange(period=(9, 200, 17), offset=(0.15, 0.85, 0.35),
                               sigma=(3.0, 9.0, 3.0))
    combos = _combos(sweep)
    alma.set_policy(policy)
    out = np.full(len(combos) * x.size, -1.0, dtype=np.float32)
    rows, cols, got = alma.alma_batch_into_host_f32(x, sweep, out)
    assert (rows, cols) == (len(combos), x.size)
    assert [(c.period, c.offset, c.sigma) for c in got] == combos
    ref = O.alma_batch(x.astype(np.float64), combos)
    g = out.reshape(rows, cols)
    nan_mask_equal(g, ref)
    assert max_rel_err(g, ref) <= 1e-5
    # the same values as the device-resident form
    d = alma.alma_batch_dev(x, sweep).to_host()
    assert np.array_equal(d.view(np.uint32), g.view(np.uint32))
    with pytest.raises(vta.VtaError):
        alma.alma_batch_into_host_f32(x, sweep, np.empty(out.size - 1, dtype=np.float32))


def test_many_series_period_above_lds_floats(alma):
    """A period longer than the LDS holds floats (17,000 > 65,536 B / 4) takes
    the direct many-series kernel instead of sizing an LDS tile from a wrapped
    unsigned difference."""
    rows = 20_000
    x = datagen.gbm_close(rows, 0.01, seed=27)
    x[:5] = np.nan
    p = vta.AlmaParams(17_000, 0.85, 6.0)
    g = alma.alma_multi_series_one_param_time_major_dev(x.reshape(-1, 1), 1, rows, p).to_host()
    ref, st = O.many_series("alma", x.reshape(-1, 1).astype(np.float64), 17_000, 0.85, 6.0)
    assert (st == 0).all()
    nan_mask_equal(g, ref)
    assert max_rel_err(g, ref) <= 1e-5


def test_f16x3_zero_crossing_noise_guard(alma):
    """Split-F16 sweep on white noise (zero crossings everywhere): a tile whose
    split-error bound exceeds 2e-6 |y| for any output is recomputed with the
    FP32 dot, so every output meets the atol + rtol 5e-6 form; AUTO on host
    prices that change sign resolves to the FP32 kernel."""
    rng = np.random.default_rng(5)
    z = rng.normal(0.0, 1.0, 50_000).astype(np.float32)
    sweep = vta.AlmaBatchRange(period=(10, 250, 40), offset=(0.05, 0.95, 0.3),
                               sigma=(2.0, 10.0, 4.0))
    ref = O.alma_batch(z.astype(np.float64), _combos(sweep), nthreads=0)
    try:
        for policy, expect in (("mfma_f16x3", "mfma_f16x3"), ("auto", "fp32")):
            alma.set_policy(policy)
            g = alma.alma_batch_dev(z, sweep).to_host()
            assert alma.selected_batch_kernel() == expect
            nan_mask_equal(g, ref)
            assert_close(g, ref, rtol=5e-6, atol=5e-6, what=f"noise {policy}")
    finally:
        alma.set_policy("auto")


def test_auto_selects_checked_split_on_price_series(alma):
    """AUTO on a positive price series with 16+ combos and periods <= 256 runs
    the split-F16 sweep, within 1e-5 of the oracle with exact NaN prefixes."""
    x = datagen.gbm_close(30_000, 0.015, seed=71)
    x[:6] = np.nan
    sweep = vta.AlmaBatchRange(period=(10, 250, 20), offset=(0.05, 0.95, 0.45),
                               sigma=(2.0, 10.0, 4.0))
    alma.set_policy("auto")
    g = alma.alma_batch_dev(x, sweep).to_host()
    assert alma.selected_batch_kernel() == "mfma_f16x3"
    ref = O.alma_batch(x.astype(np.float64), _combos(sweep), nthreads=0)
    nan_mask_equal(g, ref)
    err = max_rel_err(g, ref)
    print(f"alma auto (split-F16) max rel err {err:.3e}")
    assert err <= 1e-5
