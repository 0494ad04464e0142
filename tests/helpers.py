"""Shared comparison helpers for the parity tests."""
import numpy as np


def nan_mask_equal(gpu, ref):
    """NaN positions must match bit-for-bit (warm-up prefixes, poison rules)."""
    g = np.isnan(np.asarray(gpu, dtype=np.float64))
    r = np.isnan(np.asarray(ref, dtype=np.float64))
    bad = np.flatnonzero(g != r)
    assert bad.size == 0, f"NaN mask differs at {bad[:10]} (gpu nan={g.ravel()[bad[:10]]})"


def max_rel_err(gpu, ref, floor=0.0):
    """max |gpu - ref| / max(|ref|, floor) over positions where ref is finite."""
    g = np.asarray(gpu, dtype=np.float64).ravel()
    r = np.asarray(ref, dtype=np.float64).ravel()
    m = np.isfinite(r)
    if not m.any():
        return 0.0
    den = np.maximum(np.abs(r[m]), floor)
    return float(np.max(np.abs(g[m] - r[m]) / den))


def assert_close(gpu, ref, rtol, atol=0.0, what=""):
    """|gpu - ref| <= atol + rtol*|ref| elementwise, NaN masks equal."""
    nan_mask_equal(gpu, ref)
    g = np.asarray(gpu, dtype=np.float64).ravel()
    r = np.asarray(ref, dtype=np.float64).ravel()
    m = np.isfinite(r)
    err = np.abs(g[m] - r[m])
    lim = atol + rtol * np.abs(r[m])
    bad = np.flatnonzero(err > lim)
    if bad.size:
        i = bad[0]
        idx = np.flatnonzero(m)[i]
        raise AssertionError(
            f"{what}: {bad.size} values out of tolerance (rtol={rtol}, atol={atol}); first at "
            f"flat index {idx}: gpu={g[idx]!r} ref={r[idx]!r} rel={err[i] / max(abs(r[idx]), 1e-300):.3e}")
