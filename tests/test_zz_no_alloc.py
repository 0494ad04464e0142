"""The allocation-free ALMA sweep (vta_alma_batch_from_device_prices_noalloc)
through the C ABI: on two calls in a row, the second with a shorter series and
a smaller grid, its output is bit-identical to alma_batch_from_device_prices
(same kernels, same tables, now in the static scratch array) and within 1e-5
relative of the f64 oracle with exact NaN prefixes; grids whose tables exceed
the 16 MiB scratch and wrong output lengths are refused."""
import ctypes

import numpy as np
import pytest

import oracle as O
import vectorta_amd as vta
from vectorta_amd import datagen
from vectorta_amd._lib import VtaError, check, lib
from tests.conftest import require_gpu
from tests.helpers import max_rel_err, nan_mask_equal

pytestmark = pytest.mark.gpu


def _noalloc(alma, d_x, n, fv, sweep, d_out, out_len):
    rc = sweep.to_c()
    return lib().vta_alma_batch_from_device_prices_noalloc(
        alma.ctx.handle, ctypes.c_void_p(d_x.ptr), int(n), int(fv), ctypes.byref(rc),
        ctypes.c_void_p(d_out.ptr), int(out_len))


@pytest.mark.parametrize("policy", ["mfma", "fp32"])
def test_noalloc_matches_existing_entry_point(policy):
    require_gpu()
    alma = vta.CudaAlma(0, policy=policy)
    cases = [
        (datagen.gbm_close(40_000, 0.015, seed=91), 7,
         vta.AlmaBatchRange(period=(10, 250, 10), offset=(0.05, 0.95, 0.3),
                            sigma=(2.0, 10.0, 4.0))),
        (datagen.gbm_close(12_000, 0.01, seed=92), 3,
         vta.AlmaBatchRange(period=(9, 40, 3), offset=(0.85, 0.85, 0.0), sigma=(6.0, 6.0, 0.0))),
    ]
    for x, fv, sweep in cases:
        x[:fv] = np.nan
        combos = [(c.period, c.offset, c.sigma) for c in vta.expand_grid(sweep)]
        C, T = len(combos), x.size
        d_x = alma.ctx.to_device(x)
        d_out = alma.ctx.alloc(C * T * 4)
        check(lib().vta_memcpy_htod(alma.ctx.handle, ctypes.c_void_p(d_out.ptr),
                                    np.full(C * T, -7.0, np.float32).ctypes.data_as(ctypes.c_void_p),
                                    C * T * 4))
        check(_noalloc(alma, d_x, T, fv, sweep, d_out, C * T))
        got = d_out.copy_to_host().reshape(C, T)
        assert alma.selected_batch_kernel() == ("mfma_bf16x3" if policy == "mfma" else "fp32")
        ref_dev = alma.alma_batch_from_device_prices(d_x, T, fv, sweep).to_host()
        assert np.array_equal(got.view(np.uint32), ref_dev.view(np.uint32)), (policy, C, T)
        ref = O.alma_batch(x.astype(np.float64), combos, nthreads=0)
        nan_mask_equal(got, ref)
        err = max_rel_err(got, ref)
        print(f"noalloc {policy} {C} x {T}: max rel err {err:.3e}")
        assert err <= 1e-5
        d_out.free()
        d_x.free()


def test_noalloc_refuses_oversized_tables_and_wrong_lengths():
    require_gpu()
    alma = vta.CudaAlma(0, policy="fp32")
    x = datagen.gbm_close(6_000, 0.01, seed=93)
    d_x = alma.ctx.to_device(x)
    d_out = alma.ctx.alloc(16)
    # 1,000 combos x 5,000 taps of f32 weights = 20 MB of tables > 16 MiB
    big = vta.AlmaBatchRange(period=(5_000, 5_000, 0), offset=(0.0, 0.999, 0.001),
                             sigma=(6.0, 6.0, 0.0))
    st = _noalloc(alma, d_x, x.size, 0, big, d_out, 1_000 * x.size)
    assert st == 4 and b"scratch" in lib().vta_last_error()
    small = vta.AlmaBatchRange(period=(9, 9, 0))
    st = _noalloc(alma, d_x, x.size, 0, small, d_out, x.size - 1)
    assert st == 4 and b"wrong length" in lib().vta_last_error()
    with pytest.raises(VtaError):
        check(_noalloc(alma, d_x, x.size, x.size - 3, small, d_out, x.size))
