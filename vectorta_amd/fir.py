"""SMA / WMA / SWMA operator surfaces: mirrors of CudaSma
(src/cuda/moving_averages/sma_wrapper.rs), CudaWma (wma_wrapper.rs) and CudaSwma
(swma_wrapper.rs) over the finite-window kernels (kernels/fir.hip)."""
from __future__ import annotations

import ctypes
from typing import List, Tuple

import numpy as np

from ._lib import DeviceArrayF32C, check, lib
from .device import DeviceArrayF32, DeviceBuffer, HipContext, f32_ptr, require_len
from .rec import _expand, _rng


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _vp(buf):
    return ctypes.c_void_p(buf.ptr) if buf is not None else None


class _Fir:
    _name = ""

    def __init__(self, device_id: int = 0):
        self.ctx = HipContext(device_id)

    def synchronize(self) -> None:
        self.ctx.synchronize()

    @classmethod
    def expand_grid(cls, period: Tuple[int, int, int]) -> List[int]:
        return _expand(getattr(lib(), f"vta_{cls._name}_expand_grid"), period)

    def _batch_dev(self, prices, period) -> DeviceArrayF32:
        x = _f32(prices)
        rc = _rng(period)
        out = DeviceArrayF32C()
        check(getattr(lib(), f"vta_{self._name}_batch_dev")(self.ctx.handle, f32_ptr(x), x.size,
                                                            ctypes.byref(rc), ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)

    def _batch_device(self, d_prices: DeviceBuffer, d_periods: DeviceBuffer, series_len: int,
                      n_combos: int, first_valid: int, max_period: int,
                      d_out: DeviceBuffer) -> None:
        require_len(d_prices, series_len, "prices")
        require_len(d_periods, n_combos, "periods")
        require_len(d_out, int(n_combos) * int(series_len), "output")
        check(getattr(lib(), f"vta_{self._name}_batch_device")(
            self.ctx.handle, _vp(d_prices), _vp(d_periods), int(series_len), int(n_combos),
            int(first_valid), int(max_period), _vp(d_out)))

    def _ms_dev(self, data_tm, cols, rows, period) -> DeviceArrayF32:
        x = _f32(data_tm).reshape(-1)
        out = DeviceArrayF32C()
        fn = getattr(lib(), f"vta_{self._name}_many_series_one_param_time_major_dev")
        check(fn(self.ctx.handle, f32_ptr(x), int(cols), int(rows), int(period),
                 ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)


class CudaSma(_Fir):
    _name = "sma"

    def sma_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev(prices_f32, period)

    def sma_batch_device(self, *args) -> None:
        """Enqueue-only SMA batch (no reference counterpart): (d_prices,
        d_periods, series_len, n_combos, first_valid, max_period, d_out)."""
        self._batch_device(*args)

    def sma_batch_dev_from_device(self, d_prices: DeviceBuffer, len_: int, periods,
                                  first_valid: int) -> DeviceArrayF32:
        """sma_wrapper.rs:830: device prices, host combos (periods)."""
        require_len(d_prices, len_, "prices")
        ps = (ctypes.c_size_t * max(len(periods), 1))(*[int(p) for p in periods])
        out = DeviceArrayF32C()
        check(lib().vta_sma_batch_dev_from_device(self.ctx.handle, _vp(d_prices), int(len_), ps,
                                                  len(periods), int(first_valid),
                                                  ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)

    def sma_multi_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev(data_tm, cols, rows, period)

    def sma_many_series_one_param_device(self, d_prices_tm, d_first_valids, cols, rows, period,
                                         d_out_tm) -> None:
        """Enqueue-only SMA many-series form (no reference counterpart)."""
        _need_ms(d_prices_tm, d_first_valids, d_out_tm, cols, rows)
        check(lib().vta_sma_many_series_one_param_device(
            self.ctx.handle, _vp(d_prices_tm), _vp(d_first_valids), int(cols), int(rows),
            int(period), _vp(d_out_tm)))

    def sma_multi_series_one_param_time_major_dev_from_device(self, d_prices_tm, d_first_valids,
                                                              cols, rows,
                                                              period) -> DeviceArrayF32:
        """sma_wrapper.rs:982."""
        require_len(d_prices_tm, int(cols) * int(rows), "time-major input")
        require_len(d_first_valids, cols, "first_valids")
        out = DeviceArrayF32C()
        check(lib().vta_sma_multi_series_one_param_time_major_dev_from_device(
            self.ctx.handle, _vp(d_prices_tm), _vp(d_first_valids), int(cols), int(rows),
            int(period), ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)


def _need_ms(d_prices_tm, d_first_valids, d_out_tm, cols, rows) -> None:
    require_len(d_prices_tm, int(cols) * int(rows), "time-major input")
    require_len(d_out_tm, int(cols) * int(rows), "time-major output")
    require_len(d_first_valids, cols, "first_valids")


class CudaWma(_Fir):
    _name = "wma"

    def wma_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev(prices_f32, period)

    def wma_batch_device(self, *args) -> None:
        """wma_wrapper.rs:747: (d_prices, d_periods, series_len, n_combos,
        first_valid, max_period, d_out)."""
        self._batch_device(*args)

    def wma_multi_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev(data_tm, cols, rows, period)

    def wma_multi_series_one_param_time_major_device(self, d_prices_tm, period: int,
                                                     num_series: int, series_len: int,
                                                     d_first_valids, d_out_tm) -> None:
        """wma_wrapper.rs:1049."""
        _need_ms(d_prices_tm, d_first_valids, d_out_tm, num_series, series_len)
        check(lib().vta_wma_multi_series_one_param_time_major_device(
            self.ctx.handle, _vp(d_prices_tm), int(period), int(num_series), int(series_len),
            _vp(d_first_valids), _vp(d_out_tm)))


class CudaSwma(_Fir):
    _name = "swma"

    def swma_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev(prices_f32, period)

    def swma_batch_device(self, d_prices, d_periods, d_warms, series_len: int, n_combos: int,
                          max_period: int, d_out) -> None:
        """swma_wrapper.rs:663: first_valid = d_warms[0] - periods[0] + 1 on the
        device; every warm index checked (reported by the next synchronize)."""
        require_len(d_prices, series_len, "prices")
        require_len(d_periods, n_combos, "periods")
        require_len(d_warms, n_combos, "warms")
        require_len(d_out, int(n_combos) * int(series_len), "output")
        check(lib().vta_swma_batch_device(self.ctx.handle, _vp(d_prices), _vp(d_periods),
                                          _vp(d_warms), int(series_len), int(n_combos),
                                          int(max_period), _vp(d_out)))

    def swma_multi_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev(data_tm, cols, rows, period)

    def swma_multi_series_one_param_device(self, d_prices_tm, d_weights, period: int,
                                           num_series: int, series_len: int, d_first_valids,
                                           d_out_tm) -> None:
        """swma_wrapper.rs:787: the weights follow from the period (d_weights is
        not read)."""
        _need_ms(d_prices_tm, d_first_valids, d_out_tm, num_series, series_len)
        check(lib().vta_swma_multi_series_one_param_device(
            self.ctx.handle, _vp(d_prices_tm), _vp(d_weights), int(period), int(num_series),
            int(series_len), _vp(d_first_valids), _vp(d_out_tm)))
