"""EMA / RSI / ATR / ADX operator surfaces: mirrors of CudaEma
(src/cuda/moving_averages/ema_wrapper.rs), CudaRsi (src/cuda/rsi_wrapper.rs),
CudaAtr (src/cuda/atr_wrapper.rs) and CudaAdx (src/cuda/adx_wrapper.rs), plus
the fused EMA + RSI + ATR many-series launch."""
from __future__ import annotations

import ctypes
from typing import List, Sequence, Tuple

import numpy as np

from ._lib import DeviceArrayF32C, RangeUsize, RecJobC, VtaError, check, lib
from .device import DeviceArrayF32, DeviceBuffer, HipContext, f32_ptr, require_len


def _rng(r: Sequence[int]) -> RangeUsize:
    return RangeUsize(*[int(v) for v in r])


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _expand(fn, period: Sequence[int]) -> List[int]:
    rc = _rng(period)
    n = ctypes.c_size_t()
    check(fn(ctypes.byref(rc), None, 0, ctypes.byref(n)))
    arr = (ctypes.c_size_t * max(n.value, 1))()
    check(fn(ctypes.byref(rc), arr, n.value, ctypes.byref(n)))
    return [int(v) for v in arr[:n.value]]


class _OneParam:
    """Shared body of the one-parameter wrappers."""
    _name = ""
    _hlc = False

    def __init__(self, device_id: int = 0):
        self.ctx = HipContext(device_id)

    def synchronize(self) -> None:
        self.ctx.synchronize()

    @classmethod
    def expand_grid(cls, period: Tuple[int, int, int]) -> List[int]:
        return _expand(getattr(lib(), f"vta_{cls._name}_expand_grid"), period)

    def _batch_dev(self, arrays, period) -> DeviceArrayF32:
        arrs = [_f32(a) for a in arrays]
        n = arrs[0].size
        rc = _rng(period)
        out = DeviceArrayF32C()
        fn = getattr(lib(), f"vta_{self._name}_batch_dev")
        check(fn(self.ctx.handle, *[f32_ptr(a) for a in arrs], n, ctypes.byref(rc),
                 ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)

    def _ms_dev(self, arrays, cols: int, rows: int, period: int) -> DeviceArrayF32:
        arrs = [_f32(a).reshape(-1) for a in arrays]
        out = DeviceArrayF32C()
        fn = getattr(lib(), f"vta_{self._name}_many_series_one_param_time_major_dev")
        check(fn(self.ctx.handle, *[f32_ptr(a) for a in arrs], int(cols), int(rows), int(period),
                 ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)


def _need_batch(d_prices, d_periods, d_out, series_len, n_combos) -> None:
    """One series x n_combos periods: series_len bars in, n_combos int32
    periods, an n_combos x series_len output (ema_wrapper.rs:282-296)."""
    require_len(d_prices, series_len, "prices")
    require_len(d_periods, n_combos, "periods")
    require_len(d_out, int(n_combos) * int(series_len), "output")


def _need_ms(d_prices_tm, d_first_valids, d_out_tm, d_clean, num_series, series_len) -> None:
    """Time-major many-series forms: series_len x num_series in and out, one
    int32 first_valid (and clean flag) per series (ema_wrapper.rs:400-409)."""
    total = int(num_series) * int(series_len)
    require_len(d_prices_tm, total, "time-major input")
    require_len(d_out_tm, total, "time-major output")
    require_len(d_first_valids, num_series, "first_valids")
    require_len(d_clean, num_series, "clean flags")


class CudaEma(_OneParam):
    _name = "ema"

    def ema_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev([prices_f32], period)

    def ema_batch_device(self, d_prices: DeviceBuffer, d_periods: DeviceBuffer, max_period: int,
                         series_len: int, first_valid: int, n_combos: int,
                         d_out: DeviceBuffer) -> None:
        _need_batch(d_prices, d_periods, d_out, series_len, n_combos)
        check(lib().vta_ema_batch_device(self.ctx.handle, ctypes.c_void_p(d_prices.ptr),
                                         ctypes.c_void_p(d_periods.ptr), None, int(max_period),
                                         int(series_len), int(first_valid), int(n_combos),
                                         ctypes.c_void_p(d_out.ptr)))

    def ema_many_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev([data_tm], cols, rows, period)

    def ema_many_series_one_param_device(self, d_prices_tm, d_first_valids, period, num_series,
                                         series_len, d_out_tm, d_clean=None) -> None:
        _need_ms(d_prices_tm, d_first_valids, d_out_tm, d_clean, num_series, series_len)
        check(lib().vta_ema_many_series_one_param_device(
            self.ctx.handle, ctypes.c_void_p(d_prices_tm.ptr), ctypes.c_void_p(d_first_valids.ptr),
            int(period), int(num_series), int(series_len),
            ctypes.c_void_p(d_clean.ptr) if d_clean is not None else None,
            ctypes.c_void_p(d_out_tm.ptr)))


class CudaRsi(_OneParam):
    _name = "rsi"

    def rsi_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev([prices_f32], period)

    def rsi_many_series_one_param_time_major_dev(self, prices_tm, cols, rows, period):
        return self._ms_dev([prices_tm], cols, rows, period)


class CudaAtr(_OneParam):
    _name = "atr"

    def atr_batch_dev(self, high, low, close, length: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev([high, low, close], length)

    def atr_many_series_one_param_time_major_dev(self, high_tm, low_tm, close_tm, cols, rows,
                                                 period):
        return self._ms_dev([high_tm, low_tm, close_tm], cols, rows, period)

    def tr_from_hlc_device(self, d_high: DeviceBuffer, d_low: DeviceBuffer,
                           d_close: DeviceBuffer, series_len: int, first_valid: int,
                           d_tr_out: DeviceBuffer) -> None:
        """atr_wrapper.rs:566: enqueue-only true range of one device-resident series."""
        for b, what in ((d_high, "high"), (d_low, "low"), (d_close, "close"),
                        (d_tr_out, "tr output")):
            require_len(b, series_len, what)
        check(lib().vta_tr_from_hlc_device(self.ctx.handle, ctypes.c_void_p(d_high.ptr),
                                           ctypes.c_void_p(d_low.ptr), ctypes.c_void_p(d_close.ptr),
                                           int(series_len), int(first_valid),
                                           ctypes.c_void_p(d_tr_out.ptr)))

    def atr_batch_device_with_tr(self, d_tr: DeviceBuffer, d_periods: DeviceBuffer,
                                 max_period: int, series_len: int, first_valid: int,
                                 n_combos: int, d_out: DeviceBuffer) -> None:
        """atr_wrapper.rs:626: enqueue-only ATR of every period over a TR series."""
        _need_batch(d_tr, d_periods, d_out, series_len, n_combos)
        check(lib().vta_atr_batch_device_with_tr(self.ctx.handle, ctypes.c_void_p(d_tr.ptr),
                                                 ctypes.c_void_p(d_periods.ptr), int(max_period),
                                                 int(series_len), int(first_valid),
                                                 int(n_combos), ctypes.c_void_p(d_out.ptr)))


class CudaAdx(_OneParam):
    _name = "adx"

    def adx_batch_dev(self, high, low, close, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev([high, low, close], period)

    def adx_many_series_one_param_time_major_dev(self, high_tm, low_tm, close_tm, cols, rows,
                                                 period):
        return self._ms_dev([high_tm, low_tm, close_tm], cols, rows, period)


class _OneInputMa(_OneParam):
    """SMMA / Wilders: one-series batch, enqueue-only batch and many-series forms."""

    def _batch(self, prices_f32, period) -> DeviceArrayF32:
        return self._batch_dev([prices_f32], period)

    def _batch_device(self, d_prices: DeviceBuffer, d_periods: DeviceBuffer, max_period: int,
                      series_len: int, first_valid: int, n_combos: int,
                      d_out: DeviceBuffer) -> None:
        _need_batch(d_prices, d_periods, d_out, series_len, n_combos)
        fn = getattr(lib(), f"vta_{self._name}_batch_device")
        check(fn(self.ctx.handle, ctypes.c_void_p(d_prices.ptr), ctypes.c_void_p(d_periods.ptr),
                 int(max_period), int(series_len), int(first_valid), int(n_combos),
                 ctypes.c_void_p(d_out.ptr)))

    def _ms_device(self, d_prices_tm, d_first_valids, period, num_series, series_len, d_out_tm,
                   d_clean=None) -> None:
        _need_ms(d_prices_tm, d_first_valids, d_out_tm, d_clean, num_series, series_len)
        fn = getattr(lib(), f"vta_{self._name}_many_series_one_param_device")
        check(fn(self.ctx.handle, ctypes.c_void_p(d_prices_tm.ptr),
                 ctypes.c_void_p(d_first_valids.ptr), int(period), int(num_series),
                 int(series_len), ctypes.c_void_p(d_clean.ptr) if d_clean is not None else None,
                 ctypes.c_void_p(d_out_tm.ptr)))


class CudaSmma(_OneInputMa):
    """src/cuda/moving_averages/smma_wrapper.rs (:398 batch_dev, :377 batch_device,
    :623 many-series _dev, :599 many-series _device)."""
    _name = "smma"

    def smma_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch(prices_f32, period)

    def smma_batch_device(self, *args) -> None:
        self._batch_device(*args)

    def smma_multi_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev([data_tm], cols, rows, period)

    def smma_multi_series_one_param_device(self, *args, **kw) -> None:
        self._ms_device(*args, **kw)


class CudaWilders(_OneInputMa):
    """src/cuda/moving_averages/wilders_wrapper.rs (:134 batch_dev, :252 batch_device,
    :600 many-series _dev, :662 many-series _device)."""
    _name = "wilders"

    def wilders_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch(prices_f32, period)

    def wilders_batch_device(self, *args) -> None:
        self._batch_device(*args)

    def wilders_many_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev([data_tm], cols, rows, period)

    def wilders_many_series_one_param_device(self, *args, **kw) -> None:
        self._ms_device(*args, **kw)


def ema_rsi_atr_many_series_dev(ctx: HipContext, close_tm, high_tm, low_tm, cols: int,
                                rows: int, ema_period: int = 20, rsi_period: int = 14,
                                atr_period: int = 14):
    """Fused EMA + RSI + ATR over time-major series; returns (ema, rsi, atr) device arrays."""
    c, h, lo = (_f32(a).reshape(-1) for a in (close_tm, high_tm, low_tm))
    outs = (DeviceArrayF32C * 3)()
    check(lib().vta_ema_rsi_atr_many_series_dev(ctx.handle, f32_ptr(c), f32_ptr(h), f32_ptr(lo),
                                                int(cols), int(rows), int(ema_period),
                                                int(rsi_period), int(atr_period), outs))
    return tuple(DeviceArrayF32(ctx, outs[i]) for i in range(3))


class ManySeriesOhlcPlan:
    """Device-resident time-major close/high/low with prepared per-series
    metadata, for repeated enqueue-only fused launches."""

    def __init__(self, ctx: HipContext, close_tm, high_tm, low_tm):
        c, h, lo = (_f32(a) for a in (close_tm, high_tm, low_tm))
        self.rows, self.cols = c.shape
        self.ctx = ctx
        self.d_c, self.d_h, self.d_l = ctx.to_device(c), ctx.to_device(h), ctx.to_device(lo)
        n = self.cols
        self.d_fv, self.d_fv3, self.d_clean = ctx.alloc(4 * n), ctx.alloc(4 * n), ctx.alloc(4 * n)
        check(lib().vta_many_series_prepare(
            ctx.handle, ctypes.c_void_p(self.d_c.ptr), ctypes.c_void_p(self.d_h.ptr),
            ctypes.c_void_p(self.d_l.ptr), n, self.rows, 3, ctypes.c_void_p(self.d_fv.ptr),
            ctypes.c_void_p(self.d_fv3.ptr), ctypes.c_void_p(self.d_clean.ptr)))
        ctx.synchronize()
        nbytes = 4 * self.rows * self.cols
        self.d_out = [ctx.alloc(nbytes) for _ in range(3)]

    def launch(self, ema_period=20, rsi_period=14, atr_period=14) -> None:
        check(lib().vta_ema_rsi_atr_many_series_device(
            self.ctx.handle, ctypes.c_void_p(self.d_c.ptr), ctypes.c_void_p(self.d_h.ptr),
            ctypes.c_void_p(self.d_l.ptr), self.cols, self.rows, int(ema_period), int(rsi_period),
            int(atr_period), ctypes.c_void_p(self.d_fv.ptr), ctypes.c_void_p(self.d_fv3.ptr),
            ctypes.c_void_p(self.d_clean.ptr), *[ctypes.c_void_p(b.ptr) for b in self.d_out]))

    def outputs(self):
        return [b.copy_to_host().reshape(self.rows, self.cols) for b in self.d_out]


REC_KINDS = {"ema": 0, "rsi": 1, "atr": 2, "adx": 3, "smma": 4, "wilders": 5}


class RecPack:
    """Many one-series batches of one recurrence indicator run together: one
    launch per scan pass over a device job table (vta_rec_pack_*).  No
    reference counterpart (its wrappers run one series per call).

    jobs: dicts with device buffers ``close`` (and ``high``, ``low`` for ATR /
    ADX), ``periods`` (int32), ``out`` (n_combos x series_len f32) and the ints
    ``series_len``, ``first_valid``, ``n_combos``, ``max_period``.  The
    buffers must outlive the pack; ``launch()`` is enqueue-only."""

    def __init__(self, ctx: HipContext, kind: str, jobs: Sequence[dict]):
        if kind not in REC_KINDS:
            raise VtaError(4, f"unknown indicator kind {kind!r}")
        self.ctx = ctx
        self.kind = kind
        arr = (RecJobC * len(jobs))()
        for i, j in enumerate(jobs):
            T, C = int(j["series_len"]), int(j["n_combos"])
            require_len(j["close"], T, "close")
            require_len(j["periods"], C, "periods")
            require_len(j["out"], C * T, "output")
            require_len(j.get("high"), T, "high")
            require_len(j.get("low"), T, "low")
            arr[i] = RecJobC(j["close"].ptr, j["high"].ptr if j.get("high") else None,
                             j["low"].ptr if j.get("low") else None, j["periods"].ptr,
                             j["out"].ptr, int(j["series_len"]), int(j["first_valid"]),
                             int(j["n_combos"]), int(j["max_period"]))
        self._p = ctypes.c_void_p()
        check(lib().vta_rec_pack_create(ctx.handle, REC_KINDS[kind], arr, len(jobs),
                                        ctypes.byref(self._p)))

    def launch(self) -> None:
        if not self._p:
            raise VtaError(4, "pack is destroyed")
        check(lib().vta_rec_pack_launch(self.ctx.handle, self._p))

    def close(self) -> None:
        if self._p:
            lib().vta_rec_pack_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
