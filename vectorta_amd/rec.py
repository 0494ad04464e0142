"""EMA / RSI / ATR / ADX operator surfaces: mirrors of CudaEma
(src/cuda/moving_averages/ema_wrapper.rs), CudaRsi (src/cuda/rsi_wrapper.rs),
CudaAtr (src/cuda/atr_wrapper.rs) and CudaAdx (src/cuda/adx_wrapper.rs), plus
the fused EMA + RSI + ATR many-series launch."""
from __future__ import annotations

import ctypes
from typing import List, Sequence, Tuple

import numpy as np

from dataclasses import dataclass
from typing import Optional

from ._lib import (DeviceArrayF32C, KernelPolicyC, KernelSelectedC, RangeUsize, RecJobC,
                   VtaError, check, lib)
from .device import DeviceArrayF32, DeviceBuffer, HipContext, f32_ptr, require_len

_KIND = {"auto": 0, "plain": 1, "oned": 1}
_KIND_NAME = {0: "auto", 1: "plain"}
_SELECTED = {1: "plain", 2: "chunked"}


@dataclass
class KernelPolicy:
    """CudaEmaPolicy / CudaRsiPolicy / CudaAdxPolicy / CudaPnrPolicy
    (ema_wrapper.rs:49-90): per call family "auto" or "plain" (the reference's
    Plain { block_x } / OneD { block_x }: one lane per combo or series over
    the whole series, no time chunking).  block_x is 0 or 64 (64-lane waves)."""
    batch: str = "auto"
    batch_block_x: int = 0
    many_series: str = "auto"
    many_series_block_x: int = 0

    def to_c(self) -> KernelPolicyC:
        for k in (self.batch, self.many_series):
            if k not in _KIND:
                raise VtaError(5, f"invalid policy: {k}")
        return KernelPolicyC(_KIND[self.batch], int(self.batch_block_x),
                             _KIND[self.many_series], int(self.many_series_block_x))


class _PolicyMixin:
    """set_policy / policy / selected_batch_kernel / selected_many_series_kernel
    (percentile_nearest_rank_wrapper.rs:143-152); new_with_policy is the
    constructor's policy argument (ema_wrapper.rs:172)."""
    _policy_name = ""

    def set_policy(self, policy: KernelPolicy) -> None:
        pc = policy.to_c()
        check(getattr(lib(), f"vta_{self._policy_name}_set_policy")(self.ctx.handle,
                                                                    ctypes.byref(pc)))

    def policy(self) -> KernelPolicy:
        pc = KernelPolicyC()
        check(getattr(lib(), f"vta_{self._policy_name}_policy")(self.ctx.handle,
                                                                ctypes.byref(pc)))
        return KernelPolicy(_KIND_NAME[pc.batch], pc.batch_block_x,
                            _KIND_NAME[pc.many_series], pc.many_series_block_x)

    def _selected(self, which: str):
        sc = KernelSelectedC()
        check(getattr(lib(), f"vta_{self._policy_name}_selected_{which}_kernel")(
            self.ctx.handle, ctypes.byref(sc)))
        return (_SELECTED[sc.kind], sc.block_x) if sc.kind in _SELECTED else None

    def selected_batch_kernel(self):
        """None before the first batch call, else ("plain" | "chunked", 64)."""
        return self._selected("batch")

    def selected_many_series_kernel(self):
        return self._selected("many_series")


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


def _vp(buf):
    return ctypes.c_void_p(buf.ptr) if buf is not None else None


def _into_host(fn, out: np.ndarray, *args) -> None:
    if out.dtype != np.float32 or not out.flags.c_contiguous:
        raise VtaError(4, "out must be a contiguous float32 array")
    check(fn(*args, f32_ptr(out), out.size))


class CudaEma(_OneParam, _PolicyMixin):
    _name = "ema"
    _policy_name = "ema"

    def __init__(self, device_id: int = 0, policy: Optional[KernelPolicy] = None):
        super().__init__(device_id)
        if policy is not None:
            self.set_policy(policy)

    def ema_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev([prices_f32], period)

    def ema_batch_device(self, d_prices: DeviceBuffer, d_periods: DeviceBuffer,
                         d_alphas: Optional[DeviceBuffer], series_len: int, first_valid: int,
                         n_combos: int, d_out: DeviceBuffer) -> None:
        """ema_wrapper.rs:256: enqueue-only; d_alphas[c] is combo c's smoothing
        factor (None derives 2/(p+1) in f64)."""
        _need_batch(d_prices, d_periods, d_out, series_len, n_combos)
        require_len(d_alphas, n_combos, "alphas")
        check(lib().vta_ema_batch_device(self.ctx.handle, _vp(d_prices), _vp(d_periods),
                                         _vp(d_alphas), int(series_len), int(first_valid),
                                         int(n_combos), _vp(d_out)))

    def ema_batch_into_host_f32(self, prices_f32, period, out: np.ndarray) -> None:
        x = _f32(prices_f32)
        rc = _rng(period)
        _into_host(lib().vta_ema_batch_into_host_f32, out, self.ctx.handle, f32_ptr(x), x.size,
                   ctypes.byref(rc))

    def ema_many_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev([data_tm], cols, rows, period)

    def ema_many_series_one_param_device(self, d_prices_tm, d_first_valids, period: int,
                                         alpha: float, num_series: int, series_len: int,
                                         d_out_tm) -> None:
        """ema_wrapper.rs:381: enqueue-only, the caller's smoothing factor."""
        _need_ms(d_prices_tm, d_first_valids, d_out_tm, None, num_series, series_len)
        check(lib().vta_ema_many_series_one_param_device(
            self.ctx.handle, _vp(d_prices_tm), _vp(d_first_valids), int(period), float(alpha),
            int(num_series), int(series_len), _vp(d_out_tm)))

    def ema_many_series_one_param_time_major_into_host_f32(self, data_tm, num_series,
                                                           series_len, period,
                                                           out_tm: np.ndarray) -> None:
        x = _f32(data_tm).reshape(-1)
        _into_host(lib().vta_ema_many_series_one_param_time_major_into_host_f32, out_tm,
                   self.ctx.handle, f32_ptr(x), int(num_series), int(series_len), int(period))


class CudaRsi(_OneParam, _PolicyMixin):
    _name = "rsi"
    _policy_name = "rsi"

    def __init__(self, device_id: int = 0, policy: Optional[KernelPolicy] = None):
        super().__init__(device_id)
        if policy is not None:
            self.set_policy(policy)

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


class CudaAdx(_OneParam, _PolicyMixin):
    _name = "adx"
    _policy_name = "adx"

    def __init__(self, device_id: int = 0, policy: Optional[KernelPolicy] = None):
        super().__init__(device_id)
        if policy is not None:
            self.set_policy(policy)

    def adx_batch_dev(self, high, low, close, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev([high, low, close], period)

    def adx_batch_into_host_f32(self, high, low, close, period, out: np.ndarray):
        """adx_wrapper.rs:253: returns (rows, cols, periods) like the reference."""
        h, lo, c = (_f32(a) for a in (high, low, close))
        rc = _rng(period)
        _into_host(lib().vta_adx_batch_into_host_f32, out, self.ctx.handle, f32_ptr(h),
                   f32_ptr(lo), f32_ptr(c), c.size, ctypes.byref(rc))
        ps = self.expand_grid(period)
        return len(ps), c.size, ps

    def adx_many_series_one_param_time_major_dev(self, high_tm, low_tm, close_tm, cols, rows,
                                                 period):
        return self._ms_dev([high_tm, low_tm, close_tm], cols, rows, period)

    def adx_many_series_one_param_time_major_into_host_f32(self, high_tm, low_tm, close_tm,
                                                           cols, rows, period,
                                                           out_tm: np.ndarray) -> None:
        h, lo, c = (_f32(a).reshape(-1) for a in (high_tm, low_tm, close_tm))
        _into_host(lib().vta_adx_many_series_one_param_time_major_into_host_f32, out_tm,
                   self.ctx.handle, f32_ptr(h), f32_ptr(lo), f32_ptr(c), int(cols), int(rows),
                   int(period))


class CudaSmma(_OneParam):
    """src/cuda/moving_averages/smma_wrapper.rs (:398 batch_dev, :377 batch_device,
    :623 many-series _dev, :599 many-series _device)."""
    _name = "smma"

    def smma_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev([prices_f32], period)

    def smma_batch_device(self, d_prices, d_periods, d_warms, first_valid: int, series_len: int,
                          n_combos: int, d_out) -> None:
        """Enqueue-only; d_warms[c] must be first_valid + period - 1 (checked on
        the device, reported by the next synchronize)."""
        _need_batch(d_prices, d_periods, d_out, series_len, n_combos)
        require_len(d_warms, n_combos, "warms")
        check(lib().vta_smma_batch_device(self.ctx.handle, _vp(d_prices), _vp(d_periods),
                                          _vp(d_warms), int(first_valid), int(series_len),
                                          int(n_combos), _vp(d_out)))

    def smma_multi_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev([data_tm], cols, rows, period)

    def smma_multi_series_one_param_device(self, d_prices_tm, period: int, num_series: int,
                                           series_len: int, d_first_valids, d_out_tm) -> None:
        _need_ms(d_prices_tm, d_first_valids, d_out_tm, None, num_series, series_len)
        check(lib().vta_smma_multi_series_one_param_device(
            self.ctx.handle, _vp(d_prices_tm), int(period), int(num_series), int(series_len),
            _vp(d_first_valids), _vp(d_out_tm)))


class CudaWilders(_OneParam):
    """src/cuda/moving_averages/wilders_wrapper.rs (:134 batch_dev, :252 batch_device,
    :600 many-series _dev, :662 many-series _device)."""
    _name = "wilders"

    def wilders_batch_dev(self, prices_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        return self._batch_dev([prices_f32], period)

    def wilders_batch_device(self, d_prices, d_periods, d_alphas, d_warm, series_len: int,
                             first_valid: int, n_combos: int, d_out) -> None:
        """Enqueue-only; the alphas are used, d_warm[c] must be
        first_valid + period - 1 (checked on the device)."""
        _need_batch(d_prices, d_periods, d_out, series_len, n_combos)
        require_len(d_alphas, n_combos, "alphas")
        require_len(d_warm, n_combos, "warms")
        check(lib().vta_wilders_batch_device(self.ctx.handle, _vp(d_prices), _vp(d_periods),
                                             _vp(d_alphas), _vp(d_warm), int(series_len),
                                             int(first_valid), int(n_combos), _vp(d_out)))

    def wilders_many_series_one_param_time_major_dev(self, data_tm, cols, rows, period):
        return self._ms_dev([data_tm], cols, rows, period)

    def wilders_many_series_one_param_device(self, d_prices_tm, d_first_valids, period: int,
                                             alpha: float, num_series: int, series_len: int,
                                             d_out_tm) -> None:
        _need_ms(d_prices_tm, d_first_valids, d_out_tm, None, num_series, series_len)
        check(lib().vta_wilders_many_series_one_param_device(
            self.ctx.handle, _vp(d_prices_tm), _vp(d_first_valids), int(period), float(alpha),
            int(num_series), int(series_len), _vp(d_out_tm)))


MA_KINDS = {"ema": 0, "smma": 1, "wilders": 2, "sma": 3, "wma": 4, "swma": 5}


def ma_many_series_prepared_device(ctx: HipContext, kind: str, d_prices_tm, d_first_valids,
                                   d_meta, period: int, num_series: int, series_len: int,
                                   d_out_tm, alpha: float = 0.0) -> None:
    """Enqueue-only many-series EMA / SMMA / Wilders / SMA / WMA / SWMA over
    metadata prepared once with vta_many_series_prepare (clean flags, mode 0,
    for EMA / SMMA / Wilders; first non-finite bars, mode 4, for the others)."""
    if kind not in MA_KINDS:
        raise VtaError(4, f"unknown moving-average kind {kind!r}")
    total = int(num_series) * int(series_len)
    require_len(d_prices_tm, total, "time-major input")
    require_len(d_out_tm, total, "time-major output")
    require_len(d_first_valids, num_series, "first_valids")
    require_len(d_meta, num_series, "metadata")
    check(lib().vta_ma_many_series_prepared_device(
        ctx.handle, MA_KINDS[kind], _vp(d_prices_tm), _vp(d_first_valids), _vp(d_meta),
        int(period), float(alpha), int(num_series), int(series_len), _vp(d_out_tm)))


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
