"""Rolling-window operator surfaces: mirrors of CudaBollingerBands
(src/cuda/bollinger_bands_wrapper.rs), CudaPercentileNearestRank
(src/cuda/percentile_nearest_rank_wrapper.rs) and CudaMediumAd
(src/cuda/medium_ad_wrapper.rs)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import (BbBatchRangeC, BbJobC, BbParamsC, DeviceArrayF32C, PnrBatchRangeC, PnrParamsC,
                   RangeF64, RangeUsize, check, lib)
from .device import DeviceArrayF32, DeviceBuffer, HipContext, f32_ptr, require_len
from .rec import KernelPolicy, _PolicyMixin


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _ru(r: Sequence[int]) -> RangeUsize:
    return RangeUsize(*[int(v) for v in r])


def _rf(r: Sequence[float]) -> RangeF64:
    return RangeF64(*[float(v) for v in r])


@dataclass
class BollingerBandsBatchRange:
    """period (usize axis), devup and devdn (f64 axes); matype "sma" and
    devtype 0 are the only values the GPU path accepts."""
    period: Tuple[int, int, int] = (20, 20, 0)
    devup: Tuple[float, float, float] = (2.0, 2.0, 0.0)
    devdn: Tuple[float, float, float] = (2.0, 2.0, 0.0)

    def to_c(self) -> BbBatchRangeC:
        return BbBatchRangeC(_ru(self.period), _rf(self.devup), _rf(self.devdn))


@dataclass
class PercentileNearestRankBatchRange:
    length: Tuple[int, int, int] = (15, 15, 0)
    percentage: Tuple[float, float, float] = (50.0, 50.0, 0.0)

    def to_c(self) -> PnrBatchRangeC:
        return PnrBatchRangeC(_ru(self.length), _rf(self.percentage))


class _Rolling:
    def __init__(self, device_id: int = 0):
        self.ctx = HipContext(device_id)

    def synchronize(self) -> None:
        self.ctx.synchronize()


class CudaBollingerBands(_Rolling):
    @staticmethod
    def expand_grid(sweep: BollingerBandsBatchRange) -> List[Tuple[int, float, float]]:
        rc = sweep.to_c()
        n = ctypes.c_size_t()
        check(lib().vta_bb_expand_grid(ctypes.byref(rc), None, 0, ctypes.byref(n)))
        arr = (BbParamsC * max(n.value, 1))()
        check(lib().vta_bb_expand_grid(ctypes.byref(rc), arr, n.value, ctypes.byref(n)))
        return [(int(p.period), float(p.devup), float(p.devdn)) for p in arr[:n.value]]

    def bollinger_bands_batch_dev(self, data_f32, sweep: BollingerBandsBatchRange):
        """Returns (upper, middle, lower) device arrays, each [combo][len]."""
        x = _f32(data_f32)
        rc = sweep.to_c()
        outs = (DeviceArrayF32C * 3)()
        check(lib().vta_bollinger_bands_batch_dev(self.ctx.handle, f32_ptr(x), x.size,
                                                  ctypes.byref(rc), outs))
        return tuple(DeviceArrayF32(self.ctx, outs[i]) for i in range(3))

    def bollinger_bands_many_series_one_param_time_major_dev(self, data_tm_f32, cols: int,
                                                             rows: int, period: int,
                                                             devup: float, devdn: float):
        x = _f32(data_tm_f32).reshape(-1)
        outs = (DeviceArrayF32C * 3)()
        check(lib().vta_bollinger_bands_many_series_one_param_time_major_dev(
            self.ctx.handle, f32_ptr(x), int(cols), int(rows), int(period), float(devup),
            float(devdn), outs))
        return tuple(DeviceArrayF32(self.ctx, outs[i]) for i in range(3))

    def bollinger_bands_many_series_one_param_device(
            self, d_data_tm: DeviceBuffer, d_first_valids: DeviceBuffer,
            d_first_nonfinite: Optional[DeviceBuffer], cols: int, rows: int, period: int,
            devup: float, devdn: float, d_upper: DeviceBuffer, d_middle: DeviceBuffer,
            d_lower: DeviceBuffer) -> None:
        total = int(cols) * int(rows)
        for b, what in ((d_data_tm, "time-major input"), (d_upper, "upper"),
                        (d_middle, "middle"), (d_lower, "lower")):
            require_len(b, total, what)
        require_len(d_first_valids, cols, "first_valids")
        require_len(d_first_nonfinite, cols, "first_nonfinite")
        check(lib().vta_bollinger_bands_many_series_one_param_device(
            self.ctx.handle, ctypes.c_void_p(d_data_tm.ptr), ctypes.c_void_p(d_first_valids.ptr),
            ctypes.c_void_p(d_first_nonfinite.ptr) if d_first_nonfinite is not None else None,
            int(cols), int(rows), int(period), float(devup), float(devdn),
            ctypes.c_void_p(d_upper.ptr), ctypes.c_void_p(d_middle.ptr),
            ctypes.c_void_p(d_lower.ptr)))


class CudaPercentileNearestRank(_Rolling, _PolicyMixin):
    """CudaPnr (percentile_nearest_rank_wrapper.rs), with its policy surface
    (set_policy / policy / selected_*_kernel, :143-152)."""
    _policy_name = "pnr"

    def __init__(self, device_id: int = 0, policy: Optional[KernelPolicy] = None):
        super().__init__(device_id)
        if policy is not None:
            self.set_policy(policy)

    @staticmethod
    def expand_grid(sweep: PercentileNearestRankBatchRange) -> List[Tuple[int, float]]:
        rc = sweep.to_c()
        n = ctypes.c_size_t()
        check(lib().vta_pnr_expand_grid(ctypes.byref(rc), None, 0, ctypes.byref(n)))
        arr = (PnrParamsC * max(n.value, 1))()
        check(lib().vta_pnr_expand_grid(ctypes.byref(rc), arr, n.value, ctypes.byref(n)))
        return [(int(p.length), float(p.percentage)) for p in arr[:n.value]]

    def pnr_batch_dev(self, data_f32, sweep: PercentileNearestRankBatchRange):
        """Returns (device array [combo][len], combos) as the reference does."""
        x = _f32(data_f32)
        rc = sweep.to_c()
        out = DeviceArrayF32C()
        check(lib().vta_pnr_batch_dev(self.ctx.handle, f32_ptr(x), x.size, ctypes.byref(rc),
                                      ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out), self.expand_grid(sweep)

    def pnr_many_series_one_param_time_major_dev(self, data_tm_f32, cols: int, rows: int,
                                                 length: int, percentage: float):
        x = _f32(data_tm_f32).reshape(-1)
        out = DeviceArrayF32C()
        check(lib().vta_pnr_many_series_one_param_time_major_dev(
            self.ctx.handle, f32_ptr(x), int(cols), int(rows), int(length), float(percentage),
            ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)


class CudaMediumAd(_Rolling):
    @staticmethod
    def expand_grid(period: Tuple[int, int, int]) -> List[int]:
        rc = _ru(period)
        n = ctypes.c_size_t()
        check(lib().vta_medium_ad_expand_grid(ctypes.byref(rc), None, 0, ctypes.byref(n)))
        arr = (ctypes.c_size_t * max(n.value, 1))()
        check(lib().vta_medium_ad_expand_grid(ctypes.byref(rc), arr, n.value, ctypes.byref(n)))
        return [int(v) for v in arr[:n.value]]

    def medium_ad_batch_dev(self, data_f32, period: Tuple[int, int, int]) -> DeviceArrayF32:
        x = _f32(data_f32)
        rc = _ru(period)
        out = DeviceArrayF32C()
        check(lib().vta_medium_ad_batch_dev(self.ctx.handle, f32_ptr(x), x.size, ctypes.byref(rc),
                                            ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)

    def medium_ad_many_series_one_param_time_major_dev(self, data_tm_f32, cols: int, rows: int,
                                                       period: int) -> DeviceArrayF32:
        x = _f32(data_tm_f32).reshape(-1)
        out = DeviceArrayF32C()
        check(lib().vta_medium_ad_many_series_one_param_time_major_dev(
            self.ctx.handle, f32_ptr(x), int(cols), int(rows), int(period), ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)


class RollingManySeriesPlan:
    """Device-resident time-major [rows, cols] series with per-series
    first_valid / first non-finite bar tables, for repeated enqueue-only
    launches of Bollinger Bands, percentile_nearest_rank and rolling MAD."""

    def __init__(self, ctx: HipContext, data_tm):
        x = _f32(data_tm)
        self.rows, self.cols = x.shape
        self.ctx = ctx
        fin = ~np.isnan(x)
        if not fin.any(axis=0).all():
            raise ValueError("a series is all NaN")
        fv = np.argmax(fin, axis=0).astype(np.int32)
        nf = ~np.isfinite(x) & (np.arange(self.rows)[:, None] >= fv[None, :])
        z = np.where(nf.any(axis=0), np.argmax(nf, axis=0), self.rows).astype(np.int32)
        self.first_valid = fv
        self.d_x = ctx.to_device(x)
        self.d_fv = ctx.to_device(fv)
        self.d_z = ctx.to_device(z)
        nbytes = 4 * self.rows * self.cols
        self.d_out = [ctx.alloc(nbytes) for _ in range(3)]
        self._bb = CudaBollingerBands.__new__(CudaBollingerBands)
        self._bb.ctx = ctx

    def _need(self, n: int) -> None:
        if int(self.rows - self.first_valid.min()) < n:
            raise ValueError("not enough valid data for a series")

    def launch_bb(self, period: int = 20, devup: float = 2.0, devdn: float = 2.0) -> None:
        self._need(period)
        self._bb.bollinger_bands_many_series_one_param_device(
            self.d_x, self.d_fv, self.d_z, self.cols, self.rows, period, devup, devdn,
            *self.d_out)

    def launch_pnr(self, length: int = 100, percentage: float = 70.0) -> None:
        self._need(length)
        check(lib().vta_pnr_many_series_one_param_device(
            self.ctx.handle, ctypes.c_void_p(self.d_x.ptr), ctypes.c_void_p(self.d_fv.ptr),
            self.cols, self.rows, int(length), float(percentage),
            ctypes.c_void_p(self.d_out[0].ptr)))

    def launch_mad(self, period: int = 64) -> None:
        self._need(period)
        check(lib().vta_medium_ad_many_series_one_param_device(
            self.ctx.handle, ctypes.c_void_p(self.d_x.ptr), ctypes.c_void_p(self.d_fv.ptr),
            self.cols, self.rows, int(period), ctypes.c_void_p(self.d_out[0].ptr)))

    def outputs(self, n: int = 1):
        return [b.copy_to_host().reshape(self.rows, self.cols) for b in self.d_out[:n]]


class BbPack:
    """Many one-series Bollinger Bands batches in one launch
    (vta_bb_pack_*; no reference counterpart).  jobs: dicts with device
    buffers ``prices``, ``periods`` (int32), ``devup``, ``devdn`` (f64),
    ``upper``, ``middle``, ``lower`` (n_combos x series_len f32 each) and the
    ints ``series_len``, ``first_valid``, ``first_nonfinite``, ``n_combos``,
    ``max_period``.  The buffers must outlive the pack."""

    def __init__(self, ctx: HipContext, jobs: Sequence[dict]):
        self.ctx = ctx
        arr = (BbJobC * len(jobs))()
        for i, j in enumerate(jobs):
            T, C = int(j["series_len"]), int(j["n_combos"])
            require_len(j["prices"], T, "prices")
            require_len(j["periods"], C, "periods")
            require_len(j["devup"], C, "devup", itemsize=8)
            require_len(j["devdn"], C, "devdn", itemsize=8)
            for band in ("upper", "middle", "lower"):
                require_len(j[band], C * T, band)
            arr[i] = BbJobC(j["prices"].ptr, j["periods"].ptr, j["devup"].ptr, j["devdn"].ptr,
                            j["upper"].ptr, j["middle"].ptr, j["lower"].ptr,
                            int(j["series_len"]), int(j["first_valid"]),
                            int(j["first_nonfinite"]), int(j["n_combos"]), int(j["max_period"]))
        self._p = ctypes.c_void_p()
        check(lib().vta_bb_pack_create(ctx.handle, arr, len(jobs), ctypes.byref(self._p)))

    def launch(self) -> None:
        check(lib().vta_bb_pack_launch(self.ctx.handle, self._p))

    def close(self) -> None:
        if self._p:
            lib().vta_bb_pack_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
