"""ALMA operator surface: mirror of CudaAlma (src/cuda/moving_averages/alma_wrapper.rs).

Method names and argument meaning follow the reference wrapper so its tests
(tests/alma_cuda.rs, tests/alma_cuda_variants.rs) read the same here.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from ._lib import (AlmaBatchRangeC, AlmaJobC, AlmaParamsC, DeviceArrayF32C, RangeF64,
                   RangeUsize, VtaError, check, lib)
from .device import DeviceArrayF32, DeviceBuffer, HipContext, f32_ptr, require_len


@dataclass
class AlmaParams:
    """AlmaParams (alma.rs:60-80); defaults period 9, offset 0.85, sigma 6."""
    period: int = 9
    offset: float = 0.85
    sigma: float = 6.0


@dataclass
class AlmaBatchRange:
    """AlmaBatchRange (alma.rs:1067-1082): (start, end, step) per axis."""
    period: Tuple[int, int, int] = (9, 258, 1)
    offset: Tuple[float, float, float] = (0.85, 0.85, 0.0)
    sigma: Tuple[float, float, float] = (6.0, 6.0, 0.0)

    def to_c(self) -> AlmaBatchRangeC:
        return AlmaBatchRangeC(RangeUsize(*[int(v) for v in self.period]),
                               RangeF64(*[float(v) for v in self.offset]),
                               RangeF64(*[float(v) for v in self.sigma]))


_POLICY = {"auto": 0, "fp32": 1, "plain": 1, "mfma": 2, "mfma_bf16x3": 2, "mfma_f16x3": 3,
           "f16x3": 3}
_SELECTED = {0: None, 1: "fp32", 2: "mfma_bf16x3", 3: "mfma_f16x3"}


def expand_grid(sweep: AlmaBatchRange) -> List[AlmaParams]:
    """expand_grid (alma_wrapper.rs:1550-1631), computed by the C++ runtime."""
    rc = sweep.to_c()
    n = ctypes.c_size_t()
    check(lib().vta_alma_expand_grid(ctypes.byref(rc), None, 0, ctypes.byref(n)))
    arr = (AlmaParamsC * max(n.value, 1))()
    check(lib().vta_alma_expand_grid(ctypes.byref(rc), arr, n.value, ctypes.byref(n)))
    return [AlmaParams(int(a.period), float(a.offset), float(a.sigma)) for a in arr[:n.value]]


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class AlmaSweepPlan:
    """Device-resident per-combo constants for repeated sweeps (enqueue-only launch)."""

    def __init__(self, alma: "CudaAlma", sweep: AlmaBatchRange, series_len: int,
                 kernel: str = "auto"):
        self.alma = alma
        self._p = ctypes.c_void_p()
        rc = sweep.to_c()
        check(lib().vta_alma_sweep_plan_create(alma.ctx.handle, ctypes.byref(rc),
                                               int(series_len), _POLICY[kernel],
                                               ctypes.byref(self._p)))
        self.series_len = int(series_len)
        self.n_combos = int(lib().vta_alma_sweep_plan_combos(self._p))

    def launch(self, d_prices: int, first_valid: int, d_out: int) -> None:
        check(lib().vta_alma_sweep_launch(self.alma.ctx.handle, self._p,
                                          ctypes.c_void_p(d_prices), int(first_valid),
                                          ctypes.c_void_p(d_out)))

    def close(self) -> None:
        if self._p:
            lib().vta_alma_sweep_plan_destroy(self.alma.ctx.handle, self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            if self.alma.ctx._ptr:
                self.close()
        except Exception:
            pass


class AlmaPack:
    """Many one-series ALMA sweeps (uneven lengths, own grids) in one launch
    (vta_alma_pack_*; no reference counterpart).  jobs: dicts with device
    buffers ``prices`` (series_len f32) and ``out`` (n_combos x series_len
    f32), an AlmaBatchRange ``sweep`` and the ints ``series_len`` and
    ``first_valid``.  kernel: "auto", "fp32", "mfma" (split-BF16) or
    "mfma_f16x3".  The buffers must outlive the pack."""

    def __init__(self, ctx: HipContext, jobs, kernel: str = "auto"):
        self.ctx = ctx
        arr = (AlmaJobC * max(len(jobs), 1))()
        self.combos = []
        for i, j in enumerate(jobs):
            T = int(j["series_len"])
            C = len(expand_grid(j["sweep"]))
            require_len(j["prices"], T, "prices")
            require_len(j["out"], C * T, "output")
            arr[i] = AlmaJobC(j["prices"].ptr, j["out"].ptr, j["sweep"].to_c(), T,
                              int(j["first_valid"]))
        self._p = ctypes.c_void_p()
        check(lib().vta_alma_pack_create(ctx.handle, _POLICY[kernel], arr, len(jobs),
                                         ctypes.byref(self._p)))
        self.kernel = _SELECTED[int(lib().vta_alma_pack_kernel(self._p))]
        self.combos = [int(lib().vta_alma_pack_job_combos(self._p, i)) for i in range(len(jobs))]

    def launch(self) -> None:
        check(lib().vta_alma_pack_launch(self.ctx.handle, self._p))

    def close(self) -> None:
        if self._p:
            lib().vta_alma_pack_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CudaAlma:
    """ALMA on one MI355X: new / set_policy / *_batch_dev / *_device / many-series."""

    def __init__(self, device_id: int = 0, policy: str = "auto"):
        self.ctx = HipContext(device_id)
        self.set_policy(policy)

    # CudaAlma::set_policy / selected_batch_kernel (alma_wrapper.rs:224-245)
    def set_policy(self, batch: str) -> None:
        if batch not in _POLICY:
            raise VtaError(5, f"invalid policy: {batch}")
        check(lib().vta_alma_set_batch_policy(self.ctx.handle, _POLICY[batch]))

    def selected_batch_kernel(self):
        return _SELECTED.get(int(lib().vta_alma_selected_batch_kernel(self.ctx.handle)))

    def synchronize(self) -> None:
        self.ctx.synchronize()

    def alma_batch_dev(self, data_f32, sweep: AlmaBatchRange) -> DeviceArrayF32:
        x = _f32(data_f32)
        rc = sweep.to_c()
        out = DeviceArrayF32C()
        check(lib().vta_alma_batch_dev(self.ctx.handle, f32_ptr(x), x.size, ctypes.byref(rc),
                                       ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)

    def alma_batch_from_device_prices(self, d_prices: DeviceBuffer, series_len: int,
                                      first_valid: int, sweep: AlmaBatchRange) -> DeviceArrayF32:
        require_len(d_prices, series_len, "prices")
        rc = sweep.to_c()
        out = DeviceArrayF32C()
        check(lib().vta_alma_batch_from_device_prices(
            self.ctx.handle, ctypes.c_void_p(d_prices.ptr), int(series_len), int(first_valid),
            ctypes.byref(rc), ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)

    def alma_batch_into_host_f32(self, data_f32, sweep: AlmaBatchRange, out: np.ndarray):
        x = _f32(data_f32)
        rc = sweep.to_c()
        if out.dtype != np.float32 or not out.flags.c_contiguous:
            raise VtaError(4, "out must be a contiguous float32 array")
        check(lib().vta_alma_batch_into_host_f32(self.ctx.handle, f32_ptr(x), x.size,
                                                 ctypes.byref(rc), f32_ptr(out), out.size))
        combos = expand_grid(sweep)
        return len(combos), x.size, combos

    def alma_batch_device(self, d_prices, d_weights, d_periods, d_inv_norms, max_period: int,
                          series_len: int, n_combos: int, first_valid: int, d_out,
                          time_major: bool = False) -> None:
        # alma_wrapper.rs:625-669 shapes: weights [n_combos][max_period]
        require_len(d_prices, series_len, "prices")
        require_len(d_weights, int(n_combos) * int(max_period), "weights")
        require_len(d_periods, n_combos, "periods")
        require_len(d_inv_norms, n_combos, "inv_norms")
        require_len(d_out, int(n_combos) * int(series_len), "output")
        fn = lib().vta_alma_batch_device_tm if time_major else lib().vta_alma_batch_device
        check(fn(self.ctx.handle, ctypes.c_void_p(d_prices.ptr), ctypes.c_void_p(d_weights.ptr),
                 ctypes.c_void_p(d_periods.ptr), ctypes.c_void_p(d_inv_norms.ptr),
                 int(max_period), int(series_len), int(n_combos), int(first_valid),
                 ctypes.c_void_p(d_out.ptr)))

    def sweep_plan(self, sweep: AlmaBatchRange, series_len: int,
                   kernel: str = "auto") -> AlmaSweepPlan:
        return AlmaSweepPlan(self, sweep, series_len, kernel)

    def alma_multi_series_one_param_time_major_dev(self, data_tm_f32, cols: int, rows: int,
                                                   params: AlmaParams) -> DeviceArrayF32:
        x = _f32(data_tm_f32).reshape(-1)
        if x.size != cols * rows:
            raise VtaError(4, f"data length {x.size} != cols*rows {cols * rows}")
        pc = AlmaParamsC(int(params.period), float(params.offset), float(params.sigma))
        out = DeviceArrayF32C()
        check(lib().vta_alma_many_series_one_param_time_major_dev(
            self.ctx.handle, f32_ptr(x), int(cols), int(rows), ctypes.byref(pc),
            ctypes.byref(out)))
        return DeviceArrayF32(self.ctx, out)


HipAlma = CudaAlma
