"""Double-ALMA crossover backtest reduction over fast x slow ALMA parameter
grids (no reference counterpart; the reference's README.md:9 names the
workload).  Positions are sign(ALMA_fast - ALMA_slow), equal to the f64
ALMA's signs; per pair: trades, bars long, bars short, an exact sign checksum
sum(pos_t * (t + 1)) and the pnl sum(pos_t * (price[t+1] - price[t]))."""
from __future__ import annotations

import ctypes
from typing import Tuple

import numpy as np

from ._lib import check, lib
from .alma import AlmaBatchRange
from .device import HipContext, f32_ptr

STAT_FIELDS = ("trades", "n_long", "n_short", "checksum")


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class DoubleAlmaCrossover:
    def __init__(self, device_id: int = 0):
        self.ctx = HipContext(device_id)

    def n_pairs(self, fast: AlmaBatchRange, slow: AlmaBatchRange) -> int:
        fr, sr = fast.to_c(), slow.to_c()
        n = ctypes.c_size_t()
        check(lib().vta_double_alma_crossover(self.ctx.handle, None, 0, ctypes.byref(fr),
                                              ctypes.byref(sr), None, None, 0, ctypes.byref(n)))
        return n.value

    def run(self, prices, fast: AlmaBatchRange, slow: AlmaBatchRange
            ) -> Tuple[np.ndarray, np.ndarray]:
        """Returns (stats int64 [n_fast * n_slow, 4], pnl float64 [n_fast * n_slow]);
        pair i * n_slow + j is (fast combo i, slow combo j)."""
        x = _f32(prices)
        n = self.n_pairs(fast, slow)
        stats = np.zeros((n, 4), dtype=np.int64)
        pnl = np.zeros(n, dtype=np.float64)
        fr, sr = fast.to_c(), slow.to_c()
        got = ctypes.c_size_t()
        check(lib().vta_double_alma_crossover(
            self.ctx.handle, f32_ptr(x), x.size, ctypes.byref(fr), ctypes.byref(sr),
            stats.ctypes.data_as(ctypes.c_void_p), pnl.ctypes.data_as(ctypes.c_void_p), n,
            ctypes.byref(got)))
        return stats, pnl


class DoubleAlmaPlan:
    """Prices resident on the device and tables built once; launch() enqueues
    the f64 rows, the pair walk and the reduction; results() waits and copies."""

    def __init__(self, ctx: HipContext, prices, fast: AlmaBatchRange, slow: AlmaBatchRange):
        self.ctx = ctx
        x = _f32(prices)
        fr, sr = fast.to_c(), slow.to_c()
        h = ctypes.c_void_p()
        check(lib().vta_double_alma_plan_create(ctx.handle, f32_ptr(x), x.size, ctypes.byref(fr),
                                                ctypes.byref(sr), ctypes.byref(h)))
        self.handle = h
        self.pairs = int(lib().vta_double_alma_plan_pairs(h))

    def launch(self) -> None:
        check(lib().vta_double_alma_plan_launch(self.handle))

    def results(self) -> Tuple[np.ndarray, np.ndarray]:
        stats = np.zeros((self.pairs, 4), dtype=np.int64)
        pnl = np.zeros(self.pairs, dtype=np.float64)
        check(lib().vta_double_alma_plan_results(self.handle,
                                                 stats.ctypes.data_as(ctypes.c_void_p),
                                                 pnl.ctypes.data_as(ctypes.c_void_p), self.pairs))
        return stats, pnl

    def close(self) -> None:
        if self.handle:
            check(lib().vta_double_alma_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
