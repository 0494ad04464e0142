"""Continuous-batching request server over the C ABI (vta_server_*): worker
slots with their own streams, HBM-budget admission, a device-resident series
cache, device-resident results."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import (RangeUsize, VtaError, VtaRequestC, VtaResultC, VtaServerStatsC, check, lib,
                   STATUS_NAMES)
from .alma import AlmaBatchRange
from .device import HipContext
from .rolling import BollingerBandsBatchRange

KINDS = {"alma": 0, "ema": 1, "rsi": 2, "bollinger_bands": 3, "adx": 4}


@dataclass
class ServedResult:
    req_id: int
    status: int
    n_outputs: int
    n_combos: int
    outputs: List[Tuple[int, int, int]]      # (device ptr, rows, cols)
    t_submit: float
    t_start: float
    t_done: float
    t_series: float = 0.0
    t_launch: float = 0.0

    @property
    def ok(self) -> bool:
        return self.status == 0

    @property
    def latency_s(self) -> float:
        return self.t_done - self.t_submit

    @property
    def n_values(self) -> int:
        return sum(r * c for _, r, c in self.outputs)


def _f32(a) -> Optional[np.ndarray]:
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


class IndicatorServer:
    def __init__(self, device_id: int = 0, n_slots: int = 4, hbm_budget: int = 32 << 30):
        h = ctypes.c_void_p()
        check(lib().vta_server_create(int(device_id), int(n_slots), int(hbm_budget),
                                      ctypes.byref(h)))
        self.handle = h
        self.ctx = HipContext(device_id)       # for copies of results to the host

    def submit(self, kind: str, close, high=None, low=None, *, key: int = 0, cache: bool = False,
               alma: Optional[AlmaBatchRange] = None, period: Sequence[int] = (0, 0, 0),
               bb: Optional[BollingerBandsBatchRange] = None) -> int:
        c, h, l = _f32(close), _f32(high), _f32(low)
        r = VtaRequestC()
        r.kind = KINDS[kind]
        r.cache = 1 if cache else 0
        r.series_key = int(key)
        r.close = c.ctypes.data
        r.high = h.ctypes.data if h is not None else None
        r.low = l.ctypes.data if l is not None else None
        r.len = c.size
        if alma is not None:
            r.alma = alma.to_c()
        r.period = RangeUsize(*[int(v) for v in period])
        if bb is not None:
            r.bb = bb.to_c()
        rid = ctypes.c_uint64()
        check(lib().vta_server_submit(self.handle, ctypes.byref(r), ctypes.byref(rid)))
        return rid.value

    def wait(self, req_id: int, timeout_ms: int = -1) -> ServedResult:
        res = VtaResultC()
        check(lib().vta_server_wait(self.handle, ctypes.c_uint64(req_id), int(timeout_ms),
                                    ctypes.byref(res)))
        outs = [(res.out[i].ptr or 0, res.out[i].rows, res.out[i].cols)
                for i in range(res.n_outputs if res.status == 0 else 0)]
        return ServedResult(req_id, res.status, res.n_outputs, res.n_combos, outs, res.t_submit,
                            res.t_start, res.t_done, res.t_series, res.t_launch)

    def to_host(self, result: ServedResult, i: int = 0) -> np.ndarray:
        if not result.ok:
            raise VtaError(result.status, f"request {result.req_id} failed "
                                          f"({STATUS_NAMES.get(result.status)})")
        ptr, rows, cols = result.outputs[i]
        out = np.empty((rows, cols), dtype=np.float32)
        check(lib().vta_memcpy_dtoh(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_void_p(ptr), out.nbytes))
        return out

    def release(self, req_id: int) -> None:
        check(lib().vta_server_release(self.handle, ctypes.c_uint64(req_id)))

    def stats(self) -> dict:
        st = VtaServerStatsC()
        check(lib().vta_server_get_stats(self.handle, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in VtaServerStatsC._fields_}

    def now(self) -> float:
        return float(lib().vta_server_now(self.handle))

    def close(self) -> None:
        if self.handle:
            check(lib().vta_server_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
