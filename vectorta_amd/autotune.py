"""Autotuner for launch shapes: picks each kernel family's knob by timing the
caller's own workload on the device, and remembers the winner per shape.

The reference has one shape knob, ``ALMA_TILE`` read from the environment
(src/cuda/moving_averages/alma_wrapper.rs:314), plus fixed kernel policies
(:18-70).  Here the knobs live on a context (``vta_ctx_set_tuning``), the
timing loop is native (``vta_autotune``: one untimed run per candidate, then
``reps`` runs back to back between two HIP events on the context's stream),
and this module adds the candidate lists and a cache keyed by device, knob and
a shape bucket (powers of two of the sizes that set the launch shape), kept in
memory and optionally in a JSON file, so a shape is measured once.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
from typing import Callable, Dict, Optional, Sequence, Tuple

from ._lib import VtaError, check, lib

ALMA_MFMA_VARIANT = 0
SCAN_WAVES = 1
FIR_GROUP = 2
MS_CHUNK = 3
ROLL_FORM = 4
KNOB_NAMES = {ALMA_MFMA_VARIANT: "alma_mfma_variant", SCAN_WAVES: "scan_waves",
              FIR_GROUP: "fir_group", MS_CHUNK: "ms_chunk", ROLL_FORM: "roll_form"}

# Candidate values per knob (every value is valid for vta_ctx_set_tuning).
CANDIDATES: Dict[int, Tuple[int, ...]] = {
    ALMA_MFMA_VARIANT: (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11),
    SCAN_WAVES: (1024, 2048, 4096, 8192, 16384),
    FIR_GROUP: (16, 32, 64, 128),
    MS_CHUNK: (512, 1024, 2048, 4096),
    ROLL_FORM: (1, 2),
}

ENQUEUE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)


def set_tuning(ctx, knob: int, value: int) -> None:
    check(lib().vta_ctx_set_tuning(ctx.handle, int(knob), int(value)))


def get_tuning(ctx, knob: int) -> int:
    v = ctypes.c_int64()
    check(lib().vta_ctx_get_tuning(ctx.handle, int(knob), ctypes.byref(v)))
    return int(v.value)


def shape_bucket(*sizes: int) -> Tuple[int, ...]:
    """Rounds each size to the nearest power of two (log2 rounded), so shapes
    that launch alike share one tuning decision."""
    return tuple(int(round(math.log2(max(int(s), 1)))) for s in sizes)


def measure(ctx, knob: int, candidates: Sequence[int], enqueue: Callable[[], None],
            reps: int = 5) -> Tuple[int, Dict[int, Optional[float]]]:
    """Runs vta_autotune over `candidates` with `enqueue` (a Python callable
    that enqueues the workload on ctx, e.g. a sweep plan launch) and returns
    (best value, {candidate: mean ms or None if it could not launch}).  The
    context's knob is left at the best value."""
    err: list = []

    def _cb(_ctx, _user):
        try:
            enqueue()
            return 0
        except VtaError as e:
            err.append(e)
            return int(e.status)
        except Exception as e:  # a Python error must not cross the C frame
            err.append(e)
            return 4

    cb = ENQUEUE_FN(_cb)
    n = len(candidates)
    cand = (ctypes.c_int64 * n)(*[int(c) for c in candidates])
    ms = (ctypes.c_float * n)()
    best = ctypes.c_int64()
    st = lib().vta_autotune(ctx.handle, int(knob), cand, n, ctypes.cast(cb, ctypes.c_void_p),
                            None, int(reps), ctypes.byref(best), ms)
    if st != 0:
        if err and not isinstance(err[-1], VtaError):
            raise err[-1]
        check(st)
    times = {int(c): (float(ms[i]) if ms[i] >= 0 else None) for i, c in enumerate(candidates)}
    return int(best.value), times


class Autotuner:
    """Per-device cache of tuning decisions.

    ``tune(ctx, knob, shape, enqueue)`` returns the cached value for
    (device, knob, shape_bucket(*shape)) after setting it on ctx, or measures
    the knob's candidates on `enqueue` first.  With `path`, decisions are read
    from and written to that JSON file."""

    def __init__(self, path: Optional[str] = None, reps: int = 5):
        self.path = path
        self.reps = int(reps)
        self.table: Dict[str, dict] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    @staticmethod
    def key(device: str, knob: int, shape: Sequence[int]) -> str:
        return f"{device}|{KNOB_NAMES[knob]}|{','.join(map(str, shape_bucket(*shape)))}"

    def lookup(self, device: str, knob: int, shape: Sequence[int]) -> Optional[int]:
        e = self.table.get(self.key(device, knob, shape))
        return None if e is None else int(e["value"])

    def record(self, device: str, knob: int, shape: Sequence[int], value: int,
               times: Dict[int, Optional[float]]) -> None:
        self.table[self.key(device, knob, shape)] = {
            "value": int(value), "ms": {str(k): v for k, v in times.items()}}
        if self.path:
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)

    def tune(self, ctx, knob: int, shape: Sequence[int], enqueue: Callable[[], None],
             candidates: Optional[Sequence[int]] = None, device: str = "gfx950") -> int:
        v = self.lookup(device, knob, shape)
        if v is not None:
            set_tuning(ctx, knob, v)
            return v
        v, times = measure(ctx, knob, candidates or CANDIDATES[knob], enqueue, self.reps)
        self.record(device, knob, shape, v, times)
        return v

    # convenience forms for the workloads whose launch shape has a knob
    def tune_alma_sweep(self, plan, d_prices: int, first_valid: int, d_out: int) -> int:
        """The split-BF16 sweep kernel variant for this plan's shape (series
        length, combos)."""
        return self.tune(plan.alma.ctx, ALMA_MFMA_VARIANT, (plan.series_len, plan.n_combos),
                         lambda: plan.launch(d_prices, first_valid, d_out))
