"""CPU-side checks of the C-ABI library: it loads, exports every symbol that
include/vta.h declares, ships a gfx950 code object holding every kernel the
runtime looks up, and its host logic (grid expansion, validation, device probe)
behaves like the reference wrappers.  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

import vectorta_amd as vta
from vectorta_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vta.h")
PKG = os.path.join(ROOT, "vectorta_amd")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # function-pointer typedefs name a type, not an exported symbol
    text = re.sub(r"typedef\s+\w+\s*\(\s*\*[^;]*;", "", text)
    return sorted(set(re.findall(r"\b(vta_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers every declared symbol
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def _kernel_names(path):
    data = open(path, "rb").read()
    return set(re.findall(rb"[a-z_0-9]+", data))


def test_code_objects_hold_kernels():
    from vectorta_amd._lib import LIB_PATH
    src = open(os.path.join(PKG, "csrc", "runtime.cpp")).read()
    # every (module, symbol) pair vta_ctx_preload checks
    pairs = re.findall(r'\{"([a-z]+)", "([a-z0-9_]+)"\}', src)
    assert pairs
    for module, sym in pairs:
        path = os.path.join(PKG, f"vta_{module}.hsaco")
        assert os.path.exists(path), path
        assert sym.encode() in _kernel_names(path), (module, sym)
        assert b"gfx950" in open(path, "rb").read()
    assert os.path.exists(LIB_PATH)


def test_version_and_probe_forced_off(monkeypatch):
    assert b"gfx950" in _lib.lib().vta_version()
    monkeypatch.setenv("VTA_FORCE_SKIP", "1")
    assert _lib.lib().vta_device_available() == 0
    p = ctypes.c_void_p()
    st = _lib.lib().vta_ctx_create(0, ctypes.byref(p))
    assert st == 10 and not p.value          # VTA_ERR_NO_DEVICE
    assert b"no HIP device" in _lib.lib().vta_last_error()


def _rust_axis_f64(start, end, step):
    # alma.rs:1220-1250 semantics: x accumulates by repeated addition
    if abs(step) < 1e-12 or abs(start - end) < 1e-12:
        return [start]
    out = []
    if start < end:
        x = start
        while x <= end + 1e-12:
            out.append(x)
            x += abs(step)
        return out
    x = start
    while x + 1e-12 >= end:
        out.append(x)
        x -= abs(step)
    return out


def test_alma_expand_grid_matches_reference_semantics():
    r = vta.AlmaBatchRange(period=(10, 250, 10), offset=(0.05, 0.95, 0.1), sigma=(2.0, 10.0, 2.0))
    combos = vta.expand_grid(r)
    assert len(combos) == 1250
    offs = _rust_axis_f64(0.05, 0.95, 0.1)
    assert len(offs) == 10
    # exact bits of the accumulated offsets (0.15000000000000002, ...)
    assert [c.offset for c in combos[:50:5]] == offs
    assert combos[1].offset == 0.05 and combos[5].offset == offs[1]
    assert offs[1] == 0.15000000000000002
    # order: period -> offset -> sigma
    assert [c.period for c in combos[::50]] == list(range(10, 251, 10))
    assert [c.sigma for c in combos[:5]] == [2.0, 4.0, 6.0, 8.0, 10.0]
    # descending period axis walks down from start (alma_wrapper.rs:1557-1567)
    d = vta.expand_grid(vta.AlmaBatchRange(period=(30, 10, 7)))
    assert [c.period for c in d] == [30, 23, 16]
    s = vta.expand_grid(vta.AlmaBatchRange(period=(9, 9, 0)))
    assert [c.period for c in s] == [9]


def test_alma_invalid_policy_rejected_without_device():
    with pytest.raises(KeyError):
        vta.alma._POLICY["bogus"]


def test_rolling_expand_grids_match_reference_semantics():
    # bollinger_bands_wrapper.rs:179-281: period x devup x devdn, descending usize
    bb = vta.CudaBollingerBands.expand_grid(
        vta.BollingerBandsBatchRange(period=(30, 10, 10), devup=(1.0, 2.0, 0.5),
                                     devdn=(2.0, 2.0, 0.0)))
    assert [c[0] for c in bb] == [30] * 3 + [20] * 3 + [10] * 3
    assert [c[1] for c in bb[:3]] == [1.0, 1.5, 2.0]
    assert all(c[2] == 2.0 for c in bb)
    # percentile_nearest_rank_wrapper.rs:200-260: descending f64 axis
    pnr = vta.CudaPercentileNearestRank.expand_grid(
        vta.PercentileNearestRankBatchRange(length=(10, 20, 5), percentage=(90.0, 70.0, 10.0)))
    assert pnr == [(l, p) for l in (10, 15, 20) for p in (90.0, 80.0, 70.0)]
    # medium_ad_wrapper.rs:120-150
    assert vta.CudaMediumAd.expand_grid((64, 60, 2)) == [64, 62, 60]
    assert vta.CudaMediumAd.expand_grid((5, 9, 0)) == [5]


def test_moving_average_grids_match_reference_semantics():
    # expand_grid_sma (sma.rs:705-755), swma_wrapper.rs:275-301, smma.rs:430-471,
    # wilders_wrapper.rs:909-940: ascending start..=end, or counting down to end
    for op in (vta.CudaSma, vta.CudaSwma, vta.CudaSmma, vta.CudaWilders):
        e = op.expand_grid
        assert e((5, 20, 5)) == [5, 10, 15, 20]
        assert e((5, 22, 5)) == [5, 10, 15, 20]
        assert e((10, 5, 3)) == [10, 7]
        assert e((10, 4, 3)) == [10, 7, 4]
        assert e((9, 9, 3)) == [9] and e((9, 30, 0)) == [9]
    # wma_wrapper.rs:324-350: the countdown appends end when it misses it
    e = vta.CudaWma.expand_grid
    assert e((10, 5, 3)) == [10, 7, 5]
    assert e((10, 4, 3)) == [10, 7, 4]
    assert e((4, 12, 4)) == [4, 8, 12]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device_type", ctypes.c_int32),
                ("device_id", ctypes.c_int32), ("ndim", ctypes.c_int32),
                ("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16),
                ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManaged(ctypes.Structure):
    _fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                ("deleter", ctypes.c_void_p)]


class _DLManagedVersioned(ctypes.Structure):
    _fields_ = [("major", ctypes.c_uint32), ("minor", ctypes.c_uint32),
                ("manager_ctx", ctypes.c_void_p), ("deleter", ctypes.c_void_p),
                ("flags", ctypes.c_uint64), ("dl_tensor", _DLTensor)]


@pytest.mark.parametrize("versioned", [0, 1])
def test_dlpack_export_layout_without_device(versioned):
    """The DLPack structs follow the public ABI (kDLROCM = 10, float32, row-major
    strides in elements) and ownership moves out of the array; an empty array
    exports without a device (dlpack_cuda.rs:13-276 is the reference form)."""
    arr = _lib.DeviceArrayF32C()
    arr.ptr = None
    arr.rows, arr.cols = 0, 7
    mt = ctypes.c_void_p()
    _lib.check(_lib.lib().vta_device_array_to_dlpack(None, ctypes.byref(arr), versioned,
                                                     ctypes.byref(mt)))
    m = ctypes.cast(mt, ctypes.POINTER(_DLManagedVersioned if versioned else _DLManaged))[0]
    if versioned:
        assert (m.major, m.minor, m.flags) == (1, 0, 0)
    d = m.dl_tensor
    assert (d.device_type, d.device_id, d.ndim) == (10, 0, 2)
    assert (d.code, d.bits, d.lanes, d.byte_offset) == (2, 32, 1, 0)
    assert (d.shape[0], d.shape[1], d.strides[0], d.strides[1]) == (0, 7, 7, 1)
    assert not d.data and m.deleter
    assert arr.rows == 0 and arr.cols == 0
    _lib.lib().vta_dlpack_delete(mt, versioned)


# Entry points whose null-handle result is a documented no-op or sentinel
# rather than VTA_ERR_INVALID_INPUT: destructors accept NULL like free(), and
# the getters return their "none" value.
_NULL_SENTINELS = {
    "vta_ctx_destroy": 0, "vta_device_free": 0, "vta_device_array_free": 0,
    "vta_dlpack_delete": None, "vta_alma_sweep_plan_destroy": 0, "vta_rec_pack_destroy": 0,
    "vta_bb_pack_destroy": 0, "vta_graph_destroy": 0, "vta_double_alma_plan_destroy": 0,
    "vta_server_destroy": 0, "vta_ctx_device_id": -1, "vta_ctx_stream": None,
    "vta_alma_selected_batch_kernel": 0, "vta_alma_sweep_plan_combos": 0,
    "vta_graph_num_nodes": 0, "vta_double_alma_plan_pairs": 0, "vta_server_now": 0.0,
    "vta_alma_pack_destroy": 0, "vta_alma_pack_kernel": 0, "vta_alma_pack_job_combos": 0,
}

_NULL_PROBE = r"""
import ctypes, json, sys
from vectorta_amd import _lib
L = _lib.lib()
ints = (ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t, ctypes.c_uint32,
        ctypes.c_uint64)
for name, (res, args) in sorted(_lib.SIGNATURES.items()):
    if not args:
        continue
    vals = [0.0 if a in (ctypes.c_float, ctypes.c_double) else 0 if a in ints else None
            for a in args]
    print("calling", name, flush=True)
    r = getattr(L, name)(*vals)
    e = L.vta_last_error()
    print(json.dumps({"name": name, "ret": r, "err": e.decode() if e else ""}), flush=True)
"""


def test_null_arguments_rejected_without_device():
    """Every entry point given null handles and zero sizes returns
    VTA_ERR_INVALID_INPUT with a message before touching HIP (the reference
    wrappers' InvalidInput, alma_wrapper.rs:72-101), or its documented no-op /
    sentinel; none crashes.  Runs in a child process so a crash names the
    entry point it happened in."""
    import json
    import subprocess
    import sys
    p = subprocess.run([sys.executable, "-c", _NULL_PROBE], cwd=ROOT, capture_output=True,
                       text=True, timeout=120)
    lines = p.stdout.splitlines()
    assert p.returncode == 0, (p.returncode, lines[-1:] if lines else p.stderr[-500:])
    seen = {}
    for ln in lines:
        if ln.startswith("{"):
            d = json.loads(ln)
            seen[d["name"]] = d
    expect = {n for n, (_, a) in _lib.SIGNATURES.items() if a}
    assert set(seen) == expect
    for name, d in seen.items():
        if name in _NULL_SENTINELS:
            assert d["ret"] == _NULL_SENTINELS[name], d
        else:
            assert d["ret"] == 4 and d["err"], d


_GRID_PROBE = r"""
import ctypes, json
from vectorta_amd import _lib
L = _lib.lib()
U, F = _lib.RangeUsize, _lib.RangeF64
M, INF = 2**64 - 1, float("inf")
n = ctypes.c_size_t()
def run(tag, call):
    print("calling", tag, flush=True)
    st = call()
    print(json.dumps({"tag": tag, "st": st, "n": n.value}), flush=True)
alma = lambda p, o, s: L.vta_alma_expand_grid(
    ctypes.byref(_lib.AlmaBatchRangeC(U(*p), F(*o), F(*s))), None, 0, ctypes.byref(n))
run("alma_huge_period", lambda: alma((1, M, 1), (0.85, 0.85, 0), (6, 6, 0)))
run("alma_inf_offset", lambda: alma((9, 9, 0), (0.0, INF, 0.1), (6, 6, 0)))
run("alma_dense_sigma", lambda: alma((9, 9, 0), (0.85, 0.85, 0), (1.0, 100.0, 1e-9)))
run("alma_grid_product", lambda: alma((1, 4096, 1), (0.0, 0.4095, 1e-4), (1.0, 2.0, 1.0)))
run("ema_limit", lambda: L.vta_ema_expand_grid(ctypes.byref(U(1, 1 << 24, 1)), None, 0,
                                              ctypes.byref(n)))
run("ema_over_limit", lambda: L.vta_ema_expand_grid(ctypes.byref(U(1, (1 << 24) + 1, 1)), None,
                                                   0, ctypes.byref(n)))
run("mad_huge_down", lambda: L.vta_medium_ad_expand_grid(ctypes.byref(U(M, 1, 1)), None, 0,
                                                        ctypes.byref(n)))
run("bb_inf_devdn", lambda: L.vta_bb_expand_grid(
    ctypes.byref(_lib.BbBatchRangeC(U(20, 20, 0), F(2, 2, 0), F(-INF, 2.0, 0.5))), None, 0,
    ctypes.byref(n)))
run("bb_grid_product", lambda: L.vta_bb_expand_grid(
    ctypes.byref(_lib.BbBatchRangeC(U(1, 4096, 1), F(0, 4095, 1), F(1, 2, 1))), None, 0,
    ctypes.byref(n)))
run("pnr_dense_pct", lambda: L.vta_pnr_expand_grid(
    ctypes.byref(_lib.PnrBatchRangeC(U(20, 20, 0), F(0.0, 100.0, 1e-6))), None, 0,
    ctypes.byref(n)))
run("alma_stuck_sigma", lambda: alma((9, 9, 0), (0.85, 0.85, 0), (1e17, 1e17 + 4096, 1.0)))
run("alma_stuck_offset_down", lambda: alma((9, 9, 0), (1e17 + 4096, 1e17, 1.0), (6, 6, 0)))
run("bb_stuck_devup", lambda: L.vta_bb_expand_grid(
    ctypes.byref(_lib.BbBatchRangeC(U(20, 20, 0), F(1e17, 1e17 + 4096, 1.0), F(2, 2, 0))), None,
    0, ctypes.byref(n)))
"""


def test_oversized_and_infinite_grids_refused():
    """Parameter ranges whose expansion would exhaust host memory or never end
    (a usize axis up to SIZE_MAX, an infinite f64 bound, a step of 1e-9 over a
    span of 100, a grid product above 2^24, a step below half the rounding
    unit of the start, which never moves x += step) return InvalidInput instead of
    aborting the process; an axis of exactly 2^24 values is still expanded.
    Grids below the limit follow the reference's rules unchanged (the grid
    tests above)."""
    import json
    import subprocess
    import sys
    p = subprocess.run([sys.executable, "-c", _GRID_PROBE], cwd=ROOT, capture_output=True,
                       text=True, timeout=120)
    lines = p.stdout.splitlines()
    assert p.returncode == 0, (p.returncode, lines[-1:], p.stderr[-300:])
    got = {d["tag"]: d for d in (json.loads(ln) for ln in lines if ln.startswith("{"))}
    assert len(got) == 13
    assert got["ema_limit"]["st"] == 0 and got["ema_limit"]["n"] == 1 << 24
    for tag, d in got.items():
        if tag != "ema_limit":
            assert d["st"] == 4, d
