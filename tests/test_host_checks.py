"""Buffer-length checks of the Python layer's enqueue-only (_device) forms and
packs, the counterpart of the reference wrappers' DeviceBuffer length checks
(ema_wrapper.rs:282-296, 400-409): a caller-owned device buffer shorter than
what the launch reads or writes raises InvalidInput before the C library is
called.  CPU only: the wrappers get a context whose handle is null, so a call
whose buffers are exactly the needed size gets past the length checks and is
then refused by the library's own null-context check; one element short, it
never reaches the library."""
from types import SimpleNamespace as NS

import pytest

import vectorta_amd as vta
from vectorta_amd import VtaError

T, C, MAXP = 50, 3, 7
COLS, ROWS = 4, 20
CR = COLS * ROWS


def buf(n, itemsize=4):
    return NS(ptr=0, nbytes=int(n) * itemsize)


def op(cls):
    o = object.__new__(cls)
    o.ctx = NS(handle=None)
    return o


def _refused_by_library(fn):
    with pytest.raises(VtaError) as e:
        fn()
    assert "buffer length mismatch" not in str(e.value), e.value
    assert e.value.status == 4


def _refused_by_length(fn, what):
    with pytest.raises(VtaError) as e:
        fn()
    assert "buffer length mismatch" in str(e.value) and what in str(e.value), e.value


def _cases():
    """(name, call(sizes) -> None, exact sizes, {argument: what})."""
    ema, smma, atr = op(vta.CudaEma), op(vta.CudaSmma), op(vta.CudaAtr)
    wil, swma, wma = op(vta.CudaWilders), op(vta.CudaSwma), op(vta.CudaWma)
    sma, bb, alma = op(vta.CudaSma), op(vta.CudaBollingerBands), op(vta.CudaAlma)
    batch = {"prices": T, "periods": C, "output": C * T}
    ms = {"time-major input": CR, "first_valids": COLS, "time-major output": CR}
    return [
        ("ema_batch_device",
         lambda s: ema.ema_batch_device(buf(s["prices"]), buf(s["periods"]), buf(s["alphas"]),
                                        T, 0, C, buf(s["output"])),
         dict(batch, alphas=C)),
        ("ema_many_series_one_param_device",
         lambda s: ema.ema_many_series_one_param_device(
             buf(s["time-major input"]), buf(s["first_valids"]), 5, 0.3, COLS, ROWS,
             buf(s["time-major output"])), ms),
        ("smma_batch_device",
         lambda s: smma.smma_batch_device(buf(s["prices"]), buf(s["periods"]), buf(s["warms"]),
                                          0, T, C, buf(s["output"])),
         dict(batch, warms=C)),
        ("smma_multi_series_one_param_device",
         lambda s: smma.smma_multi_series_one_param_device(
             buf(s["time-major input"]), 5, COLS, ROWS, buf(s["first_valids"]),
             buf(s["time-major output"])), ms),
        ("wilders_batch_device",
         lambda s: wil.wilders_batch_device(buf(s["prices"]), buf(s["periods"]),
                                            buf(s["alphas"]), buf(s["warms"]), T, 0, C,
                                            buf(s["output"])),
         dict(batch, alphas=C, warms=C)),
        ("wilders_many_series_one_param_device",
         lambda s: wil.wilders_many_series_one_param_device(
             buf(s["time-major input"]), buf(s["first_valids"]), 5, 0.2, COLS, ROWS,
             buf(s["time-major output"])), ms),
        ("swma_batch_device",
         lambda s: swma.swma_batch_device(buf(s["prices"]), buf(s["periods"]), buf(s["warms"]),
                                          T, C, MAXP, buf(s["output"])),
         dict(batch, warms=C)),
        ("swma_multi_series_one_param_device",
         lambda s: swma.swma_multi_series_one_param_device(
             buf(s["time-major input"]), None, 5, COLS, ROWS, buf(s["first_valids"]),
             buf(s["time-major output"])), ms),
        ("wma_multi_series_one_param_time_major_device",
         lambda s: wma.wma_multi_series_one_param_time_major_device(
             buf(s["time-major input"]), 5, COLS, ROWS, buf(s["first_valids"]),
             buf(s["time-major output"])), ms),
        ("tr_from_hlc_device",
         lambda s: atr.tr_from_hlc_device(buf(s["high"]), buf(s["low"]), buf(s["close"]), T, 0,
                                          buf(s["tr output"])),
         {"high": T, "low": T, "close": T, "tr output": T}),
        ("atr_batch_device_with_tr",
         lambda s: atr.atr_batch_device_with_tr(buf(s["prices"]), buf(s["periods"]), MAXP, T, 0,
                                                C, buf(s["output"])), batch),
        ("sma_batch_device",
         lambda s: sma.sma_batch_device(buf(s["prices"]), buf(s["periods"]), T, C, 0, MAXP,
                                        buf(s["output"])), batch),
        ("sma_many_series_one_param_device",
         lambda s: sma.sma_many_series_one_param_device(
             buf(s["time-major input"]), buf(s["first_valids"]), COLS, ROWS, 5,
             buf(s["time-major output"])), ms),
        ("bollinger_bands_many_series_one_param_device",
         lambda s: bb.bollinger_bands_many_series_one_param_device(
             buf(s["time-major input"]), buf(s["first_valids"]), buf(s["first_nonfinite"]),
             COLS, ROWS, 5, 2.0, 2.0, buf(s["upper"]), buf(s["middle"]), buf(s["lower"])),
         {"time-major input": CR, "first_valids": COLS, "first_nonfinite": COLS,
          "upper": CR, "middle": CR, "lower": CR}),
        ("alma_batch_device",
         lambda s: alma.alma_batch_device(buf(s["prices"]), buf(s["weights"]), buf(s["periods"]),
                                          buf(s["inv_norms"]), MAXP, T, C, 0, buf(s["output"])),
         {"prices": T, "weights": C * MAXP, "periods": C, "inv_norms": C, "output": C * T}),
        ("alma_batch_from_device_prices",
         lambda s: alma.alma_batch_from_device_prices(buf(s["prices"]), T, 0,
                                                      vta.AlmaBatchRange(period=(5, 9, 2))),
         {"prices": T}),
        ("RecPack",
         lambda s: vta.RecPack(NS(handle=None), "adx", [
             {"close": buf(s["close"]), "high": buf(s["high"]), "low": buf(s["low"]),
              "periods": buf(s["periods"]), "out": buf(s["output"]), "series_len": T,
              "first_valid": 0, "n_combos": C, "max_period": MAXP}]),
         {"close": T, "high": T, "low": T, "periods": C, "output": C * T}),
        ("BbPack",
         lambda s: vta.BbPack(NS(handle=None), [
             {"prices": buf(s["prices"]), "periods": buf(s["periods"]),
              "devup": buf(s["devup"], 8), "devdn": buf(s["devdn"], 8),
              "upper": buf(s["upper"]), "middle": buf(s["middle"]), "lower": buf(s["lower"]),
              "series_len": T, "first_valid": 0, "first_nonfinite": T, "n_combos": C,
              "max_period": MAXP}]),
         {"prices": T, "periods": C, "devup": C, "devdn": C, "upper": C * T, "middle": C * T,
          "lower": C * T}),
    ]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_device_buffer_lengths_checked(case):
    name, call, exact = case
    _refused_by_library(lambda: call(dict(exact)))
    for arg, n in exact.items():
        short = dict(exact, **{arg: n - 1})
        _refused_by_length(lambda: call(short), arg)


def test_unsized_buffers_pass_through():
    """Raw pointers (no byte size) and absent optional buffers are not checked."""
    ema = op(vta.CudaEma)
    _refused_by_library(lambda: ema.ema_batch_device(NS(ptr=0), NS(ptr=0), None, T, 0, C,
                                                     NS(ptr=0)))
    _refused_by_library(lambda: ema.ema_many_series_one_param_device(
        NS(ptr=0), buf(COLS), 5, 0.3, COLS, ROWS, buf(CR)))
