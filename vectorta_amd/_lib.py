"""ctypes binding of libvta_hip.so (C ABI declared in include/vta.h).

The library and its gfx950 code objects (vta_<family>.hsaco) live next to this
file.  Loading fails loudly when they are missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvta_hip.so")

VTA_OK = 0
STATUS_NAMES = {
    0: "Ok", 1: "Hip", 2: "OutOfMemory", 3: "MissingKernelSymbol", 4: "InvalidInput",
    5: "InvalidPolicy", 6: "LaunchConfigTooLarge", 7: "DeviceMismatch", 8: "NotImplemented",
    9: "ArithmeticOverflow", 10: "NoDevice", 11: "Timeout",
}


class VtaError(RuntimeError):
    """Mirrors the reference wrapper error enums (alma_wrapper.rs:72-101)."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.kind = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.kind}: {message}")


class RangeUsize(ctypes.Structure):
    _fields_ = [("start", ctypes.c_size_t), ("end", ctypes.c_size_t), ("step", ctypes.c_size_t)]


class RangeF64(ctypes.Structure):
    _fields_ = [("start", ctypes.c_double), ("end", ctypes.c_double), ("step", ctypes.c_double)]


class DeviceArrayF32C(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("rows", ctypes.c_size_t), ("cols", ctypes.c_size_t)]


class AlmaParamsC(ctypes.Structure):
    _fields_ = [("period", ctypes.c_size_t), ("offset", ctypes.c_double),
                ("sigma", ctypes.c_double)]


class AlmaBatchRangeC(ctypes.Structure):
    _fields_ = [("period", RangeUsize), ("offset", RangeF64), ("sigma", RangeF64)]


class AlmaJobC(ctypes.Structure):
    """vta_alma_job (include/vta.h)."""
    _fields_ = [("d_prices", ctypes.c_void_p), ("d_out", ctypes.c_void_p),
                ("range", AlmaBatchRangeC), ("series_len", ctypes.c_size_t),
                ("first_valid", ctypes.c_size_t)]


class RecJobC(ctypes.Structure):
    """vta_rec_job (include/vta.h)."""
    _fields_ = [("d_close", ctypes.c_void_p), ("d_high", ctypes.c_void_p),
                ("d_low", ctypes.c_void_p), ("d_periods", ctypes.c_void_p),
                ("d_out", ctypes.c_void_p), ("series_len", ctypes.c_int32),
                ("first_valid", ctypes.c_int32), ("n_combos", ctypes.c_int32),
                ("max_period", ctypes.c_int32)]


class BbJobC(ctypes.Structure):
    """vta_bb_job (include/vta.h)."""
    _fields_ = [("d_prices", ctypes.c_void_p), ("d_periods", ctypes.c_void_p),
                ("d_devup", ctypes.c_void_p), ("d_devdn", ctypes.c_void_p),
                ("d_upper", ctypes.c_void_p), ("d_middle", ctypes.c_void_p),
                ("d_lower", ctypes.c_void_p), ("series_len", ctypes.c_int32),
                ("first_valid", ctypes.c_int32), ("first_nonfinite", ctypes.c_int32),
                ("n_combos", ctypes.c_int32), ("max_period", ctypes.c_int32)]


class BbParamsC(ctypes.Structure):
    _fields_ = [("period", ctypes.c_size_t), ("devup", ctypes.c_double),
                ("devdn", ctypes.c_double)]


class BbBatchRangeC(ctypes.Structure):
    _fields_ = [("period", RangeUsize), ("devup", RangeF64), ("devdn", RangeF64)]


class PnrParamsC(ctypes.Structure):
    _fields_ = [("length", ctypes.c_size_t), ("percentage", ctypes.c_double)]


class PnrBatchRangeC(ctypes.Structure):
    _fields_ = [("length", RangeUsize), ("percentage", RangeF64)]


class VtaRequestC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("cache", ctypes.c_int32),
                ("series_key", ctypes.c_uint64), ("close", ctypes.c_void_p),
                ("high", ctypes.c_void_p), ("low", ctypes.c_void_p), ("len", ctypes.c_size_t),
                ("alma", AlmaBatchRangeC), ("period", RangeUsize), ("bb", BbBatchRangeC)]


class VtaResultC(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("n_outputs", ctypes.c_int32),
                ("n_combos", ctypes.c_size_t), ("out", DeviceArrayF32C * 3),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_done", ctypes.c_double), ("t_series", ctypes.c_double),
                ("t_launch", ctypes.c_double)]


class KernelPolicyC(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int32), ("batch_block_x", ctypes.c_uint32),
                ("many_series", ctypes.c_int32), ("many_series_block_x", ctypes.c_uint32)]


class KernelSelectedC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("block_x", ctypes.c_uint32)]


class VtaServerStatsC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_size_t) for n in ("budget", "in_use", "peak_in_use", "queued",
                                                "in_flight", "peak_in_flight", "cached_series",
                                                "packs", "packed_requests")]


# Every symbol include/vta.h declares, with (restype, argtypes).
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I32 = ctypes.c_int32
_F32 = ctypes.c_float
_ST = ctypes.c_int
SIGNATURES = {
    "vta_last_error": (ctypes.c_char_p, []),
    "vta_version": (ctypes.c_char_p, []),
    "vta_device_available": (ctypes.c_int, []),
    "vta_device_count": (ctypes.c_int, []),
    "vta_ctx_create": (_ST, [ctypes.c_int, ctypes.POINTER(_P)]),
    "vta_ctx_destroy": (_ST, [_P]),
    "vta_ctx_synchronize": (_ST, [_P]),
    "vta_ctx_stream": (_P, [_P]),
    "vta_ctx_device_id": (ctypes.c_int, [_P]),
    "vta_ctx_preload": (_ST, [_P]),
    "vta_ctx_set_tuning": (_ST, [_P, ctypes.c_int, ctypes.c_int64]),
    "vta_ctx_get_tuning": (_ST, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
    "vta_autotune": (_ST, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), _SZ, _P, _P,
                           ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                           ctypes.POINTER(ctypes.c_float)]),
    "vta_rec_pack_create": (_ST, [_P, ctypes.c_int, ctypes.POINTER(RecJobC), _SZ,
                                  ctypes.POINTER(_P)]),
    "vta_rec_pack_launch": (_ST, [_P, _P]),
    "vta_rec_pack_destroy": (_ST, [_P]),
    "vta_alma_pack_create": (_ST, [_P, ctypes.c_int, ctypes.POINTER(AlmaJobC), _SZ,
                                   ctypes.POINTER(_P)]),
    "vta_alma_pack_launch": (_ST, [_P, _P]),
    "vta_alma_pack_kernel": (ctypes.c_int, [_P]),
    "vta_alma_pack_job_combos": (_SZ, [_P, _SZ]),
    "vta_alma_pack_destroy": (_ST, [_P]),
    "vta_bb_pack_create": (_ST, [_P, ctypes.POINTER(BbJobC), _SZ, ctypes.POINTER(_P)]),
    "vta_bb_pack_launch": (_ST, [_P, _P]),
    "vta_bb_pack_destroy": (_ST, [_P]),
    "vta_ctx_begin_capture": (_ST, [_P]),
    "vta_ctx_end_capture": (_ST, [_P, ctypes.POINTER(_P)]),
    "vta_graph_launch": (_ST, [_P, _P]),
    "vta_graph_num_nodes": (_SZ, [_P]),
    "vta_graph_destroy": (_ST, [_P]),
    "vta_device_array_free": (_ST, [_P, ctypes.POINTER(DeviceArrayF32C)]),
    "vta_device_array_copy_to_host": (_ST, [_P, ctypes.POINTER(DeviceArrayF32C), _P, _SZ]),
    "vta_device_alloc": (_ST, [_P, _SZ, ctypes.POINTER(_P)]),
    "vta_device_free": (_ST, [_P, _P]),
    "vta_memcpy_htod": (_ST, [_P, _P, _P, _SZ]),
    "vta_memcpy_dtoh": (_ST, [_P, _P, _P, _SZ]),
    # ALMA
    "vta_alma_expand_grid": (_ST, [ctypes.POINTER(AlmaBatchRangeC), ctypes.POINTER(AlmaParamsC),
                                   _SZ, ctypes.POINTER(_SZ)]),
    "vta_alma_set_batch_policy": (_ST, [_P, ctypes.c_int]),
    "vta_alma_selected_batch_kernel": (ctypes.c_int, [_P]),
    "vta_alma_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(AlmaBatchRangeC),
                                 ctypes.POINTER(DeviceArrayF32C)]),
    "vta_alma_batch_from_device_prices": (_ST, [_P, _P, _SZ, _SZ, ctypes.POINTER(AlmaBatchRangeC),
                                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_alma_batch_into_host_f32": (_ST, [_P, _P, _SZ, ctypes.POINTER(AlmaBatchRangeC), _P,
                                           _SZ]),
    "vta_alma_batch_from_device_prices_noalloc": (_ST, [_P, _P, _SZ, _SZ,
                                                        ctypes.POINTER(AlmaBatchRangeC), _P,
                                                        _SZ]),
    "vta_alma_batch_device": (_ST, [_P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "vta_alma_batch_device_tm": (_ST, [_P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "vta_alma_sweep_plan_create": (_ST, [_P, ctypes.POINTER(AlmaBatchRangeC), _SZ, ctypes.c_int,
                                         ctypes.POINTER(_P)]),
    "vta_alma_sweep_plan_destroy": (_ST, [_P, _P]),
    "vta_alma_sweep_plan_combos": (_SZ, [_P]),
    "vta_alma_sweep_launch": (_ST, [_P, _P, _P, _SZ, _P]),
    "vta_alma_many_series_one_param_time_major_dev": (
        _ST, [_P, _P, _SZ, _SZ, ctypes.POINTER(AlmaParamsC), ctypes.POINTER(DeviceArrayF32C)]),
    "vta_alma_many_series_one_param_time_major_device": (
        _ST, [_P, _P, _P, _SZ, _F32, _SZ, _SZ, _P, _P]),
    "vta_device_array_to_dlpack": (_ST, [_P, ctypes.POINTER(DeviceArrayF32C), ctypes.c_int,
                                         ctypes.POINTER(_P)]),
    "vta_dlpack_delete": (None, [_P, ctypes.c_int]),
    # recurrences
    "vta_ema_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_ema_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_ema_batch_device": (_ST, [_P, _P, _P, _P, _SZ, _SZ, _SZ, _P]),
    "vta_ema_batch_into_host_f32": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize), _P, _SZ]),
    "vta_ema_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_ema_many_series_one_param_device": (_ST, [_P, _P, _P, _I32, _F32, _SZ, _SZ, _P]),
    "vta_ema_many_series_one_param_time_major_into_host_f32": (_ST, [_P, _P, _SZ, _SZ, _SZ, _P,
                                                                     _SZ]),
    "vta_smma_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_smma_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_smma_batch_device": (_ST, [_P, _P, _P, _P, _SZ, _SZ, _SZ, _P]),
    "vta_smma_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_smma_multi_series_one_param_device": (_ST, [_P, _P, _I32, _I32, _I32, _P, _P]),
    "vta_wilders_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_wilders_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_wilders_batch_device": (_ST, [_P, _P, _P, _P, _P, _I32, _I32, _I32, _P]),
    "vta_wilders_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_wilders_many_series_one_param_device": (_ST, [_P, _P, _P, _I32, _F32, _SZ, _SZ, _P]),
    "vta_ma_many_series_prepared_device": (_ST, [_P, ctypes.c_int, _P, _P, _P, _SZ, _F32, _SZ,
                                                 _SZ, _P]),
    "vta_tr_from_hlc_device": (_ST, [_P, _P, _P, _P, _I32, _I32, _P]),
    "vta_atr_batch_device_with_tr": (_ST, [_P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    # finite-window averages
    "vta_sma_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_sma_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_sma_batch_device": (_ST, [_P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "vta_sma_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_sma_many_series_one_param_device": (_ST, [_P, _P, _P, _SZ, _SZ, _SZ, _P]),
    "vta_sma_batch_dev_from_device": (_ST, [_P, _P, _SZ, _P, _SZ, _SZ,
                                            ctypes.POINTER(DeviceArrayF32C)]),
    "vta_sma_multi_series_one_param_time_major_dev_from_device": (
        _ST, [_P, _P, _P, _SZ, _SZ, _SZ, ctypes.POINTER(DeviceArrayF32C)]),
    "vta_wma_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_wma_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_wma_batch_device": (_ST, [_P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "vta_wma_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_wma_multi_series_one_param_time_major_device": (_ST, [_P, _P, _I32, _I32, _I32, _P,
                                                                _P]),
    "vta_swma_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_swma_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_swma_batch_device": (_ST, [_P, _P, _P, _P, _SZ, _SZ, _SZ, _P]),
    "vta_swma_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_swma_multi_series_one_param_device": (_ST, [_P, _P, _P, _I32, _I32, _I32, _P, _P]),
    "vta_rsi_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_rsi_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_rsi_batch_device": (_ST, [_P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "vta_rsi_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_rsi_many_series_one_param_device": (_ST, [_P, _P, _P, _SZ, _SZ, _SZ, _P, _P]),
    "vta_atr_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_atr_batch_dev": (_ST, [_P, _P, _P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_atr_batch_device": (_ST, [_P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "vta_atr_many_series_one_param_time_major_dev": (_ST, [_P, _P, _P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_atr_many_series_one_param_device": (_ST, [_P, _P, _P, _P, _P, _SZ, _SZ, _SZ, _P, _P]),
    "vta_adx_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ, ctypes.POINTER(_SZ)]),
    "vta_adx_batch_dev": (_ST, [_P, _P, _P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_adx_many_series_one_param_time_major_dev": (_ST, [_P, _P, _P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_adx_batch_into_host_f32": (_ST, [_P, _P, _P, _P, _SZ, ctypes.POINTER(RangeUsize), _P,
                                          _SZ]),
    "vta_adx_many_series_one_param_time_major_into_host_f32": (_ST, [_P, _P, _P, _P, _SZ, _SZ,
                                                                     _SZ, _P, _SZ]),
    "vta_ema_set_policy": (_ST, [_P, ctypes.POINTER(KernelPolicyC)]),
    "vta_ema_policy": (_ST, [_P, ctypes.POINTER(KernelPolicyC)]),
    "vta_ema_selected_batch_kernel": (_ST, [_P, ctypes.POINTER(KernelSelectedC)]),
    "vta_ema_selected_many_series_kernel": (_ST, [_P, ctypes.POINTER(KernelSelectedC)]),
    "vta_rsi_set_policy": (_ST, [_P, ctypes.POINTER(KernelPolicyC)]),
    "vta_rsi_policy": (_ST, [_P, ctypes.POINTER(KernelPolicyC)]),
    "vta_rsi_selected_batch_kernel": (_ST, [_P, ctypes.POINTER(KernelSelectedC)]),
    "vta_rsi_selected_many_series_kernel": (_ST, [_P, ctypes.POINTER(KernelSelectedC)]),
    "vta_adx_set_policy": (_ST, [_P, ctypes.POINTER(KernelPolicyC)]),
    "vta_adx_policy": (_ST, [_P, ctypes.POINTER(KernelPolicyC)]),
    "vta_adx_selected_batch_kernel": (_ST, [_P, ctypes.POINTER(KernelSelectedC)]),
    "vta_adx_selected_many_series_kernel": (_ST, [_P, ctypes.POINTER(KernelSelectedC)]),
    "vta_pnr_set_policy": (_ST, [_P, ctypes.POINTER(KernelPolicyC)]),
    "vta_pnr_policy": (_ST, [_P, ctypes.POINTER(KernelPolicyC)]),
    "vta_pnr_selected_batch_kernel": (_ST, [_P, ctypes.POINTER(KernelSelectedC)]),
    "vta_pnr_selected_many_series_kernel": (_ST, [_P, ctypes.POINTER(KernelSelectedC)]),
    "vta_many_series_prepare": (_ST, [_P, _P, _P, _P, _SZ, _SZ, ctypes.c_int, _P, _P, _P]),
    "vta_ema_rsi_atr_many_series_device": (_ST, [_P, _P, _P, _P, _SZ, _SZ, _SZ, _SZ, _SZ, _P,
                                                 _P, _P, _P, _P, _P]),
    "vta_ema_rsi_atr_many_series_dev": (_ST, [_P, _P, _P, _P, _SZ, _SZ, _SZ, _SZ, _SZ,
                                              ctypes.POINTER(DeviceArrayF32C)]),
    # rolling windows
    "vta_bb_expand_grid": (_ST, [ctypes.POINTER(BbBatchRangeC), ctypes.POINTER(BbParamsC), _SZ,
                                 ctypes.POINTER(_SZ)]),
    "vta_bollinger_bands_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(BbBatchRangeC),
                                            ctypes.POINTER(DeviceArrayF32C)]),
    "vta_bollinger_bands_many_series_one_param_time_major_dev": (
        _ST, [_P, _P, _SZ, _SZ, _SZ, ctypes.c_double, ctypes.c_double,
              ctypes.POINTER(DeviceArrayF32C)]),
    "vta_bollinger_bands_many_series_one_param_device": (
        _ST, [_P, _P, _P, _P, _SZ, _SZ, _SZ, ctypes.c_double, ctypes.c_double, _P, _P, _P]),
    "vta_pnr_expand_grid": (_ST, [ctypes.POINTER(PnrBatchRangeC), ctypes.POINTER(PnrParamsC),
                                  _SZ, ctypes.POINTER(_SZ)]),
    "vta_pnr_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(PnrBatchRangeC),
                                ctypes.POINTER(DeviceArrayF32C)]),
    "vta_pnr_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                           ctypes.c_double,
                                                           ctypes.POINTER(DeviceArrayF32C)]),
    "vta_pnr_many_series_one_param_device": (_ST, [_P, _P, _P, _SZ, _SZ, _SZ, ctypes.c_double,
                                                   _P]),
    "vta_medium_ad_many_series_one_param_device": (_ST, [_P, _P, _P, _SZ, _SZ, _SZ, _P]),
    "vta_double_alma_crossover": (_ST, [_P, _P, _SZ, ctypes.POINTER(AlmaBatchRangeC),
                                        ctypes.POINTER(AlmaBatchRangeC), _P, _P, _SZ,
                                        ctypes.POINTER(_SZ)]),
    "vta_double_alma_plan_create": (_ST, [_P, _P, _SZ, ctypes.POINTER(AlmaBatchRangeC),
                                          ctypes.POINTER(AlmaBatchRangeC), ctypes.POINTER(_P)]),
    "vta_double_alma_plan_pairs": (_SZ, [_P]),
    "vta_double_alma_plan_launch": (_ST, [_P]),
    "vta_double_alma_plan_results": (_ST, [_P, _P, _P, _SZ]),
    "vta_double_alma_plan_destroy": (_ST, [_P]),
    "vta_adx_batch_device": (_ST, [_P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "vta_bollinger_bands_batch_device": (_ST, [_P, _P, _I32, _I32, _I32, _P, _P, _P, _I32,
                                               _I32, _P, _P, _P]),
    "vta_server_create": (_ST, [ctypes.c_int, ctypes.c_int, _SZ, ctypes.POINTER(_P)]),
    "vta_server_destroy": (_ST, [_P]),
    "vta_server_submit": (_ST, [_P, ctypes.POINTER(VtaRequestC), ctypes.POINTER(ctypes.c_uint64)]),
    "vta_server_wait": (_ST, [_P, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(VtaResultC)]),
    "vta_server_release": (_ST, [_P, ctypes.c_uint64]),
    "vta_server_get_stats": (_ST, [_P, ctypes.POINTER(VtaServerStatsC)]),
    "vta_server_now": (ctypes.c_double, [_P]),
    "vta_medium_ad_expand_grid": (_ST, [ctypes.POINTER(RangeUsize), _P, _SZ,
                                        ctypes.POINTER(_SZ)]),
    "vta_medium_ad_batch_dev": (_ST, [_P, _P, _SZ, ctypes.POINTER(RangeUsize),
                                      ctypes.POINTER(DeviceArrayF32C)]),
    "vta_medium_ad_many_series_one_param_time_major_dev": (_ST, [_P, _P, _SZ, _SZ, _SZ,
                                                                 ctypes.POINTER(DeviceArrayF32C)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads libvta_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C vectorta_amd/csrc); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != VTA_OK:
        msg = lib().vta_last_error()
        raise VtaError(status, msg.decode() if msg else "")
