"""vectorta_amd -- MI355X-native (gfx950) implementation of VectorTA's GPU
batch-indicator path.

The product is ``libvta_hip.so`` (C ABI in ``include/vta.h``) plus one gfx950
code object per indicator family; this package is the Python host layer over
it, mirroring the reference's ``src/cuda`` wrapper classes (``CudaAlma`` ...).
"""
from ._lib import VtaError, lib
from .alma import (AlmaBatchRange, AlmaPack, AlmaParams, AlmaSweepPlan, CudaAlma, HipAlma,
                   expand_grid)
from .cross import DoubleAlmaCrossover, DoubleAlmaPlan
from .device import (DeviceArrayF32, DeviceBuffer, HipContext, HipGraph, cuda_available,
                     cuda_device_count)
from .fir import CudaSma, CudaSwma, CudaWma
from .rec import (CudaAdx, CudaAtr, CudaEma, CudaRsi, CudaSmma, CudaWilders, KernelPolicy,
                  ManySeriesOhlcPlan, RecPack, ema_rsi_atr_many_series_dev)
from .server import IndicatorServer, ServedResult
from .rolling import (BbPack, BollingerBandsBatchRange, CudaBollingerBands, CudaMediumAd,
                      CudaPercentileNearestRank, PercentileNearestRankBatchRange,
                      RollingManySeriesPlan)

__all__ = [
    "VtaError", "lib", "AlmaBatchRange", "AlmaPack", "AlmaParams", "AlmaSweepPlan", "CudaAlma", "HipAlma",
    "expand_grid", "DeviceArrayF32", "DeviceBuffer", "HipContext", "HipGraph", "cuda_available",
    "cuda_device_count", "CudaEma", "CudaRsi", "CudaAtr", "CudaAdx", "ManySeriesOhlcPlan",
    "ema_rsi_atr_many_series_dev", "BollingerBandsBatchRange", "CudaBollingerBands",
    "CudaMediumAd", "CudaPercentileNearestRank", "PercentileNearestRankBatchRange",
    "RollingManySeriesPlan", "DoubleAlmaCrossover", "DoubleAlmaPlan", "IndicatorServer",
    "ServedResult", "CudaSma", "CudaWma", "CudaSwma", "CudaSmma", "CudaWilders", "RecPack",
    "BbPack", "KernelPolicy",
]
