"""Context and device-array handles over the C ABI.

``HipContext`` plays the role of the reference wrappers' cust Context + Stream
(src/cuda/moving_averages/alma_wrapper.rs:119-169); ``DeviceArrayF32`` is the
device-resident result ``{buf, rows, cols}`` (alma_wrapper.rs:103-117).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import DeviceArrayF32C, VtaError, check, lib


def require_len(buf, count: int, what: str, itemsize: int = 4) -> None:
    """Length check of a caller-owned device buffer, as the reference wrappers
    make on their DeviceBuffer arguments (ema_wrapper.rs:282-296, 400-409): a
    buffer shorter than what the launch reads or writes is InvalidInput, raised
    before anything is enqueued.  ``None`` and objects that carry no byte size
    (raw pointers) pass unchecked; the C ABI itself only sees pointers."""
    if buf is None:
        return
    nbytes = getattr(buf, "nbytes", None)
    if nbytes is not None and int(nbytes) < int(count) * itemsize:
        raise VtaError(4, f"{what} buffer length mismatch: it holds {int(nbytes) // itemsize} "
                          f"elements, the launch needs {int(count)}")


def cuda_available() -> bool:
    """Device probe (src/cuda/mod.rs:412-582); VTA_FORCE_SKIP=1 forces False."""
    try:
        return bool(lib().vta_device_available())
    except (ImportError, OSError):
        return False


def cuda_device_count() -> int:
    return int(lib().vta_device_count())


class HipContext:
    """One device, one non-blocking stream, lazily loaded gfx950 code objects."""

    def __init__(self, device_id: int = 0):
        self._ptr = ctypes.c_void_p()
        check(lib().vta_ctx_create(int(device_id), ctypes.byref(self._ptr)))
        self.device_id = int(device_id)

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._ptr:
            raise VtaError(4, "context is closed")
        return self._ptr

    @property
    def stream(self) -> int:
        return int(lib().vta_ctx_stream(self.handle) or 0)

    def synchronize(self) -> None:
        check(lib().vta_ctx_synchronize(self.handle))

    def preload(self) -> None:
        check(lib().vta_ctx_preload(self.handle))

    def close(self) -> None:
        if self._ptr:
            lib().vta_ctx_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # captured batches
    def begin_capture(self) -> None:
        """Starts recording the context's enqueue-only calls (see HipGraph)."""
        check(lib().vta_ctx_begin_capture(self.handle))

    def end_capture(self) -> "HipGraph":
        """Ends the recording; raises if a call made while recording could not
        be captured (the capture is ended either way)."""
        g = ctypes.c_void_p()
        check(lib().vta_ctx_end_capture(self.handle, ctypes.byref(g)))
        return HipGraph(self, g)

    # raw device memory for the enqueue-only entry points
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        p = ctypes.c_void_p()
        check(lib().vta_device_alloc(self.handle, int(nbytes), ctypes.byref(p)))
        return DeviceBuffer(self, p.value, int(nbytes))

    def to_device(self, host: np.ndarray) -> "DeviceBuffer":
        host = np.ascontiguousarray(host)
        buf = self.alloc(host.nbytes)
        check(lib().vta_memcpy_htod(self.handle, ctypes.c_void_p(buf.ptr),
                                    host.ctypes.data_as(ctypes.c_void_p), host.nbytes))
        return buf


class HipGraph:
    """A replayable batch: the enqueue-only calls made between
    ``ctx.begin_capture()`` and ``ctx.end_capture()``, launched again with one
    ``launch()`` on the context's stream.  It keeps the device pointers, sizes
    and parameters of the capture; keep those buffers alive while it lives."""

    def __init__(self, ctx: HipContext, ptr: ctypes.c_void_p):
        self.ctx = ctx
        self._ptr = ptr

    @property
    def num_nodes(self) -> int:
        return int(lib().vta_graph_num_nodes(self._ptr)) if self._ptr else 0

    def launch(self) -> None:
        if not self._ptr:
            raise VtaError(4, "graph is destroyed")
        check(lib().vta_graph_launch(self.ctx.handle, self._ptr))

    def destroy(self) -> None:
        if self._ptr:
            lib().vta_graph_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DeviceBuffer:
    """Caller-owned device allocation (cust DeviceBuffer)."""

    def __init__(self, ctx: HipContext, ptr: int, nbytes: int):
        self.ctx = ctx
        self.ptr = ptr
        self.nbytes = nbytes

    def copy_to_host(self, dtype=np.float32) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        check(lib().vta_memcpy_dtoh(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_void_p(self.ptr), self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().vta_device_free(self.ctx.handle, ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            if self.ctx._ptr:
                self.free()
        except Exception:
            pass


class DeviceArrayF32:
    """Device-resident f32 matrix owned by the library; freed on close/GC."""

    def __init__(self, ctx: HipContext, c: DeviceArrayF32C):
        self.ctx = ctx
        self._c = c

    @property
    def rows(self) -> int:
        return int(self._c.rows)

    @property
    def cols(self) -> int:
        return int(self._c.cols)

    def __len__(self) -> int:
        return self.rows * self.cols

    def device_ptr(self) -> int:
        return int(self._c.ptr or 0)

    def to_host(self) -> np.ndarray:
        out = np.empty((self.rows, self.cols), dtype=np.float32)
        check(lib().vta_device_array_copy_to_host(
            self.ctx.handle, ctypes.byref(self._c), out.ctypes.data_as(ctypes.c_void_p),
            out.size))
        return out

    def free(self) -> None:
        if self._c.ptr:
            lib().vta_device_array_free(self.ctx.handle, ctypes.byref(self._c))

    def __del__(self):
        try:
            if self.ctx._ptr:
                self.free()
        except Exception:
            pass

    # ---- device-handle wire formats (src/utilities/dlpack_cuda.rs:278-374)
    @property
    def __cuda_array_interface__(self) -> dict:
        """Version-3 array interface, the form torch-rocm and CuPy consume."""
        n = self.rows * self.cols
        return {"shape": (self.rows, self.cols), "typestr": "<f4",
                "strides": (self.cols * 4, 4), "data": (self.device_ptr() if n else 0, False),
                "version": 3}

    def __dlpack_device__(self):
        return (KDL_ROCM, self.ctx.device_id)

    def __dlpack__(self, stream=None, max_version=None, dl_device=None, copy=None):
        """Exports the buffer as a DLPack capsule ("dltensor_versioned" when the
        consumer accepts DLPack >= 1, else "dltensor").  Ownership moves into
        the capsule, as in the reference (this array is left empty); the data
        are complete before return, so any consumer stream may read them."""
        if dl_device is not None and tuple(dl_device) != self.__dlpack_device__():
            raise BufferError("__dlpack__: dl_device differs and copies are not implemented"
                              if copy else "__dlpack__: dl_device mismatch")
        versioned = max_version is not None and int(tuple(max_version)[0]) >= 1
        mt = ctypes.c_void_p()
        check(lib().vta_device_array_to_dlpack(self.ctx.handle, ctypes.byref(self._c),
                                               1 if versioned else 0, ctypes.byref(mt)))
        return _capsule(mt.value, versioned)


KDL_ROCM = 10
_CAP_NAMES = {False: b"dltensor", True: b"dltensor_versioned"}
_PyCapsule_Destructor = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


def _capsule_dtor(cap):
    """Frees the tensor of a capsule no consumer took (a consumer renames it
    to used_dltensor[_versioned] and owns the tensor from then on)."""
    api = ctypes.pythonapi
    for versioned, name in _CAP_NAMES.items():
        if api.PyCapsule_IsValid(ctypes.c_void_p(cap), name):
            p = api.PyCapsule_GetPointer(ctypes.c_void_p(cap), name)
            lib().vta_dlpack_delete(ctypes.c_void_p(p), 1 if versioned else 0)
            return


_CAP_DTOR = _PyCapsule_Destructor(_capsule_dtor)


def _capsule(ptr: int, versioned: bool):
    api = ctypes.pythonapi
    api.PyCapsule_New.restype = ctypes.py_object
    api.PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _PyCapsule_Destructor]
    api.PyCapsule_IsValid.restype = ctypes.c_int
    api.PyCapsule_IsValid.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    api.PyCapsule_GetPointer.restype = ctypes.c_void_p
    api.PyCapsule_GetPointer.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    return api.PyCapsule_New(ctypes.c_void_p(ptr), _CAP_NAMES[versioned], _CAP_DTOR)


def f32_ptr(a: np.ndarray) -> ctypes.c_void_p:
    return a.ctypes.data_as(ctypes.c_void_p)
