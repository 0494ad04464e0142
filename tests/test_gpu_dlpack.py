"""Device-handle wire formats of result arrays (src/utilities/dlpack_cuda.rs):
DLPack capsules (legacy and versioned) and the version-3 array interface, read
back by a consumer written against the C ABI and by torch-rocm.

Bar: bit-identical values through every path; zero-copy (the consumer sees the
same device address); the producer array is left empty after export."""
import ctypes

import numpy as np
import pytest

import vectorta_amd as vta
from vectorta_amd import _lib, datagen
from tests.conftest import require_gpu
from tests.test_abi import _DLManaged, _DLManagedVersioned

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sma():
    require_gpu()
    return vta.CudaSma(0)


def _result(sma):
    x = datagen.gbm_close(10_000, 0.01, seed=301)
    arr = sma.sma_batch_dev(x, (5, 40, 5))
    return arr, arr.to_host()


@pytest.mark.parametrize("versioned", [False, True])
def test_capsule_consumed_through_the_c_abi(sma, versioned):
    arr, ref = _result(sma)
    ptr = arr.device_ptr()
    assert arr.__dlpack_device__() == (10, 0)
    cap = arr.__dlpack__(max_version=(1, 0) if versioned else None)
    assert arr.device_ptr() == 0 and arr.rows == 0
    api = ctypes.pythonapi
    name = b"dltensor_versioned" if versioned else b"dltensor"
    api.PyCapsule_GetPointer.restype = ctypes.c_void_p
    api.PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]
    api.PyCapsule_SetName.argtypes = [ctypes.py_object, ctypes.c_char_p]
    mt = api.PyCapsule_GetPointer(cap, name)
    m = ctypes.cast(mt, ctypes.POINTER(_DLManagedVersioned if versioned else _DLManaged))[0]
    d = m.dl_tensor
    assert d.data == ptr and (d.shape[0], d.shape[1]) == ref.shape
    out = np.empty(ref.shape, np.float32)
    _lib.check(_lib.lib().vta_memcpy_dtoh(sma.ctx.handle, out.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.c_void_p(d.data), out.nbytes))
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    # a consumer renames the capsule and runs the deleter when done
    api.PyCapsule_SetName(cap, b"used_" + name)
    _lib.lib().vta_dlpack_delete(ctypes.c_void_p(mt), 1 if versioned else 0)
    del cap


def test_unconsumed_capsule_frees_its_buffer(sma):
    for _ in range(3):
        arr, _ = _result(sma)
        cap = arr.__dlpack__()
        del cap                                   # capsule destructor runs the deleter


_TORCH_CONSUMER = r"""
import sys
import numpy as np
import torch                      # first: the package then binds to torch's HIP runtime
torch.cuda.init()
import vectorta_amd as vta
from vectorta_amd import datagen
x = datagen.gbm_close(10_000, 0.01, seed=301)
sma = vta.CudaSma(0)
arr = sma.sma_batch_dev(x, (5, 40, 5))
ref = arr.to_host()
ptr = arr.device_ptr()
t = torch.from_dlpack(arr)
assert t.device.type == "cuda" and t.dtype == torch.float32, (t.device, t.dtype)
assert tuple(t.shape) == ref.shape and t.data_ptr() == ptr, (t.shape, t.data_ptr(), ptr)
assert arr.device_ptr() == 0
got = (t * 1.0).cpu().numpy()     # a torch kernel reads the exported buffer in place
assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
del t
torch.cuda.synchronize()
print("torch-consumer-ok", flush=True)
"""


def test_torch_from_dlpack_zero_copy(sma):
    """torch-rocm takes the buffer without a copy.  One HIP runtime per process:
    torch is imported before the package (both runtime copies carry the soname
    libamdhip64.so.7, so the package then binds to torch's), which needs a
    fresh process."""
    import os
    import subprocess
    import sys
    pytest.importorskip("torch")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _TORCH_CONSUMER], cwd=root, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0 and "torch-consumer-ok" in r.stdout, (r.stdout[-2000:],
                                                                   r.stderr[-2000:])


def test_cuda_array_interface(sma):
    arr, ref = _result(sma)
    ai = arr.__cuda_array_interface__
    assert ai["shape"] == ref.shape and ai["typestr"] == "<f4" and ai["version"] == 3
    assert ai["strides"] == (ref.shape[1] * 4, 4) and ai["data"] == (arr.device_ptr(), False)
