import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def candles():
    """Bitfinex 4h candles, the reference's known-answer fixture (tests/golden)."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "bitfinex_4h.npz"))
    return {k: d[k] for k in d.files}


def require_gpu():
    """GPU tests fail (not skip) without a device so a silent no-op cannot pass."""
    import vectorta_amd as vta
    if os.environ.get("VTA_FORCE_SKIP") == "1":
        pytest.skip("VTA_FORCE_SKIP=1")
    if not vta.cuda_available():
        pytest.fail("no HIP device visible to libvta_hip")
