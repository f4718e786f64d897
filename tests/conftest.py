import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def fm():
    """The product library (ctypes mirror).  Fails loudly if it is not built."""
    import gpu_deconvolution_amd as fm
    fm.lib()
    return fm


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def relL2(a, b):
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
