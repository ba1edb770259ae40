import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def max_rel(a, b, floor=1e-4):
    """max |a-b| / max(|b|, floor*max|b|) over all elements.

    Elements below `floor` of the peak are measured against floor*peak: at
    1e-6 of the peak the reference's own single-precision result already
    differs from double precision by more than 1e-4 (C1 backprojection:
    1.4e-4), so a relative gate there would measure rounding noise."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.abs(b), floor * max(np.abs(b).max(), 1e-300))
    return float((np.abs(a - b) / scale).max())


@pytest.fixture(scope="session")
def lib():
    """The HIP library bound through the NiftyRec ABI driver (oracle.refabi)."""
    from niftyrec_amd import _lib
    from oracle import refabi
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("lib_et_array_interface.so is not built")
    return refabi.load(_lib.LIB_PATH)
