import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

try:
    # torch's ROCm wheel carries its own HIP runtime under the system one's
    # soname: it must load before the library (see niftyrec_amd._lib.load),
    # also in runs of a few test files that touch the library first
    import torch  # noqa: F401
except ImportError:
    pass


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


RTOL_L2 = 1e-5
RTOL_MAX = 1e-4


def report_parity(name, got, ref32, ref64=None, rtol_l2=RTOL_L2, rtol_max=RTOL_MAX):
    """Checks a result against the single-precision oracle (bit-identical to
    the reference CPU build) and prints its errors, with the double-precision
    oracle's where given; also appends them as one JSON line to the file named
    by $ET_PARITY_LOG.

    Gate: relative L2 <= rtol_l2 and max relative error (conftest.max_rel)
    <= rtol_max against the reference; where a double-precision result is
    given, the relative L2 against it at most twice the reference's own
    (+1e-7)."""
    import json
    e32, m32 = rel_l2(got, ref32), max_rel(got, ref32)
    rec = {"case": name, "relL2_vs_ref": e32, "maxrel_vs_ref": m32}
    msg = "%s: relL2 %.2e max_rel %.2e vs reference" % (name, e32, m32)
    if ref64 is not None:
        e64, m64 = rel_l2(got, ref64), max_rel(got, ref64)
        r64, rm64 = rel_l2(ref32, ref64), max_rel(ref32, ref64)
        rec.update(relL2_vs_fp64=e64, maxrel_vs_fp64=m64, ref_relL2_vs_fp64=r64, ref_maxrel_vs_fp64=rm64)
        msg += "; vs FP64 oracle: relL2 %.2e max_rel %.2e (reference itself: %.2e, %.2e)" % (e64, m64, r64, rm64)
    print(msg)
    path = os.environ.get("ET_PARITY_LOG")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps(rec) + "\n")
    assert e32 <= rtol_l2, msg
    assert m32 <= rtol_max, msg
    if ref64 is not None:
        assert rec["relL2_vs_fp64"] <= 2.0 * rec["ref_relL2_vs_fp64"] + 1e-7, msg
    return rec


@pytest.fixture(scope="session")
def lib():
    """The HIP library bound through the NiftyRec ABI driver (oracle.refabi)."""
    from niftyrec_amd import _lib
    from oracle import refabi
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("lib_et_array_interface.so is not built")
    return refabi.load(_lib.LIB_PATH)
