"""The single-precision oracle against the live reference CPU build
(oracle/_ref/lib_et_ref.so) at full workload shapes; skipped where the
reference build is absent.  Bit-exact equality is expected: the oracle keeps
the reference's order of operations."""
import os

import numpy as np
import pytest

from niftyrec_amd import phantoms
from oracle import orc, refabi

pytestmark = pytest.mark.skipif(not os.path.exists(refabi.REF_PATH), reason="reference build absent")


@pytest.fixture(scope="module")
def ref():
    return refabi.load_ref()


def test_spect64_projection_backprojection(ref):
    """Full size: 64^3, 64 views, PSF 7, attenuation."""
    w = phantoms.workload("spect64")
    a, m, p, c = w["activity"], w["attenuation"], w["psf"], w["cameras"]
    st, s_ref = refabi.project(ref, a, m, c, p)
    assert st == 0
    assert np.array_equal(orc.project(a, m, c, p), s_ref)
    st, b_ref = refabi.backproject(ref, s_ref, m, c, p)
    assert st == 0
    assert np.array_equal(orc.backproject(s_ref, m, c, p), b_ref)


def test_spect128_shapes_subset_of_views(ref):
    """128^3 volume, PSF 11, attenuation, on 6 of 120 views."""
    w = phantoms.workload("spect128")
    a, m, p = w["activity"], w["attenuation"], w["psf"]
    c = w["cameras"][::20]
    st, s_ref = refabi.project(ref, a, m, c, p)
    assert st == 0
    assert np.array_equal(orc.project(a, m, c, p), s_ref)


def test_mlem_step_bg(ref):
    w = phantoms.workload("small")
    a, m, p, c = w["activity"], w["attenuation"], w["psf"], w["cameras"]
    y = orc.project(a, m, c, p) * 3.0
    st, x_ref = refabi.mlem_step(ref, np.ones_like(a), y, m, c, p, 0.2, 0.01, 1e-6)
    assert st == 0
    assert np.array_equal(orc.mlem_step(np.ones_like(a), y, m, c, p, 0.2, 0.01, 1e-6), x_ref)


def test_osem_step_subset_draw(ref):
    """The oracle's rand()-driven subset draw matches the reference's osem_step."""
    import ctypes
    w = phantoms.workload("small")
    a, m, p, c = w["activity"], w["attenuation"], w["psf"], w["cameras"]
    y = orc.project(a, m, c, p)
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(77)
    st, x_ref = refabi.osem_step(ref, np.ones_like(a), y, m, c, 4, p)
    assert st == 0
    libc.srand(77)
    idx = orc.osem_subset(len(c), len(c) // 4)
    assert np.array_equal(orc.mlem_step(np.ones_like(a), y[idx], m, c[idx], p), x_ref)
