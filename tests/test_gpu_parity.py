"""Parity of the HIP library (called through the NiftyRec C ABI) with the
oracle, whose single-precision variant is bit-identical to the reference CPU
build (tests/test_oracle_reference.py).

Gate (conftest.report_parity, the project's accuracy target): relative L2
error <= 1e-5 and max relative error <= 1e-4 against the reference
restatement (elements below 1e-4 of the peak measured against 1e-4 of the
peak, conftest.max_rel); where a double-precision oracle result is given, the
relative L2 against it at most twice the reference's own.  Both errors are
printed.
"""
import ctypes

import numpy as np
import pytest

from conftest import RTOL_L2, max_rel, rel_l2, report_parity
from niftyrec_amd import phantoms
from oracle import orc, refabi

pytestmark = pytest.mark.gpu

def _case(shape=(32, 24, 32), V=16, K=5, seed=0, mu_scale=0.1):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    act, mu = phantoms.phantom(shape, 3, seed, lesion_axes=(1.5, 4.0))
    act = act + 0.1 * rng.random(shape, dtype=np.float32)
    mu = (mu / 0.15 * mu_scale).astype(np.float32)
    cams = rng.uniform(0, 2 * np.pi, V).astype(np.float32)
    psf = phantoms.gaussian_psf(nx, K, 0.5, 1.6) if K else None
    return act, mu, cams, psf


def _report(name, got, ref32, ref64=None):
    return report_parity(name, got, ref32, ref64)


@pytest.mark.parametrize("use_mu", [True, False])
@pytest.mark.parametrize("K", [0, 3, 5, 11])
def test_project_parity(lib, use_mu, K):
    act, mu, cams, psf = _case(K=K)
    m = mu if use_mu else None
    st, got = refabi.project(lib, act, m, cams, psf)
    assert st == 0
    _report("project K=%d mu=%d" % (K, use_mu), got, orc.project(act, m, cams, psf),
            orc.project(act, m, cams, psf, dtype=np.float64))


@pytest.mark.parametrize("use_mu", [True, False])
@pytest.mark.parametrize("K", [0, 5, 11])
def test_backproject_parity(lib, use_mu, K):
    act, mu, cams, psf = _case(K=K, seed=1)
    m = mu if use_mu else None
    sino = orc.project(act, m, cams, psf)
    st, got = refabi.backproject(lib, sino, m, cams, psf)
    assert st == 0
    _report("backproject K=%d mu=%d" % (K, use_mu), got, orc.backproject(sino, m, cams, psf),
            orc.backproject(sino, m, cams, psf, dtype=np.float64))


def test_project_odd_shapes(lib):
    """Sizes that are not multiples of any block or slab size."""
    for shape in [(21, 13, 21), (40, 37, 40), (7, 5, 7)]:
        act, mu, cams, psf = _case(shape=shape, V=5, K=3, seed=shape[1])
        st, got = refabi.project(lib, act, mu, cams, psf)
        assert st == 0
        _report("project %s" % (shape,), got, orc.project(act, mu, cams, psf),
                orc.project(act, mu, cams, psf, dtype=np.float64))
        st, bp = refabi.backproject(lib, got, mu, cams, psf)
        assert st == 0
        _report("backproject %s" % (shape,), bp, orc.backproject(got, mu, cams, psf),
                orc.backproject(got, mu, cams, psf, dtype=np.float64))


def test_background_nonzero(lib):
    """bg != 0 reproduces the reference resampler's background rule."""
    act, mu, cams, psf = _case(seed=2)
    st, got = refabi.project(lib, act, mu, cams, psf, bg=0.5, bga=0.02)
    assert st == 0
    _report("project bg", got, orc.project(act, mu, cams, psf, 0.5, 0.02),
            orc.project(act, mu, cams, psf, 0.5, 0.02, dtype=np.float64))
    st, bp = refabi.backproject(lib, got, mu, cams, psf, bg=0.5, bga=0.02)
    assert st == 0
    _report("backproject bg", bp, orc.backproject(got, mu, cams, psf, 0.5, 0.02),
            orc.backproject(got, mu, cams, psf, 0.5, 0.02, dtype=np.float64))


def test_three_axis_cameras(lib):
    """General rotations (3-column cameras) take the materialising path."""
    act, mu, cams, psf = _case(seed=3, V=6)
    rng = np.random.default_rng(3)
    c3 = rng.uniform(-0.6, 0.6, (3, 6)).astype(np.float32)
    st, got = refabi.project(lib, act, mu, c3, psf)
    assert st == 0
    _report("project 3-axis", got, orc.project(act, mu, c3, psf), orc.project(act, mu, c3, psf, dtype=np.float64))
    st, bp = refabi.backproject(lib, got, mu, c3, psf)
    assert st == 0
    _report("backproject 3-axis", bp, orc.backproject(got, mu, c3, psf),
            orc.backproject(got, mu, c3, psf, dtype=np.float64))


def test_wide_psf_general_path(lib):
    """PSF radius above the sweep kernels' limit (K = 25)."""
    act, mu, cams, psf = _case(shape=(32, 16, 32), V=4, K=25, seed=4)
    st, got = refabi.project(lib, act, mu, cams, psf)
    assert st == 0
    _report("project K=25", got, orc.project(act, mu, cams, psf), orc.project(act, mu, cams, psf, dtype=np.float64))
    st, bp = refabi.backproject(lib, got, mu, cams, psf)
    assert st == 0
    _report("backproject K=25", bp, orc.backproject(got, mu, cams, psf),
            orc.backproject(got, mu, cams, psf, dtype=np.float64))


def test_attenuation_only_projection(lib):
    """Activity absent: p = bg exp(-sum mu) (the reference CUDA path's result)."""
    act, mu, cams, psf = _case(seed=5)
    st, got = refabi.project(lib, None, mu, cams, None, bg=2.0)
    assert st == 0
    _report("project mu-only", got, orc.project(None, mu, cams, None, 2.0),
            orc.project(None, mu, cams, None, 2.0, dtype=np.float64))


@pytest.mark.parametrize("K", [0, 5])
def test_gradient_reference(lib, K):
    act, mu, cams, psf = _case(K=K, seed=6)
    s = np.random.default_rng(6).random((len(cams),) + act.shape[1:], dtype=np.float32)
    st, got = refabi.gradient_attenuation(lib, s, act, mu, cams, psf)
    assert st == 0
    _report("gradient ref K=%d" % K, got, orc.gradient_ref(s, act, mu, cams, psf),
            orc.gradient_ref(s, act, mu, cams, psf, dtype=np.float64))


def test_mlem_step(lib):
    act, mu, cams, psf = _case(seed=7)
    y = orc.project(act, mu, cams, psf)
    x0 = np.ones_like(act)
    st, got = refabi.mlem_step(lib, x0, y, mu, cams, psf)
    assert st == 0
    _report("mlem step", got, orc.mlem_step(x0, y, mu, cams, psf),
            orc.mlem_step(x0, y, mu, cams, psf, dtype=np.float64))


@pytest.mark.parametrize("bg, three_axis", [(0.3, False), (0.0, True)])
def test_mlem_step_general_path(lib, bg, three_axis):
    """MLEM step on the materialising path: background != 0 (the reference
    resampler's background rule) or rotations about all three axes."""
    act, mu, cams, psf = _case(seed=17, V=8)
    if three_axis:
        cams = np.random.default_rng(17).uniform(-0.5, 0.5, (3, 8)).astype(np.float32)
    y = orc.project(act, mu, cams, psf, bg=bg)
    x0 = np.ones_like(act)
    st, got = refabi.mlem_step(lib, x0, y, mu, cams, psf, bg=bg)
    assert st == 0
    _report("mlem step (bg %.1f, three-axis %s)" % (bg, three_axis), got,
            orc.mlem_step(x0, y, mu, cams, psf, bg=bg), orc.mlem_step(x0, y, mu, cams, psf, bg=bg, dtype=np.float64))


def test_osem_step_random_subset(lib):
    """et_array_osem_step draws its subset from glibc rand(): seed it, replay the
    draw in the oracle and compare with the MLEM update on that subset."""
    act, mu, cams, psf = _case(seed=8)
    y = orc.project(act, mu, cams, psf)
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1234)
    st, got = refabi.osem_step(lib, np.ones_like(act), y, mu, cams, 4, psf)
    assert st == 0
    libc.srand(1234)
    idx = orc.osem_subset(len(cams), len(cams) // 4)
    ref = orc.mlem_step(np.ones_like(act), y[idx], mu, cams[idx], psf)
    _report("osem step", got, ref, orc.mlem_step(np.ones(act.shape), y[idx], mu, cams[idx], psf, dtype=np.float64))


def test_rotate(lib):
    rng = np.random.default_rng(9)
    img = rng.random((20, 14, 18), dtype=np.float32)
    ctr = np.array([8.5, 6.5, 9.5], np.float32)
    for order in range(6):
        ang = rng.uniform(-1, 1, 3).astype(np.float32)
        st, got = refabi.rotate(lib, img, ang, ctr, 0.25, order)
        assert st == 0
        _report("rotate order %d" % order, got, orc.resample(img, orc.rotation_matrix(ang, ctr, order), 0.25),
                orc.resample(img, orc.rotation_matrix(ang, ctr, order, dtype=np.float64), 0.25, dtype=np.float64))


def test_affine(lib):
    """et_array_affine: resampling through a caller-given matrix (sample
    position = M (x, y, z, 1), first three rows, row-major)."""
    rng = np.random.default_rng(12)
    img = rng.random((14, 10, 18), dtype=np.float32)
    M = orc.rotation_matrix(rng.uniform(-0.7, 0.7, 3).astype(np.float32), np.array([8.0, 4.5, 6.0], np.float32), 3)
    M = np.ascontiguousarray(M, np.float32)
    M[:3, 3] += np.array([0.37, -0.61, 0.25], np.float32)
    out = np.zeros_like(img)
    sz = np.array(img.shape[::-1], np.int32)
    asz = np.array([4, 4], np.int32)
    F, I = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    fn = lib.et_array_affine
    fn.argtypes = [F, I, F, I, F, I, ctypes.c_float, ctypes.c_int]
    st = fn(img.ctypes.data_as(F), sz.ctypes.data_as(I), out.ctypes.data_as(F), sz.ctypes.data_as(I),
            M.ctypes.data_as(F), asz.ctypes.data_as(I), 0.5, 1)
    assert st == 0
    _report("affine", out, orc.resample(img, M, 0.5), orc.resample(img, M, 0.5, dtype=np.float64))


def test_convolve(lib):
    """et_array_convolve: per-plane separated PSF convolution (central row, then
    central column over the centre value), zero border."""
    rng = np.random.default_rng(13)
    img = rng.random((9, 20, 17), dtype=np.float32)
    psf = phantoms.gaussian_psf(9, 7, 0.6, 2.2)
    out = np.zeros_like(img)
    sz = np.array(img.shape[::-1], np.int32)
    psz = np.array(psf.shape[::-1], np.int32)
    F, I = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    fn = lib.et_array_convolve
    fn.argtypes = [F, I, F, I, F, I, ctypes.c_int]
    st = fn(img.ctypes.data_as(F), sz.ctypes.data_as(I), out.ctypes.data_as(F), sz.ctypes.data_as(I),
            np.ascontiguousarray(psf).ctypes.data_as(F), psz.ctypes.data_as(I), 1)
    assert st == 0
    _report("convolve", out, orc.convolve_planes(img, psf), orc.convolve_planes(img, psf, dtype=np.float64))


def test_spect64_full_size(lib):
    """64^3 volume, 64 views, PSF 7, attenuation."""
    w = phantoms.workload("spect64")
    act, mu, psf, cams = w["activity"], w["attenuation"], w["psf"], w["cameras"]
    st, got = refabi.project(lib, act, mu, cams, psf)
    assert st == 0
    _report("spect64 project", got, orc.project(act, mu, cams, psf), orc.project(act, mu, cams, psf, dtype=np.float64))
    st, bp = refabi.backproject(lib, got, mu, cams, psf)
    assert st == 0
    _report("spect64 backproject", bp, orc.backproject(got, mu, cams, psf),
            orc.backproject(got, mu, cams, psf, dtype=np.float64))


def test_spect256_full_size_views():
    """256^3 volume, PSF 21 (radius 10, the widest the sweep kernels take),
    attenuation 0.15/0.05/0.30, on 4 of its 180 views (the oracle's cost),
    through the device plan: projection, reference backprojection and exact
    attenuation gradient against the oracle."""
    torch = pytest.importorskip("torch")
    from niftyrec_amd.api import Plan
    w = phantoms.workload("spect256")
    act, mu, psf = w["activity"], w["attenuation"], w["psf"]
    cams = w["cameras"][[0, 37, 90, 151]]
    plan = Plan(act.shape, cams, psf)
    A, M = torch.from_numpy(act).cuda(), torch.from_numpy(mu).cuda()
    p = plan.project(A, M).cpu().numpy()
    _report("spect256 project (4 views)", p, orc.project(act, mu, cams, psf),
            orc.project(act, mu, cams, psf, dtype=np.float64))
    b = plan.backproject(torch.from_numpy(p).cuda(), M).cpu().numpy()
    _report("spect256 backproject (4 views)", b, orc.backproject(p, mu, cams, psf),
            orc.backproject(p, mu, cams, psf, dtype=np.float64))
    s = np.random.default_rng(5).random(p.shape, dtype=np.float32)
    g = plan.gradient_attenuation(torch.from_numpy(s).cuda(), A, M, exact=True, truncate=False).cpu().numpy()
    ref = orc.gradient_mu(s, act, mu, cams, psf)
    e = rel_l2(g, ref)
    print("spect256 exact mu gradient (4 views): relL2 %.2e vs oracle" % e)
    assert e <= RTOL_L2


@pytest.mark.parametrize("rotate", [1, 0])
@pytest.mark.parametrize("K, use_mu, bg", [(5, True, 0.0), (0, True, 0.3), (3, False, 0.2)])
def test_project_partial(rotate, K, use_mu, bg):
    """et_array_project_partial against the oracle's restatement (the reference
    has only a CUDA implementation, whose texture filtering is not pinned)."""
    from niftyrec_amd import api
    act, mu, cams, psf = _case(shape=(20, 12, 20), V=5, K=K, seed=30)
    mu = mu if use_mu else None
    img = (1.0 + np.random.default_rng(31).random((12, 20))).astype(np.float32) if bg else None
    sino, part = api.project_partial(act, cams, psf, mu, background=bg, background_image=img, rotate_partial=rotate)
    rs, rp = orc.project_partial(act, mu, cams, psf, bg=bg, bgimg=img, rotate_partial=rotate)
    rs64, rp64 = orc.project_partial(act, mu, cams, psf, bg=bg, bgimg=img, rotate_partial=rotate, dtype=np.float64)
    _report("partial sinogram K=%d mu=%s rotate=%d" % (K, use_mu, rotate), sino, rs, rs64)
    _report("partial sums K=%d mu=%s rotate=%d" % (K, use_mu, rotate), part, rp, rp64)


def test_project_partial_attenuation_only():
    from niftyrec_amd import api
    _, mu, cams, _ = _case(shape=(16, 10, 16), V=3, K=0, seed=32)
    sino, part = api.project_partial(None, cams, None, mu, background=1.5)
    rs, rp = orc.project_partial(None, mu, cams, None, bg=1.5)
    rs64, rp64 = orc.project_partial(None, mu, cams, None, bg=1.5, dtype=np.float64)
    _report("partial (attenuation only) sinogram", sino, rs, rs64)
    _report("partial (attenuation only) sums", part, rp, rp64)
