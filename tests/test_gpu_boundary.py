"""The drop-in boundary on the GPU: the by-value et_array_project /
et_array_backproject spellings bound exactly as the reference Python package
binds them (Python/NiftyRec/NiftyRec/NiftyRec.py:83-84, 145-146, with its
numpy shape and camera conventions), the Python 3 module niftyrec_amd.NiftyRec
(et_project, et_backproject, et_osmapem_step, Reconstructor), the EM drivers
et_array_mlem / et_array_osem / et_array_osem_ordered against repeated oracle
updates, and the OSEM caches' handling of an attenuation map that changes."""
import ctypes
from ctypes import POINTER as P, c_float, c_int

import numpy as np
import pytest

from conftest import report_parity
from niftyrec_amd import _lib, phantoms
from oracle import orc

pytestmark = pytest.mark.gpu


def _case(shape=(32, 24, 32), V=12, K=5, seed=0):
    rng = np.random.default_rng(seed)
    act, mu = phantoms.phantom(shape, 3, seed, lesion_axes=(1.5, 4.0))
    act = act + 0.1 * rng.random(shape, dtype=np.float32)
    mu = (mu / 0.15 * 0.1).astype(np.float32)
    cams = rng.uniform(0, 2 * np.pi, V).astype(np.float32)
    psf = phantoms.gaussian_psf(shape[2], K, 0.5, 1.6) if K else None
    return act, mu, cams, psf


def _reference_binding():
    """A fresh handle on the library with the reference package's argtypes."""
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in ("et_array_project", "et_array_backproject"):
        f = getattr(L, name)
        f.restype = c_int
        f.argtypes = [P(c_float), P(c_int), P(c_float), P(c_int), P(c_float), P(c_int), P(c_float), P(c_int),
                      P(c_float), P(c_int), c_float, c_float, c_int, c_int]
    return L


def test_by_value_spellings_reference_conventions():
    """Arguments built the way the reference module builds them: sizes are the
    numpy shapes, cameras the transposed (3, V) float32 matrix (so its memory
    stays [axis][camera]), the projection a zeros((N, m, V)) array."""
    L = _reference_binding()
    act, mu, cams1, psf = _case(seed=40)
    cams = np.zeros((3, len(cams1)), np.float32)
    cams[1] = cams1
    fp = P(c_float)
    ip = P(c_int)
    camsT = cams.transpose()
    cams_size = np.asarray(camsT.shape, np.int32)
    act_size = np.asarray(act.shape, np.int32)
    psf = psf.T  # MATLAB order: shape (h, h, N), x-fastest memory
    psf_size = np.asarray(psf.shape, np.int32)
    mu_size = np.asarray(mu.shape, np.int32)
    proj = np.zeros([act_size[0], act_size[1], cams_size[0]], np.float32)
    proj_size = np.asarray(proj.shape, np.int32)
    st = L.et_array_project(act.ctypes.data_as(fp), act_size.ctypes.data_as(ip), proj.ctypes.data_as(fp),
                            proj_size.ctypes.data_as(ip), camsT.ctypes.data_as(fp), cams_size.ctypes.data_as(ip),
                            psf.ctypes.data_as(fp), psf_size.ctypes.data_as(ip), mu.ctypes.data_as(fp),
                            mu_size.ctypes.data_as(ip), np.float32(0.0), np.float32(0.0), np.int32(1), np.int32(1))
    assert st == 0
    V, (nz, ny, nx) = len(cams1), act.shape
    report_parity("et_array_project (reference binding)", proj.reshape(V, ny, nx),
                  orc.project(act, mu, cams, psf.T), orc.project(act, mu, cams, psf.T, dtype=np.float64))
    sino = np.ascontiguousarray(proj)
    bk = np.zeros([proj_size[0], proj_size[1], proj_size[0]], np.float32)
    bk_size = np.asarray(bk.shape, np.int32)
    st = L.et_array_backproject(sino.ctypes.data_as(fp), proj_size.ctypes.data_as(ip), bk.ctypes.data_as(fp),
                                bk_size.ctypes.data_as(ip), camsT.ctypes.data_as(fp), cams_size.ctypes.data_as(ip),
                                psf.ctypes.data_as(fp), psf_size.ctypes.data_as(ip), mu.ctypes.data_as(fp),
                                mu_size.ctypes.data_as(ip), np.float32(0.0), np.float32(0.0), np.int32(1),
                                np.int32(1))
    assert st == 0
    s3 = proj.reshape(V, ny, nx)
    report_parity("et_array_backproject (reference binding)", bk, orc.backproject(s3, mu, cams, psf.T),
                  orc.backproject(s3, mu, cams, psf.T, dtype=np.float64))
    # absent psf and attenuation: size arrays of zeros, null pointers
    zeros3 = np.zeros(3, np.int32)
    proj2 = np.zeros_like(proj)
    st = L.et_array_project(act.ctypes.data_as(fp), act_size.ctypes.data_as(ip), proj2.ctypes.data_as(fp),
                            proj_size.ctypes.data_as(ip), camsT.ctypes.data_as(fp), cams_size.ctypes.data_as(ip),
                            None, zeros3.ctypes.data_as(ip), None, zeros3.ctypes.data_as(ip), np.float32(0.0),
                            np.float32(0.0), np.int32(1), np.int32(1))
    assert st == 0
    report_parity("et_array_project (no psf, no attenuation)", proj2.reshape(V, ny, nx), orc.project(act, None, cams),
                  orc.project(act, None, cams, dtype=np.float64))


def test_niftyrec_module_project_backproject():
    """niftyrec_amd.NiftyRec.et_project / et_backproject with the reference
    module's array conventions (NiftyRec.py:86-112)."""
    from niftyrec_amd import NiftyRec as NR
    act, mu, cams1, psf = _case(seed=41)
    cams = np.zeros((3, len(cams1)), np.float32)
    cams[1] = cams1
    V, (nz, ny, nx) = len(cams1), act.shape
    proj = NR.et_project(act, cams, psf.T, mu)  # psf in MATLAB order, (h, h, N)
    assert proj is not None and proj.shape == (nz, ny, V)
    report_parity("NiftyRec.et_project", proj.reshape(V, ny, nx), orc.project(act, mu, cams, psf),
                  orc.project(act, mu, cams, psf, dtype=np.float64))
    bk = NR.et_backproject(proj, cams, psf.T, mu)
    assert bk is not None and bk.shape == (nz, ny, nx)
    s3 = proj.reshape(V, ny, nx)
    report_parity("NiftyRec.et_backproject", bk, orc.backproject(s3, mu, cams, psf),
                  orc.backproject(s3, mu, cams, psf, dtype=np.float64))
    assert NR.et_project(act, cams, np.ones((4, 4, nz), np.float32), mu) is None  # even PSF: status 2 -> None
    assert NR.et_project(act, cams, psf, mu) is None  # (N, h, h) read as size [N, h, h]: status 2 -> None
    n, info = NR.et_list_gpus()
    assert n >= 1 and len(info) == NR.SIZE_INFO * min(n, NR.MAX_DEVICES)


def _osmapem_oracle(subset_order, x, sino, cams, mu, psf, r, beta=0.0, prior=0.0, eps=1e-6, dtype=np.float32):
    """NiftyRec.py:224-246 with the oracle as projector (numpy steps in dtype)."""
    x = np.asarray(x, dtype)
    N1, N2, N3 = x.shape
    n_cam = cams.shape[1]
    n_sub = int(np.int32(np.round(n_cam / subset_order)))
    idx = np.int32(np.round(0.5 + (n_cam - 0.5) * r) - 1)
    cs = cams[:, idx]
    norm = orc.backproject(np.ones((n_sub, N2, N3), dtype), mu, cs, psf, dtype=dtype)
    norm = norm - beta * prior
    norm = norm * (norm > 0) + dtype(eps)
    proj = orc.project(x, mu, cs, psf, dtype=dtype)
    proj = proj * (proj > 0) + dtype(eps)
    upd = orc.backproject(np.asarray(sino, dtype)[idx] / proj, mu, cs, psf, dtype=dtype)
    upd = upd * (upd > 0)
    return x * upd / norm


def test_niftyrec_osmapem_step_and_reconstructor():
    """et_osmapem_step (one OSEM step on a random subset) and the threaded
    Reconstructor (two 'osem' steps) against the same steps with the oracle."""
    from niftyrec_amd import NiftyRec as NR
    act, mu, cams1, psf = _case(shape=(32, 32, 32), V=16, seed=42)
    cams = np.zeros((3, 16), np.float32)
    cams[1] = cams1
    y = orc.project(act, mu, cams, psf)  # [camera, y, x]
    x0 = np.ones_like(act)
    r = np.random.default_rng(43).random(4)
    got = NR.et_osmapem_step(4, x0, y, cams, mu, psf.T, 0.0, 0.0, 1, rng=np.random.default_rng(43))
    report_parity("NiftyRec.et_osmapem_step", got, _osmapem_oracle(4, x0, y, cams, mu, psf, r),
                  _osmapem_oracle(4, x0, y, cams, mu, psf, r, dtype=np.float64))

    rec = NR.Reconstructor(act.shape)
    rec.set_cameras_matrix(cams)
    rec.set_sinogram(y)
    rec.set_psf_matrix(psf.T)
    rec.set_attenuation(mu)
    np.random.seed(44)
    assert rec.reconstruct("osem", {"steps": 2, "subset_order": 4}, block=True)
    np.random.seed(44)
    draws = [np.random.rand(4) for _ in range(2)]
    x, x64 = np.ones_like(act), np.ones(act.shape)
    for r in draws:
        x = _osmapem_oracle(4, x, y, cams, mu, psf, r)
        x64 = _osmapem_oracle(4, x64, y, cams, mu, psf, r, dtype=np.float64)
    report_parity("NiftyRec.Reconstructor osem (2 steps)", rec.activity, x, x64)


@pytest.fixture(scope="module")
def spect64():
    w = phantoms.workload("spect64")
    from niftyrec_amd import api
    V = len(w["cameras"])
    cams3 = api.make_cameras(0.0, 360.0 * (1 - 1.0 / V), V, 1)  # inclusive endpoints
    y = orc.project(w["activity"], w["attenuation"], cams3, w["psf"])
    y = phantoms.poisson_data(y, w["counts"], seed=5)
    return w, cams3, y


def test_et_array_mlem_driver_spect64(spect64):
    """et_array_mlem: 10 MLEM iterations of the 64^3 / 64-view / PSF 7 workload
    from a uniform start in one device-resident call, against 10 oracle
    updates (et_array_make_cameras' inclusive endpoints: last = 360 (1 - 1/64))."""
    from niftyrec_amd import api
    w, cams3, y = spect64
    mu, psf = w["attenuation"], w["psf"]
    V = len(w["cameras"])
    got = api.mlem(y, 0.0, 360.0 * (1 - 1.0 / V), 10, psf=psf, attenuation=mu)
    x32, x64 = np.ones_like(mu), np.ones(mu.shape)
    for _ in range(10):
        x32 = orc.mlem_step(x32, y, mu, cams3, psf)
        x64 = orc.mlem_step(x64, y, mu, cams3, psf, dtype=np.float64)
    report_parity("et_array_mlem (64^3, 64 views, 10 iterations)", got, x32, x64)


def test_et_array_osem_random_driver(spect64):
    """et_array_osem: 3 steps on subsets drawn from glibc rand(), replayed in
    the oracle from the same seed."""
    from niftyrec_amd import api
    w, cams3, y = spect64
    mu, psf = w["attenuation"], w["psf"]
    V = len(w["cameras"])
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(77)
    got = api.osem(y, 4, 0.0, 360.0 * (1 - 1.0 / V), 3, psf=psf, attenuation=mu)
    libc.srand(77)
    x, x64 = np.ones_like(mu), np.ones(mu.shape)
    for _ in range(3):
        idx = orc.osem_subset(V, V // 4)
        x = orc.mlem_step(x, y[idx], mu, cams3[:, idx], psf)
        x64 = orc.mlem_step(x64, y[idx], mu, cams3[:, idx], psf, dtype=np.float64)
    report_parity("et_array_osem (3 random-subset steps)", got, x, x64)


def test_et_array_osem_ordered_driver(spect64):
    """et_array_osem_ordered: 2 passes over 4 ordered subsets."""
    from niftyrec_amd import api
    w, cams3, y = spect64
    mu, psf = w["attenuation"], w["psf"]
    V = len(w["cameras"])
    got = api.osem_ordered(y, 4, 0.0, 360.0 * (1 - 1.0 / V), 2, psf=psf, attenuation=mu)
    x32, x64 = np.ones_like(mu), np.ones(mu.shape)
    for _ in range(2):
        for s in range(4):
            x32 = orc.mlem_step(x32, y[s::4], mu, cams3[:, s::4], psf)
            x64 = orc.mlem_step(x64, y[s::4], mu, cams3[:, s::4], psf, dtype=np.float64)
    report_parity("et_array_osem_ordered (4 subsets, 2 passes)", got, x32, x64)


def test_driver_general_path_and_status():
    """Drivers off the fast path (rotation about the x axis) and the first
    non-zero status on bad parameters."""
    from niftyrec_amd import api
    act, mu, _, psf = _case(shape=(24, 16, 24), V=8, K=3, seed=45)
    cams3 = api.make_cameras(0.0, 315.0, 8, 0)
    y = orc.project(act, mu, cams3, psf)
    got = api.osem_ordered(y, 2, 0.0, 315.0, 1, psf=psf, attenuation=mu, rotation_axis=0)
    x, x64 = np.ones_like(act), np.ones(act.shape)
    for s in range(2):
        x = orc.mlem_step(x, y[s::2], mu, cams3[:, s::2], psf)
        x64 = orc.mlem_step(x64, y[s::2], mu, cams3[:, s::2], psf, dtype=np.float64)
    report_parity("et_array_osem_ordered about x (general path)", got, x, x64)
    L = _lib.load()
    a = np.zeros_like(act)
    st = L.et_array_mlem(a.ctypes.data_as(_lib.fp), 24, 16, np.ascontiguousarray(y).ctypes.data_as(_lib.fp), 8,
                         0.0, 315.0, 1, 3, 1, mu.ctypes.data_as(_lib.fp), 1,
                         np.ascontiguousarray(psf).ctypes.data_as(_lib.fp), 4, 4, 0.0, 0.0, 1e-6, 1)
    assert st == _lib.ERROR_PARAMETERS


def test_osem_cache_follows_the_attenuation_map():
    """Plan.osem after the map is rewritten in place, and with another map at
    the same address, matches the oracle with the new map (the caches are keyed
    on the map's content); with the address key, invalidate() is required."""
    torch = pytest.importorskip("torch")
    from niftyrec_amd.api import Plan
    act, mu, cams, psf = _case(shape=(32, 24, 32), V=12, K=5, seed=46)
    plan = Plan(act.shape, cams, psf)
    y = orc.project(act, mu, cams, psf)
    Y = torch.from_numpy(y).cuda()

    def run(M):
        X = torch.ones(act.shape, device="cuda")
        plan.osem(X, Y, M, subsets=3, iterations=1)
        return X.cpu().numpy()

    def oracle(m, dtype=np.float32):
        x = np.ones(act.shape, dtype)
        for s in range(3):
            x = orc.mlem_step(x, y[s::3], m, cams[s::3], psf, dtype=dtype)
        return x

    def refs(m):
        return oracle(m), oracle(m, np.float64)

    M = torch.from_numpy(mu).cuda()
    report_parity("osem, first map", run(M), *refs(mu))
    M.mul_(0.5)
    report_parity("osem, map halved in place", run(M), *refs(mu * np.float32(0.5)))
    addr = M.data_ptr()
    del M
    M2 = torch.empty(act.shape, device="cuda")
    print("second map at the first map's address: %s" % (M2.data_ptr() == addr))
    mu2 = (mu * np.float32(1.5)).astype(np.float32)
    M2.copy_(torch.from_numpy(mu2))
    report_parity("osem, another map (same address: %s)" % (M2.data_ptr() == addr), run(M2), *refs(mu2))
    plan.set_mu_check(False)
    run(M2)
    M2.mul_(0.25)
    plan.invalidate()
    report_parity("osem, address key + invalidate", run(M2), *refs(mu2 * np.float32(0.25)))
    # content key (the C library's default): a map rewritten without torch
    # knowing (through a second tensor on the same memory, whose version
    # counter is its own) is still caught
    pc = Plan(act.shape, cams, psf, mu_key="content")
    Mc = torch.from_numpy(mu).cuda()
    Xc = torch.ones(act.shape, device="cuda")
    pc.osem(Xc, Y, Mc, subsets=3, iterations=1)
    alias = Mc.detach()
    alias.data.mul_(2.0)
    Xc = torch.ones(act.shape, device="cuda")
    pc.osem(Xc, Y, Mc, subsets=3, iterations=1)
    report_parity("osem, content key, map rewritten behind torch", Xc.cpu().numpy(), *refs(mu * np.float32(2.0)))
