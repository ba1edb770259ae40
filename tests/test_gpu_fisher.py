"""Fisher information of a grid of points on the GPU (et_array_fisher_grid,
et_array_fisher_grid_projection) against the golden vectors of the reference
CPU build and the oracle: from a given projection the pair sums are the
reference's float loop in its order, so the matrices are compared bit for bit;
from an activity the projection carries the projector's (parity-bounded)
differences into the matrix."""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import report_parity
from niftyrec_amd import _lib, api, phantoms
from oracle import orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))


def _load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    return {k: d[k] for k in d.files}


def test_fisher_projection_golden_bit_exact():
    g = _load("fisher_proj")
    F, P = api.fisher_grid_projection(g["projection"], g["grid"], g["cameras"], g["psf"], n=45,
                                      gpu_prior_rule=False)
    print("fisher from a projection (3-axis cameras, 5x3 PSF): max |F - F_ref| %.3e" %
          np.abs(F - g["fisher"]).max())
    assert np.array_equal(F, g["fisher"])
    assert np.array_equal(P, g["prior"])


def test_fisher_activity_golden():
    g = _load("fisher_act")
    F, P = api.fisher_grid(g["activity"], g["grid"], g["cameras"], g["psf"], n=30, attenuation=g["attenuation"],
                           gpu_prior_rule=False)
    _, F64, _ = orc.fisher_grid_activity(g["activity"].astype(np.float64), g["attenuation"], g["grid"],
                                         g["cameras"], g["psf"], n=30, dtype=np.float64)
    report_parity("fisher grid from an activity (16^3, 6 views, PSF 5)", F, g["fisher"], F64)
    assert np.array_equal(P, g["prior"])


def test_fisher_prior_rules_and_flags():
    """GPU != 0: the reference CUDA prior rule; GPU == 0: its CPU rule; the
    Fisher matrix is the same for both."""
    g = _load("fisher_proj")
    F0, P0 = api.fisher_grid_projection(g["projection"], g["grid"], g["cameras"], g["psf"], n=45,
                                        gpu_prior_rule=False)
    F1, P1 = api.fisher_grid_projection(g["projection"], g["grid"], g["cameras"], g["psf"], n=45,
                                        gpu_prior_rule=True)
    _, _, P3 = orc.fisher_grid(g["projection"], g["grid"], g["cameras"], g["psf"], n=45, prior_rule=3)
    assert np.array_equal(F0, F1)
    assert np.array_equal(P1, P3)
    assert not np.array_equal(P0, P1)
    F, P = api.fisher_grid_projection(g["projection"], g["grid"], g["cameras"], g["psf"], n=45, prior=False)
    assert P is None and np.array_equal(F, F0)


def test_fisher_larger_grid():
    """64^3, 60 views, PSF 11, 600 points: from the oracle's projection bit for
    bit against the oracle's pair sums; from the activity within the parity
    bound of the reference-order (float32) oracle, and against FP64."""
    shape, V = (64, 64, 64), 60
    rng = np.random.default_rng(90)
    act, mu = phantoms.phantom(shape, 3, 90)
    act = act + 0.1 * rng.random(shape, dtype=np.float32)
    mu = (mu / 0.15 * 0.1).astype(np.float32)
    cams = np.linspace(0, 2 * np.pi, V, endpoint=False).astype(np.float32)
    psf = phantoms.gaussian_psf(64, 11, 0.8, 2.0)
    grid = np.zeros(shape, np.float32)
    vox = rng.choice(grid.size, 600, replace=False)
    grid.ravel()[vox] = np.arange(1, 601, dtype=np.float32)
    sino = orc.project(act, mu, cams, psf)
    F, _ = api.fisher_grid_projection(sino, grid, cams, psf, prior=False, gpu_prior_rule=False)
    _, Fo, _ = orc.fisher_grid(sino, grid, cams, psf, prior_rule=0)
    print("fisher 600 points from a projection: nonzero %d, bit-exact %s" % (np.count_nonzero(F),
                                                                            np.array_equal(F, Fo)))
    assert np.array_equal(F, Fo)
    Fa, _ = api.fisher_grid(act, grid, cams, psf, attenuation=mu, prior=False)
    _, F64, _ = orc.fisher_grid_activity(act.astype(np.float64), mu, grid, cams, psf, prior_rule=0,
                                         dtype=np.float64)
    report_parity("fisher grid from an activity (64^3, 60 views, PSF 11, 600 points)", Fa, Fo, F64)


def test_fisher_statuses():
    g = _load("fisher_act")
    L = _lib.load()
    fp = _lib.fp
    act, grid, psf = g["activity"], g["grid"], g["psf"]
    cams = np.ascontiguousarray(g["cameras"])
    F = np.zeros((30, 30), np.float32)
    sz = (ctypes.c_int * 3)(16, 16, 16)

    def call(grid_arr, cams_size, psf_arr, psf_size, n=30):
        return L.et_array_fisher_grid(act.ctypes.data_as(fp), sz, cams.ctypes.data_as(fp), cams_size,
                                      None if psf_arr is None else psf_arr.ctypes.data_as(fp), psf_size,
                                      grid_arr.ctypes.data_as(fp), F.ctypes.data_as(fp), None,
                                      (ctypes.c_int * 2)(n, n), None, (ctypes.c_int * 3)(0, 0, 0), 0.0, 0.0, 0.0, 0)
    ok_cams = (ctypes.c_int * 2)(6, 1)
    psz = (ctypes.c_int * 3)(5, 5, 16)
    assert call(grid, ok_cams, psf, psz) == 0
    bad = np.zeros_like(grid)
    bad.ravel()[[5, 77, 300]] = [1.0, 3.0, 3.0]  # repeated index
    assert call(bad, ok_cams, psf, psz, n=3) == MANIFEST["fisher_repeated_index_status"] == 2
    assert call(grid, (ctypes.c_int * 2)(6, 2), psf, psz) == 1           # camera columns: the reference's check
    assert call(grid, ok_cams, psf, (ctypes.c_int * 3)(4, 5, 16)) == 1   # even PSF size: the reference's check
    assert call(grid, ok_cams, None, (ctypes.c_int * 3)(0, 0, 0)) == 2   # no PSF (the reference would crash)
