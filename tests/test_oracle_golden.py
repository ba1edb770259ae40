"""The single-precision oracle reproduces the golden vectors written by the
reference CPU build (tests/golden/make_golden.py) bit for bit."""
import json
import os

import numpy as np
import pytest

from oracle import orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))
CASES = sorted(k for k in MANIFEST["cases"] if k != "rotate" and not k.startswith("fisher"))


def _load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    return {k: d[k] for k in d.files}


@pytest.mark.parametrize("name", CASES)
def test_golden_projection(name):
    g = _load(name)
    meta = MANIFEST["cases"][name]
    got = orc.project(g["activity"], g.get("attenuation"), g["cameras"], g.get("psf"), meta["background"],
                      meta["background_attenuation"], 1)
    assert np.array_equal(got, g["projection"])


@pytest.mark.parametrize("name", CASES)
def test_golden_backprojection(name):
    g = _load(name)
    meta = MANIFEST["cases"][name]
    got = orc.backproject(g["projection"], g.get("attenuation"), g["cameras"], g.get("psf"), meta["background"],
                          meta["background_attenuation"], 1)
    assert np.array_equal(got, g["backprojection"])


def test_golden_mlem_step():
    g = _load("mlem_32")
    got = orc.mlem_step(np.ones_like(g["activity"]), g["measured"], g["attenuation"], g["cameras"], g["psf"],
                        eps=1e-6)
    assert np.array_equal(got, g["mlem_step"])


def test_golden_rotate():
    g = _load("rotate")
    for order in range(6):
        M = orc.rotation_matrix(g["angles_%d" % order], g["centers"], order)
        got = orc.resample(g["image"], M, 0.3)
        assert np.array_equal(got, g["rotated_%d" % order]), order


def test_manifest_hashes():
    import hashlib
    for name, meta in MANIFEST["cases"].items():
        g = _load(name)
        for k, h in meta["sha256"].items():
            assert hashlib.sha256(np.ascontiguousarray(g[k]).tobytes()).hexdigest() == h, (name, k)


def test_golden_fisher_grid_from_activity():
    """et_array_fisher_grid of the reference CPU build: projection, pair sums
    and the CPU prior rule, bit for bit."""
    g = _load("fisher_act")
    st, F, P = orc.fisher_grid_activity(g["activity"], g["attenuation"], g["grid"], g["cameras"], g["psf"], n=30,
                                        prior_rule=1)
    assert st == 0
    assert np.array_equal(F, g["fisher"])
    assert np.array_equal(P, g["prior"])


def test_golden_fisher_grid_from_projection():
    """et_array_fisher_grid_projection: cameras about three axes, a 5 x 3 PSF,
    grid indices left out (their points sit at (-1,-1,-1)), bit for bit."""
    g = _load("fisher_proj")
    st, F, P = orc.fisher_grid(g["projection"], g["grid"], g["cameras"], g["psf"], n=45, prior_rule=1)
    assert st == 0
    assert np.array_equal(F, g["fisher"])
    assert np.array_equal(P, g["prior"])
    assert np.count_nonzero(F) > 100


def test_fisher_repeated_index_status():
    g = _load("fisher_act")
    bad = np.zeros_like(g["grid"])
    bad.ravel()[[5, 77, 300]] = [1.0, 3.0, 3.0]
    st, _, _ = orc.fisher_grid_activity(g["activity"], g["attenuation"], bad, g["cameras"], g["psf"], n=3)
    assert st == MANIFEST["fisher_repeated_index_status"] == 2
