"""Writes golden vectors for the SPECT projector path from the reference's own
CPU build (oracle/_ref/lib_et_ref.so, compiled from the reference sources by
`make -C oracle ref`).  Run in the container that holds the reference sources:

    python tests/golden/make_golden.py

Each case is one .npz (inputs and reference outputs, float32) plus an entry
in manifest.json with the sha256 of every array.  The reference is not needed
to use the fixtures.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import refabi  # noqa: E402


def gaussian_psf(n, K, s0, s1):
    r = (K - 1) / 2.0
    t = np.arange(K) - r
    out = np.empty((n, K, K), np.float64)
    for d in range(n):
        s = s0 + (s1 - s0) * d / max(n - 1, 1)
        g = np.exp(-0.5 * (t / s) ** 2)
        out[d] = np.outer(g, g) / np.outer(g, g).sum()
    return out.astype(np.float32)


def volume(shape, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.ogrid[:nz, :ny, :nx]
    body = (((x - (nx - 1) / 2) / (0.4 * nx)) ** 2 + ((y - (ny - 1) / 2) / (0.45 * ny)) ** 2 +
            ((z - (nz - 1) / 2) / (0.4 * nz)) ** 2) <= 1
    act = (body * (1.0 + rng.random(shape)) + 0.05 * rng.random(shape)).astype(np.float32)
    mu = (body * 0.08 + 0.01 * rng.random(shape)).astype(np.float32)
    return act, mu


CASES = [
    # name, shape (nz, ny, nx), views, camera kind, psf K, mu, bg, bga
    ("axial_psf5_mu", (16, 12, 16), 7, "axial", 5, True, 0.0, 0.0),
    ("axial_psf3_nomu", (16, 12, 16), 7, "axial", 3, False, 0.0, 0.0),
    ("axial_nopsf_mu", (16, 12, 16), 7, "axial", 0, True, 0.0, 0.0),
    ("axial_nopsf_nomu", (16, 12, 16), 5, "axial", 0, False, 0.0, 0.0),
    ("axial_bg", (16, 12, 16), 7, "axial", 5, True, 0.5, 0.02),
    ("threeaxis_psf3_mu", (24, 16, 24), 6, "three", 3, True, 0.0, 0.0),
    ("axial_ragged", (21, 13, 21), 5, "axial", 3, True, 0.0, 0.0),
    ("mlem_32", (32, 32, 32), 16, "axial", 5, True, 0.0, 0.0),
]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    R = refabi.load_ref()
    manifest = {"source": "reference CPU build (oracle/_ref), SPECT_* / et_array_* array interface",
                "conventions": "volume (nz,ny,nx), sinogram (V,ny,nx), psf (nz,h,h), cameras (V,) or (3,V)",
                "cases": {}}
    for i, (name, shape, V, kind, K, use_mu, bg, bga) in enumerate(CASES):
        rng = np.random.default_rng(100 + i)
        act, mu = volume(shape, 100 + i)
        mu = mu if use_mu else None
        if kind == "axial":
            cams = rng.uniform(0, 2 * np.pi, V).astype(np.float32)
        else:
            cams = rng.uniform(-0.7, 0.7, (3, V)).astype(np.float32)
        psf = gaussian_psf(shape[2], K, 0.6, 1.5) if K else None
        arrays = {"activity": act, "cameras": cams}
        if mu is not None:
            arrays["attenuation"] = mu
        if psf is not None:
            arrays["psf"] = psf
        st, sino = refabi.project(R, act, mu, cams, psf, bg, bga, 1)
        assert st == 0
        arrays["projection"] = sino
        st, bp = refabi.backproject(R, sino, mu, cams, psf, bg, bga, 1)
        assert st == 0
        arrays["backprojection"] = bp
        if name.startswith("mlem"):
            y = np.random.default_rng(7).poisson(sino * 50.0).astype(np.float32)
            st, x1 = refabi.mlem_step(R, np.ones_like(act), y, mu, cams, psf, bg, bga, 1e-6)
            assert st == 0
            arrays["measured"] = y
            arrays["mlem_step"] = x1
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        manifest["cases"][name] = {"shape": list(shape), "views": V, "cameras": kind, "psf": K,
                                   "attenuation": use_mu, "background": bg, "background_attenuation": bga,
                                   "sha256": {k: sha(v) for k, v in arrays.items()}}
    # rotation resampler, all six axis orders, background 0.3
    rng = np.random.default_rng(42)
    img = rng.random((14, 10, 12), dtype=np.float32)
    ctr = np.array([5.5, 4.5, 6.5], np.float32)
    rot = {"image": img, "centers": ctr}
    for order in range(6):
        ang = rng.uniform(-1, 1, 3).astype(np.float32)
        st, out = refabi.rotate(R, img, ang, ctr, 0.3, order)
        assert st == 0
        rot["angles_%d" % order] = ang
        rot["rotated_%d" % order] = out
    np.savez_compressed(os.path.join(HERE, "rotate.npz"), **rot)
    manifest["cases"]["rotate"] = {"background": 0.3, "sha256": {k: sha(v) for k, v in rot.items()}}
    fisher_cases(R, manifest)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", len(manifest["cases"]), "cases")


def grid_volume(shape, n_points, n, seed):
    """Grid volume (nz, ny, nx): n_points voxels hold distinct 1-based indices
    drawn from 1..n (indices left out keep the reference's (-1,-1,-1) point)."""
    rng = np.random.default_rng(seed)
    g = np.zeros(shape, np.float32)
    vox = rng.choice(g.size, n_points, replace=False)
    g.ravel()[vox] = rng.choice(np.arange(1, n + 1), n_points, replace=False).astype(np.float32)
    return g


def fisher_cases(R, manifest):
    """et_array_fisher_grid (from an activity: projection, then the pair sums)
    and et_array_fisher_grid_projection (from a given projection, cameras about
    three axes, a 5 x 3 PSF), both with the CPU prior; and the status of a grid
    with a repeated index."""
    act, mu = volume((16, 16, 16), 300)
    rng = np.random.default_rng(300)
    cams = rng.uniform(0, 2 * np.pi, 6).astype(np.float32)
    psf = gaussian_psf(16, 5, 0.6, 1.5)
    grid = grid_volume((16, 16, 16), 30, 30, 301)
    st, F, P = refabi.fisher_grid(R, act, mu, grid, cams, psf, n=30, gpu=0)
    assert st == 0
    a = {"activity": act, "attenuation": mu, "cameras": cams, "psf": psf, "grid": grid, "fisher": F, "prior": P}
    np.savez_compressed(os.path.join(HERE, "fisher_act.npz"), **a)
    manifest["cases"]["fisher_act"] = {"n": 30, "gpu": 0, "sha256": {k: sha(v) for k, v in a.items()}}

    rng = np.random.default_rng(310)
    sino = (0.5 + rng.random((5, 20, 20))).astype(np.float32)
    cams3 = rng.uniform(-0.5, 0.5, (3, 5)).astype(np.float32)
    t = np.arange(5) - 2.0
    u = np.arange(3) - 1.0
    psf53 = np.empty((20, 3, 5), np.float32)  # (planes, py, px): px = 5, py = 3
    for d in range(20):
        s = 0.7 + 0.05 * d
        psf53[d] = np.outer(np.exp(-0.5 * (u / s) ** 2), np.exp(-0.5 * (t / s) ** 2))
    grid = grid_volume((20, 20, 20), 40, 45, 311)
    st, F, P = refabi.fisher_grid_projection(R, sino, grid, cams3, psf53, n=45, gpu=0)
    assert st == 0
    b = {"projection": sino, "cameras": cams3, "psf": psf53, "grid": grid, "fisher": F, "prior": P}
    np.savez_compressed(os.path.join(HERE, "fisher_proj.npz"), **b)
    manifest["cases"]["fisher_proj"] = {"n": 45, "gpu": 0, "sha256": {k: sha(v) for k, v in b.items()}}

    bad = grid_volume((16, 16, 16), 10, 10, 320)
    bad.ravel()[np.flatnonzero(bad)[:2]] = 3.0  # two voxels with index 3
    st, _, _ = refabi.fisher_grid(R, act, mu, bad, cams, psf, n=10, gpu=0)
    manifest["fisher_repeated_index_status"] = int(st)


if __name__ == "__main__":
    if sys.argv[1:] == ["fisher"]:  # add the Fisher-grid cases to an existing manifest
        with open(os.path.join(HERE, "manifest.json")) as f:
            m = json.load(f)
        fisher_cases(refabi.load_ref(), m)
        with open(os.path.join(HERE, "manifest.json"), "w") as f:
            json.dump(m, f, indent=1, sort_keys=True)
        print("wrote the Fisher-grid cases")
    else:
        main()
