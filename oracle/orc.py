"""ctypes front-end of the C oracle (oracle/et_oracle.c).

TEST INFRASTRUCTURE ONLY.  Array conventions (numpy C order, x fastest, so
shapes read reversed relative to the C sizes):
  volume   (nz, ny, nx)       <-> C size [nx, ny, nz]
  sinogram (V, ny, nx)        <-> C size [nx, ny, V]
  psf      (nz, h, h)         <-> C size [h, h, nz]
  cameras  (3, V) radians, row = rotation axis
  centers  (3, V) voxels (default (size-1)/2, _et_array_interface.cpp:434-439)
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE, "oracle"])
        _LIB = ctypes.CDLL(path)
    return _LIB


def _ptr(a, dt):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(dt))


def _ctype(dtype):
    return ctypes.c_float if dtype == np.float32 else ctypes.c_double


def _suffix(dtype):
    return "f32" if dtype == np.float32 else "f64"


def expand_cameras(cameras, dtype=np.float32):
    """[V] or [V,1] angles about axis 1 -> (3,V); (3,V) passes through.
    Mirrors _et_array_interface.cpp:409-419."""
    c = np.asarray(cameras)
    if c.ndim == 1 or (c.ndim == 2 and c.shape[1] == 1 and c.shape[0] != 3):
        c = c.reshape(-1)
        out = np.zeros((3, c.size), dtype=dtype)
        out[1] = c
        return out
    return np.ascontiguousarray(c, dtype=dtype)


def default_centers(shape_zyx, V, dtype=np.float32):
    nz, ny, nx = shape_zyx
    c = np.empty((3, V), dtype=dtype)
    c[0] = (nx - 1) / 2.0
    c[1] = (ny - 1) / 2.0
    c[2] = (nz - 1) / 2.0
    return c


def _prep(dtype, *arrays):
    out = []
    for a in arrays:
        out.append(None if a is None else np.ascontiguousarray(a, dtype=dtype))
    return out


def rotation_matrix(angles, centers, order=0, dtype=np.float32):
    M = np.zeros(16, dtype=dtype)
    ct = _ctype(dtype)
    getattr(lib(), "orc_rotation_matrix_" + _suffix(dtype))(
        _ptr(M, ct), ct(angles[0]), ct(angles[1]), ct(angles[2]),
        ct(centers[0]), ct(centers[1]), ct(centers[2]), ctypes.c_int(order))
    return M.reshape(4, 4)


def resample(src, M, bg=0.0, dtype=np.float32):
    src, M = _prep(dtype, src, np.asarray(M).reshape(16))
    nz, ny, nx = src.shape
    out = np.empty_like(src)
    ct = _ctype(dtype)
    getattr(lib(), "orc_resample_" + _suffix(dtype))(
        _ptr(src, ct), nx, ny, nz, _ptr(M, ct), ct(bg), _ptr(out, ct))
    return out


def resample_adj(g, M, dtype=np.float64):
    g, M = _prep(dtype, g, np.asarray(M).reshape(16))
    nz, ny, nx = g.shape
    out = np.zeros_like(g)
    ct = _ctype(dtype)
    getattr(lib(), "orc_resample_adj_" + _suffix(dtype))(
        _ptr(g, ct), nx, ny, nz, _ptr(M, ct), _ptr(out, ct))
    return out


def convolve_planes(vol, psf, transpose=False, dtype=np.float32):
    vol, psf = _prep(dtype, vol, psf)
    vol = vol.copy()
    nz, ny, nx = vol.shape
    h = psf.shape[-1]
    sep = np.empty((nz, 2 * h), dtype=dtype)
    ct = _ctype(dtype)
    s = _suffix(dtype)
    getattr(lib(), "orc_psf_separate_" + s)(_ptr(psf, ct), h, nz, _ptr(sep, ct))
    getattr(lib(), "orc_convolve_planes_" + s)(_ptr(vol, ct), nx, ny, nz, _ptr(sep, ct), h,
                                               int(transpose))
    return vol


def _geom(shape_zyx, cameras, centers, dtype):
    cams = expand_cameras(cameras, dtype)
    V = cams.shape[1]
    if centers is None:
        centers = default_centers(shape_zyx, V, dtype)
    cams, centers = _prep(dtype, cams, centers)
    return cams, centers, V


def project(act, mu, cameras, psf=None, bg=0.0, bga=0.0, truncate=1, centers=None,
            dtype=np.float32):
    ref = act if act is not None else mu
    act, mu, psf = _prep(dtype, act, mu, psf)
    nz, ny, nx = ref.shape
    cams, centers, V = _geom(ref.shape, cameras, centers, dtype)
    sino = np.zeros((V, ny, nx), dtype=dtype)
    ct = _ctype(dtype)
    h = 0 if psf is None else psf.shape[-1]
    getattr(lib(), "orc_project_" + _suffix(dtype))(
        _ptr(act, ct), _ptr(mu, ct), nx, ny, nz, _ptr(cams, ct), _ptr(centers, ct), V,
        _ptr(psf, ct), h, ct(bg), ct(bga), int(truncate), _ptr(sino, ct))
    return sino


def backproject(sino, mu, cameras, psf=None, bg=0.0, bga=0.0, truncate=1, centers=None,
                mode=0, nz=None, dtype=np.float32):
    """mode 0: reference (inverse-rotation pull); mode 1: exact adjoint."""
    sino, mu, psf = _prep(dtype, sino, mu, psf)
    V, ny, nx = sino.shape
    nz = nx if nz is None else nz
    cams, centers, V2 = _geom((nz, ny, nx), cameras, centers, dtype)
    assert V2 == V
    out = np.zeros((nz, ny, nx), dtype=dtype)
    ct = _ctype(dtype)
    h = 0 if psf is None else psf.shape[-1]
    getattr(lib(), "orc_backproject_" + _suffix(dtype))(
        _ptr(sino, ct), _ptr(mu, ct), nx, ny, nz, _ptr(cams, ct), _ptr(centers, ct), V,
        _ptr(psf, ct), h, ct(bg), ct(bga), int(truncate), int(mode), _ptr(out, ct))
    return out


def gradient_ref(sino, act, mu, cameras, psf=None, bg=0.0, bga=0.0, truncate=1, centers=None,
                 dtype=np.float32):
    sino, act, mu, psf = _prep(dtype, sino, act, mu, psf)
    nz, ny, nx = act.shape
    cams, centers, V = _geom(act.shape, cameras, centers, dtype)
    out = np.zeros_like(act)
    ct = _ctype(dtype)
    h = 0 if psf is None else psf.shape[-1]
    getattr(lib(), "orc_gradient_ref_" + _suffix(dtype))(
        _ptr(sino, ct), _ptr(act, ct), _ptr(mu, ct), nx, ny, nz, _ptr(cams, ct),
        _ptr(centers, ct), V, _ptr(psf, ct), h, ct(bg), ct(bga), int(truncate), _ptr(out, ct))
    return out


def gradient_mu(sino, act, mu, cameras, psf=None, bg=0.0, bga=0.0, centers=None,
                dtype=np.float64):
    sino, act, mu, psf = _prep(dtype, sino, act, mu, psf)
    nz, ny, nx = mu.shape
    cams, centers, V = _geom(mu.shape, cameras, centers, dtype)
    out = np.zeros_like(mu)
    ct = _ctype(dtype)
    h = 0 if psf is None else psf.shape[-1]
    getattr(lib(), "orc_gradient_mu_" + _suffix(dtype))(
        _ptr(sino, ct), _ptr(act, ct), _ptr(mu, ct), nx, ny, nz, _ptr(cams, ct),
        _ptr(centers, ct), V, _ptr(psf, ct), h, ct(bg), ct(bga), _ptr(out, ct))
    return out


def mlem_step(act, sino, mu, cameras, psf=None, bg=0.0, bga=0.0, eps=1e-6, centers=None,
              dtype=np.float32):
    act, sino, mu, psf = _prep(dtype, act, sino, mu, psf)
    act = act.copy()
    nz, ny, nx = act.shape
    cams, centers, V = _geom(act.shape, cameras, centers, dtype)
    ct = _ctype(dtype)
    h = 0 if psf is None else psf.shape[-1]
    getattr(lib(), "orc_mlem_step_" + _suffix(dtype))(
        _ptr(act, ct), _ptr(sino, ct), _ptr(mu, ct), nx, ny, nz, _ptr(cams, ct),
        _ptr(centers, ct), V, _ptr(psf, ct), h, ct(bg), ct(bga), ct(eps))
    return act


def osem_subset(V, n_sub):
    idx = np.zeros(n_sub, dtype=np.int32)
    lib().orc_osem_subset(V, n_sub, _ptr(idx, ctypes.c_int))
    return idx


def project_partial(act, mu, cameras, psf=None, bg=0.0, bgimg=None, bga=0.0, truncate=1, rotate_partial=1,
                    centers=None, dtype=np.float32):
    """(sinogram (V, ny, nx), partial sums (V, nz, ny, nx)); see orc_project_partial."""
    ref = act if act is not None else mu
    act, mu, psf, bgimg = _prep(dtype, act, mu, psf, bgimg)
    nz, ny, nx = ref.shape
    cams, centers, V = _geom(ref.shape, cameras, centers, dtype)
    sino = np.zeros((V, ny, nx), dtype=dtype)
    part = np.zeros((V, nz, ny, nx), dtype=dtype)
    ct = _ctype(dtype)
    h = 0 if psf is None else psf.shape[-1]
    getattr(lib(), "orc_project_partial_" + _suffix(dtype))(
        _ptr(act, ct), _ptr(mu, ct), nx, ny, nz, _ptr(cams, ct), _ptr(centers, ct), V, _ptr(psf, ct), h,
        ct(bg), _ptr(bgimg, ct), ct(bga), int(truncate), int(rotate_partial), _ptr(sino, ct), _ptr(part, ct))
    return sino, part


def fisher_grid(sino, grid, cameras, psf, n=None, prior_rule=1, centers=None, dtype=np.float32):
    """Fisher information matrix (n, n) of the grid points of `grid` (nz, ny, nx)
    (0, or the 1-based point index) from the projection `sino` (V, sny, snx),
    and the prior matrix (prior_rule 1: CPU rule, 3: GPU rule; 0: None).
    psf (nz, py, px).  Returns (status, fisher, prior)."""
    sino, grid, psf = _prep(dtype, sino, grid, psf)
    nz, ny, nx = grid.shape
    V, sny, snx = sino.shape
    n = int(grid.max()) if n is None else int(n)
    cams, centers, V2 = _geom((nx, sino.shape[1], nx), cameras, centers, dtype)
    assert V2 == V
    fisher = np.zeros((n, n), dtype=dtype)
    prior = np.zeros((n, n), dtype=dtype) if prior_rule else None
    ct = _ctype(dtype)
    st = getattr(lib(), "orc_fisher_grid_" + _suffix(dtype))(
        _ptr(sino, ct), snx, sny, _ptr(grid, ct), nx, ny, nz, n, _ptr(cams, ct), _ptr(centers, ct), V,
        _ptr(psf, ct), psf.shape[2], psf.shape[1], int(prior_rule), _ptr(fisher, ct), _ptr(prior, ct))
    return st, fisher, prior


def fisher_grid_activity(act, mu, grid, cameras, psf, n=None, bg=0.0, bga=0.0, prior_rule=1, dtype=np.float32):
    """Fisher grid from an activity: the projection (truncated, as the reference
    computes it, emission-lib/_et.cpp:563) followed by fisher_grid.  Centres
    from the activity size."""
    sino = project(act, mu, cameras, psf, bg, bga, 1, None, dtype)
    nz, ny, nx = act.shape
    centers = default_centers(act.shape, sino.shape[0], dtype)
    return fisher_grid(sino, grid, cameras, psf, n, prior_rule, centers, dtype)
