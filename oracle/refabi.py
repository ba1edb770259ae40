"""Drives any shared library exporting NiftyRec's array interface (the
reference CPU build in oracle/_ref or the HIP library) with numpy arrays.

TEST INFRASTRUCTURE ONLY.  Conventions as oracle/orc.py: volumes (nz, ny, nx),
sinograms (V, ny, nx), psf (nz, h, h), cameras (V,) about axis 1 or (3, V).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PATH = os.path.join(HERE, "_ref", "lib_et_ref.so")

_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)


def load(path):
    lib = ctypes.CDLL(path)
    for name in ("SPECT_project_parallelholes", "SPECT_backproject_parallelholes"):
        getattr(lib, name).restype = ctypes.c_int
        getattr(lib, name).argtypes = [_F, _I, _F, _I, _F, _I, _F, _I, _F, _I, _F, _F, _I, _I]
    lib.et_array_gradient_attenuation.argtypes = [_F, _I, _F, _I, _F, _I, _F, _I, _F, _I, _F, _I, ctypes.c_float,
                                                  ctypes.c_float, ctypes.c_int, ctypes.c_int]
    lib.et_array_mlem_step.argtypes = [_F, ctypes.c_int, ctypes.c_int, _F, ctypes.c_int, _F, ctypes.c_int, _F,
                                       ctypes.c_int, _F, ctypes.c_uint, ctypes.c_uint, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_int]
    lib.et_array_osem_step.argtypes = [_F, ctypes.c_int, ctypes.c_int, ctypes.c_uint, _F, ctypes.c_int, _F,
                                       ctypes.c_int, _F, ctypes.c_int, _F, ctypes.c_uint, ctypes.c_uint,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    lib.et_array_rotate.argtypes = [_F, _I, _F, _F, _F, ctypes.c_float, ctypes.c_int, ctypes.c_int]
    return lib


def load_ref():
    if not os.path.exists(REF_PATH):
        raise FileNotFoundError(REF_PATH + " (build with `make -C oracle ref` where the reference sources exist)")
    return load(REF_PATH)


def _f(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return None if a is None else a.ctypes.data_as(_F)


def _sz(a):
    return np.array(a.shape[::-1] if a is not None else (0, 0, 0), dtype=np.int32)


def _cams(cameras):
    c = np.asarray(cameras, dtype=np.float32)
    if c.ndim == 1:
        return np.ascontiguousarray(c), np.array([c.shape[0], 1], np.int32)
    return np.ascontiguousarray(c), np.array([c.shape[1], c.shape[0]], np.int32)


def project(lib, act, mu, cameras, psf=None, bg=0.0, bga=0.0, truncate=1, gpu=0):
    act, mu, psf = _f(act), _f(mu), _f(psf)
    ref = act if act is not None else mu
    nz, ny, nx = ref.shape
    cams, csz = _cams(cameras)
    V = int(csz[0])
    sino = np.zeros((V, ny, nx), np.float32)
    asz = _sz(act)
    st = lib.SPECT_project_parallelholes(
        _p(act), asz.ctypes.data_as(_I), _p(sino), np.array([nx, ny, V], np.int32).ctypes.data_as(_I), _p(cams),
        csz.ctypes.data_as(_I), _p(psf), _sz(psf).ctypes.data_as(_I), _p(mu), _sz(mu).ctypes.data_as(_I),
        ctypes.byref(ctypes.c_float(bg)), ctypes.byref(ctypes.c_float(bga)), ctypes.byref(ctypes.c_int(gpu)),
        ctypes.byref(ctypes.c_int(truncate)))
    return st, sino


def backproject(lib, sino, mu, cameras, psf=None, bg=0.0, bga=0.0, truncate=1, gpu=0):
    sino, mu, psf = _f(sino), _f(mu), _f(psf)
    V, ny, nx = sino.shape
    cams, csz = _cams(cameras)
    out = np.zeros((nx, ny, nx), np.float32)
    st = lib.SPECT_backproject_parallelholes(
        _p(sino), np.array([nx, ny, V], np.int32).ctypes.data_as(_I), _p(out),
        np.array([nx, ny, nx], np.int32).ctypes.data_as(_I), _p(cams), csz.ctypes.data_as(_I), _p(psf),
        _sz(psf).ctypes.data_as(_I), _p(mu), _sz(mu).ctypes.data_as(_I), ctypes.byref(ctypes.c_float(bg)),
        ctypes.byref(ctypes.c_float(bga)), ctypes.byref(ctypes.c_int(gpu)), ctypes.byref(ctypes.c_int(truncate)))
    return st, out


def gradient_attenuation(lib, sino, act, mu, cameras, psf=None, bg=0.0, bga=0.0, truncate=1, gpu=1):
    sino, act, mu, psf = _f(sino), _f(act), _f(mu), _f(psf)
    V, ny, nx = sino.shape
    cams, csz = _cams(cameras)
    out = np.zeros_like(act)
    st = lib.et_array_gradient_attenuation(
        _p(sino), np.array([nx, ny, V], np.int32).ctypes.data_as(_I), _p(act), _sz(act).ctypes.data_as(_I),
        _p(out), _sz(out).ctypes.data_as(_I), _p(cams), csz.ctypes.data_as(_I), _p(psf),
        _sz(psf).ctypes.data_as(_I), _p(mu), _sz(mu).ctypes.data_as(_I), bg, bga, gpu, truncate)
    return st, out


def _cams3(cameras):
    c = np.asarray(cameras, dtype=np.float32)
    if c.ndim == 1:
        c3 = np.zeros((3, c.shape[0]), np.float32)
        c3[1] = c
        return c3
    return np.ascontiguousarray(c)


def mlem_step(lib, act, sino, mu, cameras, psf=None, bg=0.0, bga=0.0, eps=1e-6, gpu=0):
    act = np.array(act, dtype=np.float32, copy=True)
    sino, mu, psf = _f(sino), _f(mu), _f(psf)
    c3 = _cams3(cameras)
    nz, ny, nx = act.shape
    h = psf.shape[-1] if psf is not None else 0
    st = lib.et_array_mlem_step(_p(act), nx, ny, _p(sino), c3.shape[1], _p(c3), int(mu is not None), _p(mu),
                                int(psf is not None), _p(psf), h, h, bg, bga, eps, gpu)
    return st, act


def osem_step(lib, act, sino, mu, cameras, subset_order, psf=None, bg=0.0, bga=0.0, eps=1e-6, gpu=0):
    act = np.array(act, dtype=np.float32, copy=True)
    sino, mu, psf = _f(sino), _f(mu), _f(psf)
    c3 = _cams3(cameras)
    nz, ny, nx = act.shape
    h = psf.shape[-1] if psf is not None else 0
    st = lib.et_array_osem_step(_p(act), nx, ny, subset_order, _p(sino), c3.shape[1], _p(c3), int(mu is not None),
                                _p(mu), int(psf is not None), _p(psf), h, h, bg, bga, eps, gpu)
    return st, act


def rotate(lib, img, angles, centers, bg=0.0, order=0, gpu=0):
    img = _f(img)
    out = np.zeros_like(img)
    a = np.ascontiguousarray(angles, np.float32)
    c = np.ascontiguousarray(centers, np.float32)
    st = lib.et_array_rotate(_p(img), _sz(img).ctypes.data_as(_I), _p(out), _p(a), _p(c), bg, order, gpu)
    return st, out


def _fisher_argtypes(lib):
    f, c = ctypes.c_float, ctypes.c_int
    lib.et_array_fisher_grid.restype = ctypes.c_int
    lib.et_array_fisher_grid.argtypes = [_F, _I, _F, _I, _F, _I, _F, _F, _F, _I, _F, _I, f, f, f, c]
    lib.et_array_fisher_grid_projection.restype = ctypes.c_int
    lib.et_array_fisher_grid_projection.argtypes = [_F, _I, _I, _F, _I, _F, _I, _F, _F, _F, _I, _F, _I, f, f, f, c]


def fisher_grid(lib, act, mu, grid, cameras, psf, n=None, bg=0.0, bga=0.0, gpu=0, prior=True, eps=0.0):
    """et_array_fisher_grid: returns (status, fisher (n, n), prior (n, n) or None)."""
    _fisher_argtypes(lib)
    act, mu, grid, psf = _f(act), _f(mu), _f(grid), _f(psf)
    n = int(grid.max()) if n is None else int(n)
    cams, csz = _cams(cameras)
    fisher = np.zeros((n, n), np.float32)
    pr = np.zeros((n, n), np.float32) if prior else None
    st = lib.et_array_fisher_grid(
        _p(act), _sz(act).ctypes.data_as(_I), _p(cams), csz.ctypes.data_as(_I), _p(psf),
        _sz(psf).ctypes.data_as(_I), _p(grid), _p(fisher), _p(pr), np.array([n, n], np.int32).ctypes.data_as(_I),
        _p(mu), _sz(mu).ctypes.data_as(_I), eps, bg, bga, gpu)
    return st, fisher, pr


def fisher_grid_projection(lib, sino, grid, cameras, psf, n=None, mu=None, bg=0.0, bga=0.0, gpu=0, prior=True,
                           eps=0.0):
    """et_array_fisher_grid_projection: sino (V, ny, nx), grid (nz, ny, nx)."""
    _fisher_argtypes(lib)
    sino, mu, grid, psf = _f(sino), _f(mu), _f(grid), _f(psf)
    n = int(grid.max()) if n is None else int(n)
    cams, csz = _cams(cameras)
    fisher = np.zeros((n, n), np.float32)
    pr = np.zeros((n, n), np.float32) if prior else None
    st = lib.et_array_fisher_grid_projection(
        _p(sino), _sz(sino).ctypes.data_as(_I), _sz(grid).ctypes.data_as(_I), _p(cams), csz.ctypes.data_as(_I),
        _p(psf), _sz(psf).ctypes.data_as(_I), _p(grid), _p(fisher), _p(pr),
        np.array([n, n], np.int32).ctypes.data_as(_I), _p(mu), _sz(mu).ctypes.data_as(_I), eps, bg, bga, gpu)
    return st, fisher, pr
