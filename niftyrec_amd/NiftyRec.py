"""Python 3 counterpart of the reference ctypes package
Python/NiftyRec/NiftyRec/NiftyRec.py, bound to the HIP library.

Same public functions, argument meaning and return conventions as the
reference (et_project, et_backproject, et_set_gpu, et_list_gpus,
et_is_block_multiple, et_get_block_size, et_osmapem_step, Reconstructor).
Array sizes are passed as the numpy shapes and the array memory as it is,
exactly as the reference does (NiftyRec.py:86-112): an array of shape
(a, b, c) is read by the library as size [a, b, c] with the first size
fastest.  That is MATLAB (Fortran) order: x-fastest memory with the shape
listed x first, e.g. a psf held as psf_zyx.T (shape (h, h, N)); for the
[N, m, N] volumes of a SPECT scan a C-order array indexed [z, y, x] reads the
same.  The returned projection has the reference's shape [a, b, n_cameras];
its memory holds the library's [x, y, camera] planes, i.e. it reads
correctly as `proj.reshape(n_cameras, b, a)`.

Differences from the reference module, all fixes of Python-2-era bugs:
cameras are always passed in [axis][camera] memory order (the reference
loses that order when it converts a float64 camera array), `None` checks use
`is None`, and et_osmapem_step uses this module's projector.
"""
import threading
from time import sleep

import numpy as np
from ctypes import byref, c_int

from . import _lib

MAX_DEVICES = 16
SIZE_INFO = 5
ET_BACKGROUND_ACTIVITY = 0.0
ET_BACKGROUND_ATTENUATION = 0.0
ET_EPSILON = 0.000001
BG_ACTIVITY = ET_BACKGROUND_ACTIVITY
BG_ATTENUATION = ET_BACKGROUND_ATTENUATION
EPSILON = ET_EPSILON


def _f32(a):
    """float32 with the memory layout kept (the library reads the memory as
    it is, sizes from the shape); arrays that are neither C- nor
    Fortran-contiguous are copied to Fortran order, first size fastest."""
    a = np.asarray(a, dtype=np.float32, order="K")
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
        a = np.asfortranarray(a)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_lib.fp)


def _iptr(a):
    return a.ctypes.data_as(_lib.ip)


def _cameras(cameras):
    """(3, V) camera matrix -> ([axis][camera] float32 memory, size [V, 3])."""
    c = np.asarray(cameras)
    if c.ndim == 1:
        c = c.reshape(1, -1)
    c = np.ascontiguousarray(c, dtype=np.float32)
    return c, _i32([c.shape[1], c.shape[0]])


def et_project(activity, cameras, psf=None, attenuation=None, gpu=1, background=0.0, background_attenuation=0.0,
               truncate_negative_values=1):
    """Project activity to multiple cameras (reference NiftyRec.py:58-117)."""
    L = _lib.load()
    activity = _f32(activity)
    cams, cams_size = _cameras(cameras)
    act_size = _i32(activity.shape)
    if psf is not None:
        psf = _f32(psf)
        psf_size = _i32(psf.shape)
    else:
        psf_size = _i32([0, 0, 0])
    if attenuation is not None:
        attenuation = _f32(attenuation)
        att_size = _i32(attenuation.shape)
    else:
        att_size = _i32([0, 0, 0])
    projection = np.zeros([act_size[0], act_size[1], cams_size[0]], dtype=np.float32)
    proj_size = _i32(projection.shape)
    status = L.et_array_project(_ptr(activity), _iptr(act_size), _ptr(projection), _iptr(proj_size), _ptr(cams),
                                _iptr(cams_size), _ptr(psf), _iptr(psf_size), _ptr(attenuation), _iptr(att_size),
                                float(background), float(background_attenuation), int(gpu),
                                int(truncate_negative_values))
    if status:
        return None
    return projection


def et_backproject(projections, cameras, psf=None, attenuation=None, gpu=1, background=0, background_attenuation=0.0,
                   truncate_negative_values=1):
    """Backproject from multiple cameras to body space (reference NiftyRec.py:121-180)."""
    L = _lib.load()
    projections = _f32(projections)
    cams, cams_size = _cameras(cameras)
    proj_size = _i32(projections.shape)
    if psf is not None:
        psf = _f32(psf)
        psf_size = _i32(psf.shape)
    else:
        psf_size = _i32([0, 0, 0])
    if attenuation is not None:
        attenuation = _f32(attenuation)
        att_size = _i32(attenuation.shape)
    else:
        att_size = _i32([0, 0, 0])
    bkpr = np.zeros([proj_size[0], proj_size[1], proj_size[0]], dtype=np.float32)
    bkpr_size = _i32(bkpr.shape)
    status = L.et_array_backproject(_ptr(projections), _iptr(proj_size), _ptr(bkpr), _iptr(bkpr_size), _ptr(cams),
                                    _iptr(cams_size), _ptr(psf), _iptr(psf_size), _ptr(attenuation),
                                    _iptr(att_size), float(background), float(background_attenuation), int(gpu),
                                    int(truncate_negative_values))
    if status:
        return None
    return bkpr


def et_set_gpu(gpu_id):
    """Set a GPU device with given ID (reference NiftyRec.py:184-189)."""
    return _lib.load().et_array_set_gpu(int(gpu_id))


def et_list_gpus():
    """[gpu_count, flat info list of SIZE_INFO ints per device] or None (reference NiftyRec.py:193-206)."""
    L = _lib.load()
    count = np.zeros(1, dtype=np.int32)
    info = np.zeros(MAX_DEVICES * SIZE_INFO, dtype=np.int32)
    status = L.et_array_list_gpus(_iptr(count), _iptr(info))
    if status:
        return None
    n = int(count[0])
    return [n, info[: SIZE_INFO * min(n, MAX_DEVICES)].tolist()]


def et_is_block_multiple(integer_value):
    return _lib.load().et_array_is_block_multiple(int(integer_value))


def et_get_block_size():
    return _lib.load().et_array_get_block_size()


def et_osmapem_step(subset_order, activity_old, sinogram, cameras, attenuation, psf, beta_osl, gradient_prior,
                    use_gpu, background_activity=ET_BACKGROUND_ACTIVITY,
                    background_attenuation=ET_BACKGROUND_ATTENUATION, epsilon=ET_EPSILON, rng=None):
    """Maximum a posteriori EM, one-step-late (reference NiftyRec.py:224-246).

    sinogram is indexed [camera, a, b] as in the reference; the camera subset
    is drawn with numpy's generator `rng` (default: the global numpy state)."""
    N1, N2, N3 = activity_old.shape
    cameras = np.asarray(cameras)
    n_cam = cameras.shape[1]
    n_sub = int(np.int32(np.round(n_cam / subset_order)))
    r = (rng.random(n_sub) if rng is not None else np.random.rand(n_sub))
    idx = np.int32(np.round(0.5 + (n_cam - 0.5) * r) - 1)
    cams_sub = np.zeros((3, n_sub))
    cams_sub[:, :] = cameras[:, idx]
    normalization = et_backproject(np.ones((N1, N2, n_sub), dtype=np.float32), cams_sub, psf, attenuation, use_gpu,
                                   background_activity, background_attenuation)
    normalization = normalization - beta_osl * gradient_prior
    normalization = normalization * (normalization > 0) + epsilon
    proj = et_project(activity_old, cams_sub, psf, attenuation, use_gpu, background_activity, background_attenuation)
    proj = proj * (proj > 0) + epsilon
    update = et_backproject(np.asarray(sinogram)[idx, :, :].reshape(N1, N2, n_sub) / proj, cams_sub, psf,
                            attenuation, use_gpu, background_activity, background_attenuation)
    update = update * (update > 0)
    return activity_old * update / normalization


class Reconstructor:
    """Threaded OSEM / MAP-EM driver (reference NiftyRec.py:252-438)."""

    def __init__(self, volume_size):
        self._reconstructing = False
        self.volume_size = volume_size
        self.clear_activity()
        self.N_cameras = 0
        self.cameras = None
        self.psf = None
        self.attenuation = np.zeros(self.volume_size, dtype=np.float32)
        self.sinogram = None
        self.use_gpu = True
        self.background_activity = np.float32(BG_ACTIVITY)
        self.background_attenuation = np.float32(BG_ATTENUATION)
        self.epsilon = EPSILON
        self.callback_status = lambda msg, frac: None
        self.callback_updateactivity = lambda act: None
        self._thread = None
        self._stop = False

    def set_callback_status(self, function):
        self.callback_status = function

    def set_callback_updateactivity(self, function):
        self.callback_updateactivity = function

    def set_activity(self, activity):
        self.activity = np.asarray(activity, dtype=np.float32, order="K")

    def clear_activity(self):
        self.activity = np.ones(self.volume_size, dtype=np.float32)

    def set_attenuation(self, attenuation):
        self.attenuation = np.asarray(attenuation, dtype=np.float32, order="K")

    def set_sinogram(self, sinogram):
        if sinogram.shape[1] != self.volume_size[0] or sinogram.shape[2] != self.volume_size[1]:
            print("Sinogram size not consistent with activity estimate, need reinterpolation")
            return
        self.sinogram = np.asarray(sinogram, dtype=np.float32, order="K")

    def set_psf_matrix(self, psf):
        self.psf = np.asarray(psf, dtype=np.float32, order="K")

    def set_cameras_matrix(self, cameras):
        self.cameras = np.float32(cameras)
        self.N_cameras = cameras.shape[1]

    def set_cameras_equispaced(self, first_camera, last_camera, N_cameras, axis=1):
        cams = np.zeros((3, N_cameras))
        cams[axis, :] = np.linspace(first_camera * np.pi / 180.0, last_camera * np.pi / 180.0, N_cameras)
        self.N_cameras = N_cameras
        self.cameras = np.float32(cams)

    def set_use_gpu(self, enable):
        self.use_gpu = int(enable)

    def set_background_activity(self, bg):
        self.background_activity = np.float32(bg)

    def set_background_attenuation(self, bg):
        self.background_attenuation = np.float32(bg)

    def set_epsilon(self, epsilon):
        self.epsilon = np.float32(epsilon)

    def has_all_parameters(self):
        return self.N_cameras != 0 and self.cameras is not None and self.sinogram is not None

    def is_reconstructing(self):
        return self._reconstructing

    def reconstruct(self, method, parameters, block=False):
        if self.is_reconstructing() or not self.has_all_parameters():
            return False
        self._stop = False
        self._reconstructing = True
        self._thread = threading.Thread(target=self._reconstruct, args=(method, parameters), daemon=True)
        self._thread.start()
        if block:
            self._thread.join()
        return True

    def _reconstruct(self, method, parameters):
        try:
            steps = parameters["steps"]
            subset_order = parameters["subset_order"]
            for step in range(steps):
                beta, prior = 0.0, 0.0
                if method == "quadratic_MRF":
                    from scipy.signal import convolve
                    beta = parameters["beta"] / 1000.0
                    kernel = -np.ones((3, 3, 3))
                    kernel[1, 1, 1] = 26
                    prior = convolve(self.activity, kernel, "same")
                elif method != "osem":
                    sleep(0.1)
                    if self._stop:
                        break
                    continue
                self.activity = et_osmapem_step(subset_order, self.activity, self.sinogram, self.cameras,
                                                self.attenuation, self.psf, beta, prior, self.use_gpu,
                                                self.background_activity, self.background_attenuation,
                                                self.epsilon)
                self.callback_status("%d%% %s reconstruction" % (100 * step // max(steps, 1), method),
                                     float(step) / steps)
                self.callback_updateactivity(self.activity)
                if self._stop:
                    break
        finally:
            self._reconstructing = False
            self.callback_status(" ", 0.0)

    def stop(self):
        if self.is_reconstructing():
            self._stop = True
