"""numpy and torch front-end of the HIP SPECT projector.

Conventions (numpy / torch C order, so shapes read reversed relative to the
C sizes of the array interface):
    volume    (nz, ny, nx)          C size [nx, ny, nz]
    sinogram  (V, ny, nx)           C size [nx, ny, V]
    psf       (nz, h, h)            C size [h, h, nz]
    cameras   (V,) angles about the axial axis (y), or (3, V) per-axis angles,
              or (6, V) angles followed by rotation centres; radians.

numpy functions go through the NiftyRec array interface (host buffers,
PCIe transfers included).  `Plan` keeps the geometry on the GPU and works on
torch tensors already in device memory.
"""
import ctypes

import numpy as np

from . import _lib

REFERENCE = 0
ADJOINT = 1


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _cams(cameras):
    """-> (float32 [axis][camera] memory, cols, V)."""
    c = np.asarray(cameras, dtype=np.float32)
    if c.ndim == 1:
        return np.ascontiguousarray(c), 1, c.shape[0]
    if c.shape[0] not in (1, 3, 6):
        raise ValueError("cameras must be (V,), (3, V) or (6, V)")
    return np.ascontiguousarray(c), c.shape[0], c.shape[1]


def _size(a):
    return (ctypes.c_int * 3)(*(a.shape[::-1] if a is not None else (0, 0, 0)))


def _p(a):
    return None if a is None else a.ctypes.data_as(_lib.fp)


def project(activity, cameras, psf=None, attenuation=None, background=0.0, background_attenuation=0.0,
            truncate_negative_values=True):
    """Attenuated, PSF-blurred parallel-hole projection, (V, ny, nx)."""
    L = _lib.load()
    act, mu, psf = _f32(activity), _f32(attenuation), _f32(psf)
    ref = act if act is not None else mu
    cams, cols, V = _cams(cameras)
    nz, ny, nx = ref.shape
    sino = np.zeros((V, ny, nx), dtype=np.float32)
    act_size = _size(act)
    st = L.et_array_project(_p(act), act_size, _p(sino), (ctypes.c_int * 3)(nx, ny, V), _p(cams),
                            (ctypes.c_int * 2)(V, cols), _p(psf), _size(psf), _p(mu), _size(mu),
                            float(background), float(background_attenuation), 1, int(bool(truncate_negative_values)))
    _lib.check(st, "project")
    return sino


def project_partial(activity, cameras, psf=None, attenuation=None, background=0.0, background_image=None,
                    background_attenuation=0.0, truncate_negative_values=True, rotate_partial=True):
    """Projection plus the partial line integrals of every view:
    (sinogram (V, ny, nx), partial sums (V, nz, ny, nx)); the partial sums are
    in the image frame when rotate_partial, else in each view's rotated frame."""
    L = _lib.load()
    act, mu, psf, bgi = _f32(activity), _f32(attenuation), _f32(psf), _f32(background_image)
    ref = act if act is not None else mu
    cams, cols, V = _cams(cameras)
    nz, ny, nx = ref.shape
    sino = np.zeros((V, ny, nx), dtype=np.float32)
    part = np.zeros((V, nz, ny, nx), dtype=np.float32)
    st = L.et_array_project_partial(_p(act), _size(act), _p(sino), (ctypes.c_int * 3)(nx, ny, V), _p(part),
                                    (ctypes.c_int * 4)(nx, ny, nz, V), _p(cams), (ctypes.c_int * 2)(V, cols),
                                    _p(psf), _size(psf), _p(mu), _size(mu), float(background), _p(bgi),
                                    float(background_attenuation), 1, int(bool(truncate_negative_values)),
                                    int(bool(rotate_partial)))
    _lib.check(st, "project_partial")
    return sino, part


def backproject(sinogram, cameras, psf=None, attenuation=None, background=0.0, background_attenuation=0.0,
                truncate_negative_values=True, nz=None):
    """Reference backprojection (inverse-rotation pull), volume (nz, ny, nx)."""
    L = _lib.load()
    sino, mu, psf = _f32(sinogram), _f32(attenuation), _f32(psf)
    cams, cols, V = _cams(cameras)
    _, ny, nx = sino.shape
    nz = nx if nz is None else nz
    out = np.zeros((nz, ny, nx), dtype=np.float32)
    st = L.et_array_backproject(_p(sino), (ctypes.c_int * 3)(nx, ny, V), _p(out), (ctypes.c_int * 3)(nx, ny, nz),
                                _p(cams), (ctypes.c_int * 2)(V, cols), _p(psf), _size(psf), _p(mu), _size(mu),
                                float(background), float(background_attenuation), 1,
                                int(bool(truncate_negative_values)))
    _lib.check(st, "backproject")
    return out


def gradient_attenuation(sinogram, activity, attenuation, cameras, psf=None, background=0.0,
                         background_attenuation=0.0, truncate_negative_values=True):
    """The reference attenuation gradient (et_array_gradient_attenuation)."""
    L = _lib.load()
    sino, act, mu, psf = _f32(sinogram), _f32(activity), _f32(attenuation), _f32(psf)
    cams, cols, V = _cams(cameras)
    out = np.zeros_like(act)
    nz, ny, nx = act.shape
    st = L.et_array_gradient_attenuation(_p(sino), (ctypes.c_int * 3)(nx, ny, V), _p(act), _size(act), _p(out),
                                         _size(out), _p(cams), (ctypes.c_int * 2)(V, cols), _p(psf), _size(psf),
                                         _p(mu), _size(mu), float(background), float(background_attenuation), 1,
                                         int(bool(truncate_negative_values)))
    _lib.check(st, "gradient_attenuation")
    return out


def make_cameras(first_deg, last_deg, n, axis=1):
    """(3, n) cameras with inclusive endpoints (et_array_make_cameras)."""
    L = _lib.load()
    out = np.zeros((3, n), dtype=np.float32)
    _lib.check(L.et_array_make_cameras(_p(out), float(first_deg), float(last_deg), int(n), int(axis)),
               "make_cameras")
    return out


def _cams3(cameras):
    cams, cols, V = _cams(cameras)
    if cols == 1:
        c3 = np.zeros((3, V), dtype=np.float32)
        c3[1] = cams
        return c3, V
    if cols != 3:
        raise ValueError("EM steps take (V,) or (3, V) cameras")
    return cams, V


def mlem_step(activity, sinogram, cameras, psf=None, attenuation=None, background=0.0, background_attenuation=0.0,
              epsilon=1e-6):
    """One reference MLEM update (et_array_mlem_step); returns the new activity."""
    L = _lib.load()
    act = np.array(activity, dtype=np.float32, copy=True, order="C")
    sino, mu, psf = _f32(sinogram), _f32(attenuation), _f32(psf)
    c3, V = _cams3(cameras)
    nz, ny, nx = act.shape
    h = psf.shape[-1] if psf is not None else 0
    st = L.et_array_mlem_step(_p(act), nx, ny, _p(sino), V, _p(c3), int(mu is not None), _p(mu),
                              int(psf is not None), _p(psf), h, h, float(background), float(background_attenuation),
                              float(epsilon), 1)
    _lib.check(st, "mlem_step")
    return act


def osem_step(activity, sinogram, cameras, subset_order, psf=None, attenuation=None, background=0.0,
              background_attenuation=0.0, epsilon=1e-6):
    """One reference OSEM update on a random camera subset drawn with glibc
    rand() (et_array_osem_step); returns the new activity."""
    L = _lib.load()
    act = np.array(activity, dtype=np.float32, copy=True, order="C")
    sino, mu, psf = _f32(sinogram), _f32(attenuation), _f32(psf)
    c3, V = _cams3(cameras)
    nz, ny, nx = act.shape
    h = psf.shape[-1] if psf is not None else 0
    st = L.et_array_osem_step(_p(act), nx, ny, int(subset_order), _p(sino), V, _p(c3), int(mu is not None), _p(mu),
                              int(psf is not None), _p(psf), h, h, float(background), float(background_attenuation),
                              float(epsilon), 1)
    _lib.check(st, "osem_step")
    return act


def _driver_args(sinogram, psf, attenuation):
    sino, mu, psf = _f32(sinogram), _f32(attenuation), _f32(psf)
    V, ny, nx = sino.shape
    h = psf.shape[-1] if psf is not None else 0
    act = np.zeros((nx, ny, nx), dtype=np.float32)
    return sino, mu, psf, V, ny, nx, h, act


def mlem(sinogram, first_deg, last_deg, iterations, psf=None, attenuation=None, rotation_axis=1, background=0.0,
         background_attenuation=0.0, epsilon=1e-6):
    """`iterations` MLEM updates from a uniform start (et_array_mlem): cameras
    from first_deg to last_deg inclusive about `rotation_axis`; the data stay
    in device memory for every iteration.  Returns the activity (nx, ny, nx)."""
    L = _lib.load()
    sino, mu, psf, V, ny, nx, h, act = _driver_args(sinogram, psf, attenuation)
    st = L.et_array_mlem(_p(act), nx, ny, _p(sino), V, float(first_deg), float(last_deg), int(rotation_axis),
                         int(iterations), int(mu is not None), _p(mu), int(psf is not None), _p(psf), h, h,
                         float(background), float(background_attenuation), float(epsilon), 1)
    _lib.check(st, "mlem")
    return act


def osem(sinogram, subset_order, first_deg, last_deg, iterations, psf=None, attenuation=None, rotation_axis=1,
         background=0.0, background_attenuation=0.0, epsilon=1e-6):
    """`iterations` reference OSEM steps from a uniform start (et_array_osem):
    each on V // subset_order cameras drawn with replacement from glibc rand()."""
    L = _lib.load()
    sino, mu, psf, V, ny, nx, h, act = _driver_args(sinogram, psf, attenuation)
    st = L.et_array_osem(_p(act), nx, ny, int(subset_order), _p(sino), V, float(first_deg), float(last_deg),
                         int(rotation_axis), int(iterations), int(mu is not None), _p(mu), int(psf is not None),
                         _p(psf), h, h, float(background), float(background_attenuation), float(epsilon), 1)
    _lib.check(st, "osem")
    return act


def osem_ordered(sinogram, n_subsets, first_deg, last_deg, iterations, psf=None, attenuation=None,
                 rotation_axis=1, background=0.0, background_attenuation=0.0, epsilon=1e-6):
    """`iterations` passes over `n_subsets` ordered subsets (subset s: cameras
    s, s + n_subsets, ...) from a uniform start (et_array_osem_ordered)."""
    L = _lib.load()
    sino, mu, psf, V, ny, nx, h, act = _driver_args(sinogram, psf, attenuation)
    st = L.et_array_osem_ordered(_p(act), nx, ny, int(n_subsets), _p(sino), V, float(first_deg), float(last_deg),
                                 int(rotation_axis), int(iterations), int(mu is not None), _p(mu),
                                 int(psf is not None), _p(psf), h, h, float(background),
                                 float(background_attenuation), float(epsilon), 1)
    _lib.check(st, "osem_ordered")
    return act


def rotate(image, angles, centers, background=0.0, axis_order=0):
    """Rotation resampler (et_array_rotate), image (nz, ny, nx)."""
    L = _lib.load()
    img = _f32(image)
    out = np.zeros_like(img)
    a = np.ascontiguousarray(angles, dtype=np.float32)
    c = np.ascontiguousarray(centers, dtype=np.float32)
    _lib.check(L.et_array_rotate(_p(img), _size(img), _p(out), _p(a), _p(c), float(background), int(axis_order), 1),
               "rotate")
    return out


def convolve(image, psf):
    """Depth-dependent separable convolution of every z plane (et_array_convolve)."""
    L = _lib.load()
    img, psf = _f32(image), _f32(psf)
    out = np.zeros_like(img)
    _lib.check(L.et_array_convolve(_p(img), _size(img), _p(out), _size(out), _p(psf), _size(psf), 1), "convolve")
    return out


def fisher_grid(activity, grid, cameras, psf, n=None, attenuation=None, background=0.0,
                background_attenuation=0.0, prior=True, gpu_prior_rule=True):
    """Fisher information (n, n) of the points of `grid` (nz, ny, nx): 0, or the
    1-based index of a point; n defaults to the largest index.  The activity is
    projected first (et_array_fisher_grid).  Returns (fisher, prior or None);
    gpu_prior_rule selects the reference CUDA prior rule (|d| <= 3, off the
    diagonal) over its CPU rule (|d| <= 1, with the diagonal)."""
    L = _lib.load()
    act, mu, g, psf = _f32(activity), _f32(attenuation), _f32(grid), _f32(psf)
    n = int(g.max()) if n is None else int(n)
    cams, cols, V = _cams(cameras)
    F = np.zeros((n, n), np.float32)
    P = np.zeros((n, n), np.float32) if prior else None
    st = L.et_array_fisher_grid(_p(act), _size(act), _p(cams), (ctypes.c_int * 2)(V, cols), _p(psf), _size(psf),
                                _p(g), _p(F), _p(P), (ctypes.c_int * 2)(n, n), _p(mu), _size(mu), 0.0,
                                float(background), float(background_attenuation), int(bool(gpu_prior_rule)))
    _lib.check(st, "fisher_grid")
    return F, P


def fisher_grid_projection(projection, grid, cameras, psf, n=None, prior=True, gpu_prior_rule=True):
    """fisher_grid from a given projection (V, ny, nx) of the activity
    (et_array_fisher_grid_projection); grid (nz, ny, nx)."""
    L = _lib.load()
    s, g, psf = _f32(projection), _f32(grid), _f32(psf)
    n = int(g.max()) if n is None else int(n)
    cams, cols, V = _cams(cameras)
    F = np.zeros((n, n), np.float32)
    P = np.zeros((n, n), np.float32) if prior else None
    st = L.et_array_fisher_grid_projection(_p(s), _size(s), _size(g), _p(cams), (ctypes.c_int * 2)(V, cols),
                                           _p(psf), _size(psf), _p(g), _p(F), _p(P), (ctypes.c_int * 2)(n, n), None,
                                           (ctypes.c_int * 3)(0, 0, 0), 0.0, 0.0, 0.0, int(bool(gpu_prior_rule)))
    _lib.check(st, "fisher_grid_projection")
    return F, P


class Plan:
    """Device-resident projector for torch tensors (cuda / ROCm device memory).

    shape: (nz, ny, nx); cameras as for `project`; psf (nz, h, h) or None.
    Tensors passed to the methods must be float32, contiguous, of the plan's
    volume or sinogram shape and on the plan's device (checked before every
    call: the kernels trust the sizes); work is queued on torch's current
    stream.

    mu_key: how osem knows that the attenuation map changed, so that its cached
    subset sensitivities and attenuation factors are rebuilt.  "version"
    (default): the tensor itself and its autograd version counter, which every
    in-place torch operation (and every write of this module into an `out`
    tensor) advances; no device work.  "content": a device fingerprint of the
    map at every call (one pass over it and one stream synchronisation), for
    maps written behind torch's back."""

    def __init__(self, shape, cameras, psf=None, background=0.0, background_attenuation=0.0, mu_key="version"):
        L = _lib.load()
        nz, ny, nx = shape
        self.shape = tuple(int(v) for v in shape)
        cams, cols, V = _cams(cameras)
        self.n_views = V
        self.sino_shape = (V, ny, nx)
        psf = _f32(psf)
        h = psf.shape[-1] if psf is not None else 0
        handle = ctypes.c_void_p()
        st = L.et_plan_create(ctypes.byref(handle), nx, ny, nz, _p(cams), V, cols, _p(psf), h, float(background),
                              float(background_attenuation))
        _lib.check(st, "et_plan_create")
        self._h = handle
        self._L = L
        dev = ctypes.c_int()
        _lib.check(L.et_plan_get_device(handle, ctypes.byref(dev)), "et_plan_get_device")
        self.device_index = dev.value
        if mu_key not in ("version", "content"):
            raise ValueError("mu_key must be 'version' or 'content'")
        self.mu_key = mu_key
        self._mu_seen = None  # (weak reference, version) of the map of the cached OSEM data
        _lib.check(L.et_plan_set_mu_check(handle, int(mu_key == "content")), "et_plan_set_mu_check")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.et_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers ------------------------------------------------------------
    @staticmethod
    def _ptr(t):
        if t is None:
            return None
        import torch
        if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("expected a contiguous float32 GPU tensor")
        return ctypes.c_void_p(t.data_ptr())

    def _arg(self, t, shape, name):
        """Device pointer of `t` after checking dtype, layout, device and shape."""
        if t is None:
            return None
        import torch
        if not isinstance(t, torch.Tensor):
            raise TypeError("%s must be a torch tensor" % name)
        if tuple(t.shape) != tuple(shape):
            raise ValueError("%s has shape %s, expected %s" % (name, tuple(t.shape), tuple(shape)))
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("%s: expected a contiguous float32 tensor" % name)
        if not t.is_cuda or t.device.index != self.device_index:
            raise ValueError("%s is on %s, the plan on cuda:%d" % (name, t.device, self.device_index))
        return ctypes.c_void_p(t.data_ptr())

    @staticmethod
    def _stream():
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _new(self, shape, like):
        import torch
        return torch.empty(shape, dtype=torch.float32, device=like.device)

    @staticmethod
    def _written(t):
        """Advance the version counter of a tensor this module wrote in place."""
        import torch
        if hasattr(torch.autograd.graph, "increment_version"):
            torch.autograd.graph.increment_version(t)

    def _track_mu(self, mu):
        """With mu_key "version": drop the cached OSEM data unless `mu` is the
        tensor, at the same version, that they were built from."""
        if self.mu_key != "version":
            return
        import weakref
        seen = self._mu_seen
        same = (seen is not None and
                (mu is None and seen[0] is None or
                 mu is not None and seen[0] is not None and seen[0]() is mu and seen[1] == mu._version))
        if not same:
            _lib.check(self._L.et_plan_invalidate(self._h), "et_plan_invalidate")
        self._mu_seen = (None, None) if mu is None else (weakref.ref(mu), mu._version)

    # -- operators ------------------------------------------------------------
    def project(self, activity, attenuation=None, truncate=True, out=None):
        if activity is None and attenuation is None:
            raise ValueError("project needs an activity or an attenuation map")
        ref = activity if activity is not None else attenuation
        a = self._arg(activity, self.shape, "activity")
        m = self._arg(attenuation, self.shape, "attenuation")
        out = self._new(self.sino_shape, ref) if out is None else out
        o = self._arg(out, self.sino_shape, "out")
        st = self._L.et_plan_project(self._h, a, m, o, int(bool(truncate)), self._stream())
        _lib.check(st, "et_plan_project")
        self._written(out)
        return out

    def backproject(self, sinogram, attenuation=None, mode=REFERENCE, truncate=True, out=None):
        s = self._arg(sinogram, self.sino_shape, "sinogram")
        m = self._arg(attenuation, self.shape, "attenuation")
        out = self._new(self.shape, sinogram) if out is None else out
        o = self._arg(out, self.shape, "out")
        st = self._L.et_plan_backproject(self._h, s, m, o, int(mode), int(bool(truncate)), self._stream())
        _lib.check(st, "et_plan_backproject")
        self._written(out)
        return out

    def gradient_attenuation(self, sinogram, activity, attenuation, exact=False, truncate=True, out=None):
        if activity is None or attenuation is None:
            raise ValueError("gradient_attenuation needs the activity and the attenuation map")
        s = self._arg(sinogram, self.sino_shape, "sinogram")
        a = self._arg(activity, self.shape, "activity")
        m = self._arg(attenuation, self.shape, "attenuation")
        out = self._new(self.shape, activity) if out is None else out
        o = self._arg(out, self.shape, "out")
        st = self._L.et_plan_gradient_attenuation(self._h, s, a, m, o, int(bool(exact)), int(bool(truncate)),
                                                  self._stream())
        _lib.check(st, "et_plan_gradient_attenuation")
        self._written(out)
        return out

    def project_backward(self, sinogram, activity, attenuation, truncate=False, out_activity=None,
                         out_attenuation=None):
        """Both gradients of <sinogram, project(activity, attenuation)>: the exact
        adjoint backprojection of the sinogram and the exact attenuation
        gradient, as backproject(mode=ADJOINT) and gradient_attenuation(exact=True)
        return them; right after project() of the same pair, one pass over the
        views (et_plan_project_backward).  Returns (d/d activity, d/d attenuation)."""
        if activity is None or attenuation is None:
            raise ValueError("project_backward needs the activity and the attenuation map")
        s = self._arg(sinogram, self.sino_shape, "sinogram")
        a = self._arg(activity, self.shape, "activity")
        m = self._arg(attenuation, self.shape, "attenuation")
        ga = self._new(self.shape, activity) if out_activity is None else out_activity
        gm = self._new(self.shape, activity) if out_attenuation is None else out_attenuation
        oa = self._arg(ga, self.shape, "out_activity")
        om = self._arg(gm, self.shape, "out_attenuation")
        st = self._L.et_plan_project_backward(self._h, s, a, m, oa, om, int(bool(truncate)), self._stream())
        _lib.check(st, "et_plan_project_backward")
        self._written(ga)
        self._written(gm)
        return ga, gm

    def osem(self, activity, sinogram, attenuation=None, subsets=1, iterations=1, epsilon=1e-6):
        """Ordered-subsets EM in place on `activity` (subset s: views s, s + S, ...).
        The subset sensitivities and attenuation factors are cached in the
        plan and rebuilt when the attenuation map's content changes (see
        set_mu_check)."""
        a = self._arg(activity, self.shape, "activity")
        s = self._arg(sinogram, self.sino_shape, "sinogram")
        m = self._arg(attenuation, self.shape, "attenuation")
        self._track_mu(attenuation)
        st = self._L.et_plan_osem(self._h, a, s, m, int(subsets), int(iterations), float(epsilon), self._stream())
        _lib.check(st, "et_plan_osem")
        self._written(activity)
        return activity

    def set_tuning(self, depth_chunks=0, view_batch=0):
        _lib.check(self._L.et_plan_set_tuning(self._h, int(depth_chunks), int(view_batch)), "et_plan_set_tuning")

    def set_mu_check(self, check=True):
        """True: osem keys its caches on a device fingerprint of the attenuation
        map's content (mu_key "content").  False: on its address only (the
        library's view under mu_key "version"); call invalidate() after changing
        the map behind torch's back."""
        _lib.check(self._L.et_plan_set_mu_check(self._h, int(bool(check))), "et_plan_set_mu_check")
        self.mu_key = "content" if check else "address"

    def invalidate(self):
        """Drop the cached OSEM data."""
        _lib.check(self._L.et_plan_invalidate(self._h), "et_plan_invalidate")

    def tuning(self):
        """(projection, backprojection) depth chunks of the cached OSEM data
        ((0, 0) before the first osem call)."""
        dp, db = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._L.et_plan_get_tuning(self._h, ctypes.byref(dp), ctypes.byref(db)), "et_plan_get_tuning")
        return dp.value, db.value

    def projector(self):
        """Projector of the cached OSEM data: the number of 64-column tiles of
        the tiled row-lane projector, 1 for the full-width row-lane projector,
        0 for the slab kernels or before the first osem call."""
        t = ctypes.c_int()
        _lib.check(self._L.et_plan_get_projector(self._h, ctypes.byref(t)), "et_plan_get_projector")
        return t.value

    def profile(self, enable=True, categories=None, every=1):
        """Time the plan's launches with HIP events: every category, or only the
        named ones (names of _lib.PROF_NAMES), and of those every `every`-th
        launch; resets the figures."""
        _lib.check(self._L.et_plan_profile_sample(self._h, int(every)), "et_plan_profile_sample")
        if categories is None:
            _lib.check(self._L.et_plan_profile(self._h, int(bool(enable))), "et_plan_profile")
            return
        mask = 0
        for name in categories:
            mask |= 1 << _lib.PROF_NAMES.index(name)
        _lib.check(self._L.et_plan_profile_select(self._h, mask if enable else 0), "et_plan_profile_select")

    def profile_read(self):
        """{category: (launches, total_ms, algorithmic_bytes)}."""
        res = {}
        for cat, name in enumerate(_lib.PROF_NAMES):
            n = ctypes.c_int()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            _lib.check(self._L.et_plan_profile_read(self._h, cat, ctypes.byref(n), ctypes.byref(ms),
                                                    ctypes.byref(by)), "et_plan_profile_read")
            res[name] = (n.value, ms.value, by.value)
        return res


class Batch:
    """Many plans reconstructed together: one OSEM iteration over every plan as
    grouped launches captured in one hipGraph (et_batch_osem).  The plans must
    take the fast path and stay alive while the batch is used."""

    def __init__(self, plans):
        L = _lib.load()
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        handle = ctypes.c_void_p()
        _lib.check(L.et_batch_create(ctypes.byref(handle), arr, len(self.plans)), "et_batch_create")
        self._h = handle
        self._L = L

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.et_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def osem(self, activities, sinograms, attenuations=None, subsets=1, iterations=1, epsilon=1e-6, use_graph=True):
        """OSEM in place on every activity tensor (lists in plan order)."""
        n = len(self.plans)
        if len(activities) != n or len(sinograms) != n or (attenuations is not None and len(attenuations) != n):
            raise ValueError("one activity, sinogram (and attenuation) per plan")
        P = ctypes.c_void_p * n
        acts = P(*[p._arg(a, p.shape, "activity %d" % i).value for i, (p, a) in enumerate(zip(self.plans, activities))])
        sinos = P(*[p._arg(s, p.sino_shape, "sinogram %d" % i).value
                    for i, (p, s) in enumerate(zip(self.plans, sinograms))])
        mus = None if attenuations is None else P(*[(p._arg(m, p.shape, "attenuation %d" % i).value
                                                     if m is not None else None)
                                                    for i, (p, m) in enumerate(zip(self.plans, attenuations))])
        for i, p in enumerate(self.plans):
            p._track_mu(None if attenuations is None else attenuations[i])
        st = self._L.et_batch_osem(self._h, acts, sinos, mus, int(subsets), int(iterations), float(epsilon),
                                   int(bool(use_graph)), Plan._stream())
        _lib.check(st, "et_batch_osem")
        for a in activities:
            Plan._written(a)
        return activities
