"""ctypes binding of lib_et_array_interface.so (the HIP library built in-tree).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object has not been built (`python -m niftyrec_amd.build`).
"""
import ctypes
import os

from ctypes import POINTER as P, c_char_p, c_double, c_float, c_int, c_uint, c_ushort, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "lib_et_array_interface.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

fp = P(c_float)
ip = P(c_int)

# name -> (restype, argtypes); mirrors include/et_array_interface.h
SIGNATURES = {
    "status_success": (c_int, [ip]),
    "status_io_error": (c_int, [ip]),
    "status_initialisation_error": (c_int, [ip]),
    "status_parameter_error": (c_int, [ip]),
    "status_unhandled_error": (c_int, [ip]),
    "echo": (c_int, [ip, ip]),
    "niftyrec_error_msg": (c_char_p, [c_int]),
    "et_array_list_gpus": (c_int, [ip, ip]),
    "et_array_set_gpu": (c_int, [c_int]),
    "et_array_set_gpu_pointer": (c_int, [ip]),
    "et_array_isinstalled": (c_int, []),
    "et_array_reset_gpu": (c_int, []),
    "et_array_is_block_multiple": (c_ushort, [c_ushort]),
    "et_array_get_block_size": (c_ushort, []),
    "SPECT_project_parallelholes": (c_int, [fp, ip, fp, ip, fp, ip, fp, ip, fp, ip, fp, fp, ip, ip]),
    "SPECT_backproject_parallelholes": (c_int, [fp, ip, fp, ip, fp, ip, fp, ip, fp, ip, fp, fp, ip, ip]),
    "et_array_project": (c_int, [fp, ip, fp, ip, fp, ip, fp, ip, fp, ip, c_float, c_float, c_int, c_int]),
    "et_array_backproject": (c_int, [fp, ip, fp, ip, fp, ip, fp, ip, fp, ip, c_float, c_float, c_int, c_int]),
    "et_array_gradient_attenuation": (c_int, [fp, ip, fp, ip, fp, ip, fp, ip, fp, ip, fp, ip, c_float, c_float,
                                              c_int, c_int]),
    "et_array_project_partial": (c_int, [fp, ip, fp, ip, fp, ip, fp, ip, fp, ip, fp, ip, c_float, fp, c_float,
                                         c_int, c_int, c_int]),
    "et_array_rotate": (c_int, [fp, ip, fp, fp, fp, c_float, c_int, c_int]),
    "et_array_affine": (c_int, [fp, ip, fp, ip, fp, ip, c_float, c_int]),
    "et_array_convolve": (c_int, [fp, ip, fp, ip, fp, ip, c_int]),
    "et_array_fisher_grid": (c_int, [fp, ip, fp, ip, fp, ip, fp, fp, fp, ip, fp, ip, c_float, c_float, c_float,
                                     c_int]),
    "et_array_fisher_grid_projection": (c_int, [fp, ip, ip, fp, ip, fp, ip, fp, fp, fp, ip, fp, ip, c_float,
                                                c_float, c_float, c_int]),
    "et_array_calculate_size_psf": (c_int, [P(c_uint), P(c_uint), c_float, c_float, c_float, c_float, c_float,
                                            c_float]),
    "et_array_make_psf": (c_int, [fp, c_uint, c_uint, c_float, c_float, c_float, c_float, c_float, c_float, c_uint]),
    "et_array_make_cameras": (c_int, [fp, c_float, c_float, c_uint, c_uint]),
    "et_array_mlem_step": (c_int, [fp, c_int, c_int, fp, c_int, fp, c_int, fp, c_int, fp, c_uint, c_uint, c_float,
                                   c_float, c_float, c_int]),
    "et_array_osem_step": (c_int, [fp, c_int, c_int, c_uint, fp, c_int, fp, c_int, fp, c_int, fp, c_uint, c_uint,
                                   c_float, c_float, c_float, c_int]),
    "et_array_mlem": (c_int, [fp, c_uint, c_uint, fp, c_int, c_float, c_float, c_uint, c_int, c_int, fp, c_int, fp,
                              c_uint, c_uint, c_float, c_float, c_float, c_int]),
    "et_array_osem": (c_int, [fp, c_uint, c_uint, c_uint, fp, c_int, c_float, c_float, c_uint, c_int, c_int, fp,
                              c_int, fp, c_uint, c_uint, c_float, c_float, c_float, c_int]),
    "et_array_osem_ordered": (c_int, [fp, c_uint, c_uint, c_uint, fp, c_int, c_float, c_float, c_uint, c_int, c_int,
                                      fp, c_int, fp, c_uint, c_uint, c_float, c_float, c_float, c_int]),
    "et_batch_create": (c_int, [P(c_void_p), P(c_void_p), c_int]),
    "et_batch_destroy": (c_int, [c_void_p]),
    "et_batch_osem": (c_int, [c_void_p, P(c_void_p), P(c_void_p), P(c_void_p), c_int, c_int, c_float, c_int,
                              c_void_p]),
    "et_plan_create": (c_int, [P(c_void_p), c_int, c_int, c_int, fp, c_int, c_int, fp, c_int, c_float, c_float]),
    "et_plan_destroy": (c_int, [c_void_p]),
    "et_plan_project": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "et_plan_project_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "et_plan_backproject": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "et_plan_gradient_attenuation": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                             c_void_p]),
    "et_plan_project_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                         c_void_p]),
    "et_plan_osem": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_float, c_void_p]),
    "et_plan_set_tuning": (c_int, [c_void_p, c_int, c_int]),
    "et_plan_set_mu_check": (c_int, [c_void_p, c_int]),
    "et_plan_invalidate": (c_int, [c_void_p]),
    "et_plan_get_device": (c_int, [c_void_p, P(c_int)]),
    "et_plan_get_tuning": (c_int, [c_void_p, P(c_int), P(c_int)]),
    "et_plan_get_projector": (c_int, [c_void_p, P(c_int)]),
    "et_plan_profile_select": (c_int, [c_void_p, c_uint]),
    "et_plan_profile_sample": (c_int, [c_void_p, c_int]),
    "et_plan_profile": (c_int, [c_void_p, c_int]),
    "et_plan_profile_read": (c_int, [c_void_p, c_int, ip, P(c_double), P(c_double)]),
    "et_library_version": (c_char_p, []),
}

# niftyrec status codes (include/et_array_interface.h)
SUCCESS = 0
ERROR_UNSPECIFIED = 1
ERROR_PARAMETERS = 2
ERROR_KERNEL = 3
ERROR_ALLOCGPU = 4
ERROR_ALLOCCPU = 5
ERROR_TRANSFERGPU = 6
ERROR_NOGPUBUILT = 7
ERROR_NOGPUINSTALLED = 8

PROF_PROJECT, PROF_BPU, PROF_PULL, PROF_MUSUM, PROF_OTHER, PROF_GRAD = range(6)
PROF_NAMES = ["project", "bp_rotated_frame", "bp_pull", "mu_prefix", "other", "gradient"]

_lib = None


class NiftyRecError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = "status %d" % status
        if _lib is not None:
            try:
                msg = _lib.niftyrec_error_msg(status).decode().strip()
            except Exception:
                pass
        super().__init__("%s: %s (status %d)" % (where, msg, status))


def bind(path):
    """Loads a library exporting the NiftyRec array interface and declares the
    argument types of every symbol it has from SIGNATURES."""
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    return lib


def load():
    """The product library; raises when it has not been built.

    torch (when installed) is imported first: its ROCm wheel carries its own
    HIP runtime under the same soname as the system one the library links, and
    whichever loads first serves the whole process.  Loaded in this order, the
    library and torch share torch's runtime; the other order leaves torch
    without a device."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("%s is not built: run `python -m niftyrec_amd.build`" % LIB_PATH)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = bind(LIB_PATH)
    return _lib


def check(status, where):
    if status != SUCCESS:
        raise NiftyRecError(status, where)
