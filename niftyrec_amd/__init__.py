"""MI355X (gfx950) SPECT projector with NiftyRec's array interface.

    from niftyrec_amd import api            # numpy / torch front-end
    from niftyrec_amd import NiftyRec       # Python 3 mirror of the reference package
    from niftyrec_amd.torch_ops import project   # differentiable projector

All computation runs in lib_et_array_interface.so (HIP kernels); importing a
front-end that needs the library raises ImportError when it is not built.
"""
from . import _lib

__all__ = ["api", "NiftyRec", "phantoms", "load_library"]


def load_library():
    """The ctypes handle of the HIP library (raises if it is not built)."""
    return _lib.load()
