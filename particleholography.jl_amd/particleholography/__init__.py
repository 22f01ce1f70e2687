"""MI355X-native angular-spectrum reconstruction and Gerchberg-Saxton phase retrieval.

A drop-in for the hot path of ParticleHolography.jl (src/holofunc.jl, src/types.jl): the
same function names and argument meaning over torch device tensors, executed by hand-written
HIP kernels in libholo.so (C ABI: include/holo.h). There is no CPU fallback.
"""
from . import _lib
from .holofunc import *  # noqa: F401,F403
from .holofunc import __all__ as _holofunc_all
from .holofunc import reconst_xyprojection_holo  # noqa: F401
from .frequency_filters import *  # noqa: F401,F403
from .frequency_filters import __all__ as _filters_all
from .utils import *  # noqa: F401,F403
from .utils import __all__ as _utils_all
from .ccl import *  # noqa: F401,F403
from .ccl import __all__ as _ccl_all
from .particle_detection import *  # noqa: F401,F403
from .particle_detection import __all__ as _detection_all

__all__ = (list(_holofunc_all) + list(_filters_all) + list(_utils_all) + list(_ccl_all) + list(_detection_all)
           + ["reconst_xyprojection_holo"])


def native_library_path() -> str:
    """Path of the loaded libholo.so (loads it)."""
    _lib.load()
    return str(_lib.LIB_PATH)
