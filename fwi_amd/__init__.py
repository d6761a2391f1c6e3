"""fwi_amd -- MI355X-native FWI gradient path (coupled VTI pseudo-acoustic, 8th-order, C-PML).

The compute runs in fwi_amd/libfwi.so (hand-written HIP kernels for gfx950 behind a C ABI declared in
include/fwi_serialcalls.h and include/fwi.h).  This package is the Python host side: seeded workloads
(models), the ctypes binding (_lib), the batched survey plan (survey) and the torch.autograd wrapper
(autograd).  Importing the package does not load the library; the first call that needs it does, and
raises if it is missing.
"""
from . import models  # noqa: F401

__all__ = ["models", "load_library"]


def load_library():
    from . import _lib

    return _lib.load()
