"""emix3d_amd -- MI355X-native per-RK-stage spatial operator of eMix3d.

The product is libemix3d_hip.so (HIP kernels for gfx950 behind the C ABI of
include/emix3d.h); this package builds it, binds it (ctypes) and offers a
torch-tensor host interface named after the reference's Fortran routines.
"""
from .api import Batch, Operator, SCHEMES  # noqa: F401
from . import fields  # noqa: F401

__all__ = ["Batch", "Operator", "SCHEMES", "fields"]
