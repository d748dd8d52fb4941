"""MI355X-native vertex clustering (ACVD / ACVDQ hot path) on HIP for gfx950.

The clustering work runs in libacvd_amd.so (hand-written HIP kernels behind
the C-ABI of include/acvd_c.h); this package is a thin ctypes mirror of the
reference's class API.
"""
from ._native import AcvdError, METRIC_ISOTROPIC, METRIC_QEM, lib  # noqa: F401
from .remeshing import (Context, IsotropicDiscreteRemeshing, Mesh,  # noqa: F401
                        QIsotropicDiscreteRemeshing)
from . import meshgen  # noqa: F401

__all__ = ["AcvdError", "Context", "IsotropicDiscreteRemeshing", "QIsotropicDiscreteRemeshing", "Mesh", "meshgen",
           "METRIC_ISOTROPIC", "METRIC_QEM", "lib"]
