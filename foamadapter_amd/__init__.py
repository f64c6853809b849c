"""foamadapter_amd -- MI355X-native finite-volume hot path of FoamAdapter/NeoN.

The numerical work runs in libfvhip.so (hand-written HIP kernels for gfx950
behind the C ABI in include/fvhip.h).  This package is the thin Python host
layer over that ABI, named after NeoN's fvcc / la / dsl interfaces, plus the
pressure-velocity coupling helpers (coupling) and the request server (serve).
"""
from ._lib import FvError, LIB_PATH, lib  # noqa: F401  (fails loudly if the library is missing)
from .runtime import Context, DeviceArray  # noqa: F401
from .mesh import BoundaryConditions, Mesh, PolyMesh, read_polymesh  # noqa: F401
from . import fvcc, la, dsl, coupling, serve  # noqa: F401

__all__ = ["Context", "DeviceArray", "Mesh", "PolyMesh", "BoundaryConditions", "read_polymesh",
           "fvcc", "la", "dsl", "coupling", "serve", "FvError", "LIB_PATH"]
