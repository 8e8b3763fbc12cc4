"""MI355X-native SU(3) gauge-link hot path for Gaugefields.jl (host-side mirror).

The compute path is libgfhip.so (hand-written HIP kernels for gfx950 behind the C ABI in
include/gfhip.h).  ``Lattice`` is the thin ctypes wrapper of one device-resident
configuration; ``api`` mirrors the reference's user-facing functions on top of it.
"""
from . import generators  # noqa: F401
from ._lib import GFError  # noqa: F401
from .lattice import OMELYAN_LAMBDA, Lattice  # noqa: F401
