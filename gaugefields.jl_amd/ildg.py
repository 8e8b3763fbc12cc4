"""ILDG/LIME configuration files through libgfhip's host-side reader and writer.

Host link arrays use the library's reference layout: (4, V, 9) complex128, element (i, j)
at i + 3*j, site x + NX*(y + NY*(z + NZ*t)).  File order and record types follow
ildg_format.jl:34-65, 697-746, 862-887.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .generators import volume

_D = ctypes.POINTER(ctypes.c_double)


def _path(p):
    return os.fsencode(os.fspath(p))


def _dims(dims):
    d = tuple(int(x) for x in dims)
    if len(d) != 4:
        raise ValueError(f"ILDG configurations are four dimensional; got extents {d}")
    return d, (ctypes.c_int * 4)(*d)


def ildg_info(path):
    """Header of an ILDG file: {"L": (lx, ly, lz, lt) or None, "precision": 32/64 or None,
    "payload_bytes": n} (None where the file has no ildg-format record)."""
    L = _lib.load()
    dims = (ctypes.c_int * 4)()
    prec = ctypes.c_int()
    nbytes = ctypes.c_longlong()
    _lib.check_ildg(L.gf_ildg_info(_path(path), dims, ctypes.byref(prec), ctypes.byref(nbytes)))
    ext = tuple(dims)
    return {
        "L": ext if ext[0] > 0 else None,
        "precision": prec.value if prec.value else None,
        "payload_bytes": nbytes.value,
    }


def read_ildg(path, dims):
    """(links, precision_bits) of the SU(3) configuration of extents dims stored in path."""
    d, cd = _dims(dims)
    U = np.empty((4, volume(d), 9), np.complex128)
    prec = ctypes.c_int()
    L = _lib.load()
    _lib.check_ildg(L.gf_ildg_read(_path(path), cd, U.ctypes.data_as(_D), ctypes.byref(prec)))
    return U, prec.value


def write_ildg(path, dims, links, precision=64):
    """Write links (4, V, 9) as an ILDG file of the given precision (32 or 64 bits)."""
    d, cd = _dims(dims)
    U = np.ascontiguousarray(links, dtype=np.complex128)
    if U.shape != (4, volume(d), 9):
        raise ValueError(f"links must have shape (4, {volume(d)}, 9); got {U.shape}")
    L = _lib.load()
    _lib.check_ildg(L.gf_ildg_write(_path(path), cd, U.ctypes.data_as(_D), int(precision)))
