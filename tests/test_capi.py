"""CPU-side checks of the C ABI boundary: the library loads and exports what the header declares."""
import ctypes
import os
import re

import lattice_tools as lt

ROOT = lt.ROOT
HEADER = os.path.join(ROOT, "include", "gfhip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gf_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_binding_table(gf):
    from gaugefields_jl_amd import _lib

    assert declared_symbols() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol(gf):
    from gaugefields_jl_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert _lib.load().gf_version() == 1


def test_null_arguments_are_rejected_without_a_device(gf):
    from gaugefields_jl_amd import _lib

    L = _lib.load()
    assert L.gf_create(None, 0, 0, None) == _lib.GF_EINVAL
    dims = (ctypes.c_int * 4)(4, 4, 0, 4)
    ctx = ctypes.c_void_p()
    assert L.gf_create(dims, 0, 0, ctypes.byref(ctx)) == _lib.GF_EINVAL
    dims = (ctypes.c_int * 4)(4, 4, 4, 4)
    assert L.gf_create(dims, 7, 0, ctypes.byref(ctx)) == _lib.GF_EINVAL
    assert L.gf_destroy(None) == _lib.GF_EINVAL
    assert L.gf_plaquette(None, None) == _lib.GF_EINVAL


def test_generators_layout(gf):
    gen = gf.generators
    U = gen.hot_start_links((2, 2, 2, 2), 3)
    M = gen.links_to_matrices(U)
    eye = np.eye(3)
    assert np.abs(np.conj(np.swapaxes(M, -1, -2)) @ M - eye).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14
    assert np.array_equal(gen.matrices_to_links(M), U)
    # column-major: element (i, j) at i + 3 j
    assert U[0, 0, 1] == M[0, 0, 1, 0]


import numpy as np  # noqa: E402
