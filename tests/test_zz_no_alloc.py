"""gf_plaquette_links, the entry point that allocates nothing, against gf_plaquette (collected
last: a failure here hides no other test).

The caller (this test) holds the links in device memory in the library's device layout,
allocated through the HIP runtime itself; the entry point keeps its partial sums in a static
device array.  Two calls in a row, the second on a smaller lattice, in FP64 and FP32, each
compared with gf_plaquette of a context holding the same links and with the oracle.
"""
import ctypes

import numpy as np
import pytest

import lattice_tools as lt

gen = lt.gen

H2D, D2H = 1, 2


def hip():
    """The HIP runtime libgfhip is linked against (already loaded with it)."""
    from gaugefields_jl_amd import _lib

    _lib.load()
    R = ctypes.CDLL("libamdhip64.so.7")
    R.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    R.hipFree.argtypes = [ctypes.c_void_p]
    R.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    R.hipDeviceSynchronize.argtypes = []
    for f in (R.hipMalloc, R.hipFree, R.hipMemcpy, R.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return R


def device_layout(U, precision):
    """(4, V, 9) complex reference layout (element (i, j) at i + 3 j) -> the device layout of
    an SU(3) link field: FP64 double[4][6][V][2] (plane 3 i + j of rows i = 0, 1), FP32
    float[4][3][V][4] (quad plane q: elements 2q and 2q+1 of 3 i + j)."""
    V = U.shape[1]
    rows = np.empty((4, 6, V), np.complex128)
    for i in range(2):
        for j in range(3):
            rows[:, 3 * i + j, :] = U[:, :, i + 3 * j]
    if precision == "f64":
        return np.ascontiguousarray(np.stack([rows.real, rows.imag], axis=-1))
    q = rows.reshape(4, 3, 2, V)  # quad plane q holds elements 2q, 2q+1
    out = np.stack([q[:, :, 0].real, q[:, :, 0].imag, q[:, :, 1].real, q[:, :, 1].imag], axis=-1)
    return np.ascontiguousarray(out.astype(np.float32))


def plaquette_links(R, U, dims, precision):
    """one gf_plaquette_links call on device copies of U; returns the device result"""
    from gaugefields_jl_amd import _lib

    L = _lib.load()
    host = device_layout(U, precision)
    d_links, d_sum = ctypes.c_void_p(), ctypes.c_void_p()
    assert R.hipMalloc(ctypes.byref(d_links), host.nbytes) == 0
    assert R.hipMalloc(ctypes.byref(d_sum), 8) == 0
    try:
        assert R.hipMemcpy(d_links, host.ctypes.data, host.nbytes, H2D) == 0
        cd = (ctypes.c_int * 4)(*dims)
        prec = _lib.GF_F64 if precision == "f64" else _lib.GF_F32
        rc = L.gf_plaquette_links(d_links, cd, prec, None, d_sum)
        assert rc == _lib.GF_OK, L.gf_last_error(None)
        assert R.hipDeviceSynchronize() == 0
        out = np.zeros(1)
        assert R.hipMemcpy(out.ctypes.data, d_sum, 8, D2H) == 0
        return float(out[0])
    finally:
        R.hipFree(d_links)
        R.hipFree(d_sum)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_plaquette_links_matches_gf_plaquette(gf, oracle, precision):
    R = hip()
    # 8x8x8x16 takes the t-sweep tiled block order, 4x4x4x6 the linear one
    for k, dims in enumerate([(8, 8, 8, 16), (4, 4, 4, 6)]):
        print(f"no-alloc plaquette {precision} call {k} dims {dims}", flush=True)
        L = gf.Lattice(dims, precision=precision)
        L.links_from_algebra(gen.algebra_coefficients(dims, 0.3, 700 + k))
        ref = L.plaquette_sum()  # gf_plaquette: the context's own partials buffer
        U = L.download_links()   # FP32: the float-stored links, widened exactly
        L.close()
        got = plaquette_links(R, U, dims, precision)
        assert abs(got - ref) <= 1e-12 * abs(ref), (dims, got, ref)
        orc = oracle.plaquette_sum(dims, U)
        assert abs(got - orc) <= 1e-12 * abs(orc), (dims, got, orc)


def test_plaquette_links_rejects_what_the_scratch_cannot_hold(gf):
    """Host-side checks, before any HIP call: sizes beyond the static scratch (2^25 sites),
    bad precision and null pointers are GF_EINVAL with a message on gf_last_error(NULL)."""
    from gaugefields_jl_amd import _lib

    L = _lib.load()
    big = (ctypes.c_int * 4)(128, 128, 128, 64)  # 2^27 sites
    rc = L.gf_plaquette_links(ctypes.c_void_p(256), big, _lib.GF_F32, None, ctypes.c_void_p(256))
    assert rc == _lib.GF_EINVAL
    assert b"static scratch" in L.gf_last_error(None)
    small = (ctypes.c_int * 4)(4, 4, 4, 4)
    assert L.gf_plaquette_links(ctypes.c_void_p(256), small, 5, None, ctypes.c_void_p(256)) == _lib.GF_EINVAL
    assert L.gf_plaquette_links(None, small, _lib.GF_F64, None, ctypes.c_void_p(256)) == _lib.GF_EINVAL
    assert L.gf_plaquette_links(ctypes.c_void_p(256), small, _lib.GF_F64, None, None) == _lib.GF_EINVAL
