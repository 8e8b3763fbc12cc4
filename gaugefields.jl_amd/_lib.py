"""ctypes binding of libgfhip.so (the HIP C ABI declared in include/gfhip.h).

There is no CPU fallback: if the library is missing or cannot be loaded, every call
raises.  The library is built in-tree by ``__graft_entry__.build()`` (hipcc, gfx950).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgfhip.so")

GF_F64, GF_F32, GF_F32_FORCE = 0, 1, 2
GF_OK, GF_EINVAL, GF_EDEVICE, GF_ENOMEM, GF_ESTATE = 0, -1, -2, -3, -4
GF_QPQ, GF_PQP, GF_OMELYAN2MN, GF_FUSED = 0, 1, 2, 0x100
GF_STOUT_STORE, GF_STOUT_RECOMPUTE, GF_STOUT_AUTO = 0, 1, 2
GF_TOPO_PLAQUETTE, GF_TOPO_CLOVER, GF_TOPO_IMPROVED, GF_TOPO_RECTANGLE = 0, 1, 2, 3
GF_FLOW_NOBS = 7
GF_ROUND_LINKS, GF_ROUND_PROJECT, GF_ROUND_MOMENTA = 1, 2, 4

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_F = ctypes.POINTER(ctypes.c_float)
_CTX = ctypes.c_void_p

#: every symbol the header declares, with (restype, argtypes)
SIGNATURES = {
    "gf_version": (ctypes.c_int, []),
    "gf_create": (ctypes.c_int, [_I, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_CTX)]),
    "gf_destroy": (ctypes.c_int, [_CTX]),
    "gf_last_error": (ctypes.c_char_p, [_CTX]),
    "gf_sync": (ctypes.c_int, [_CTX]),
    "gf_info": (ctypes.c_int, [_CTX, _I, _I, ctypes.POINTER(ctypes.c_size_t)]),
    "gf_set_cold": (ctypes.c_int, [_CTX]),
    "gf_upload_links": (ctypes.c_int, [_CTX, _D]),
    "gf_download_links": (ctypes.c_int, [_CTX, _D]),
    "gf_upload_momenta": (ctypes.c_int, [_CTX, _D]),
    "gf_download_momenta": (ctypes.c_int, [_CTX, _D]),
    "gf_clear_momenta": (ctypes.c_int, [_CTX]),
    "gf_download_force": (ctypes.c_int, [_CTX, _D]),
    "gf_save_links": (ctypes.c_int, [_CTX]),
    "gf_restore_links": (ctypes.c_int, [_CTX]),
    "gf_gaussian_momenta": (ctypes.c_int, [_CTX, ctypes.c_double, ctypes.c_ulonglong, ctypes.c_ulonglong]),
    "gf_hot_start": (ctypes.c_int, [_CTX, ctypes.c_ulonglong]),
    "gf_heatbath": (ctypes.c_int, [_CTX, ctypes.c_double, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_longlong)]),
    "gf_overrelaxation": (ctypes.c_int, [_CTX, ctypes.c_ulonglong, ctypes.c_ulonglong]),
    "gf_plaquette": (ctypes.c_int, [_CTX, _D]),
    "gf_plaquette_links": (ctypes.c_int, [ctypes.c_void_p, _I, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gf_action": (ctypes.c_int, [_CTX, ctypes.c_double, _D]),
    "gf_action_untraced": (ctypes.c_int, [_CTX, ctypes.c_double, _D]),
    "gf_momenta_dot": (ctypes.c_int, [_CTX, _D]),
    "gf_hamiltonian": (ctypes.c_int, [_CTX, ctypes.c_double, _D]),
    "gf_dsdu": (ctypes.c_int, [_CTX, ctypes.c_double, _D]),
    "gf_force": (ctypes.c_int, [_CTX, ctypes.c_double]),
    "gf_force_kick": (ctypes.c_int, [_CTX, ctypes.c_double, ctypes.c_double]),
    "gf_exp_update": (ctypes.c_int, [_CTX, ctypes.c_double]),
    "gf_md_update": (ctypes.c_int, [_CTX, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "gf_md_trajectory": (ctypes.c_int, [_CTX, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_double, _D]),
    "gf_flow": (ctypes.c_int, [_CTX, ctypes.c_double, ctypes.c_int, ctypes.c_int, _D]),
    "gf_energy_density": (ctypes.c_int, [_CTX, _D]),
    "gf_topological_charge": (ctypes.c_int, [_CTX, ctypes.c_int, _D]),
    "gf_topological_charge_density": (ctypes.c_int, [_CTX, ctypes.c_int, _D]),
    "gf_energy_density_clover": (ctypes.c_int, [_CTX, _D]),
    "gf_flow_measure": (ctypes.c_int, [_CTX, ctypes.c_double, ctypes.c_int, ctypes.c_int, _D]),
    "gf_stout_config": (ctypes.c_int, [_CTX, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "gf_stout_mode": (ctypes.c_int, [_CTX, _I]),
    "gf_stout_smear": (ctypes.c_int, [_CTX, _D]),
    "gf_stout_backward": (ctypes.c_int, [_CTX, _D, _D]),
    "gf_stout_force": (ctypes.c_int, [_CTX, ctypes.c_double]),
    "gf_stout_force_kick": (ctypes.c_int, [_CTX, ctypes.c_double, ctypes.c_double]),
    "gf_stout_force_add": (ctypes.c_int, [_CTX, ctypes.c_double]),
    "gf_ildg_info": (ctypes.c_int, [ctypes.c_char_p, _I, _I, ctypes.POINTER(ctypes.c_longlong)]),
    "gf_ildg_read": (ctypes.c_int, [ctypes.c_char_p, _I, _D, _I]),
    "gf_ildg_write": (ctypes.c_int, [ctypes.c_char_p, _I, _D, ctypes.c_int]),
    "gf_ildg_last_error": (ctypes.c_char_p, []),
    "gf_round_to_float": (ctypes.c_int, [_CTX, ctypes.c_int]),
    "gf_stream_copy": (ctypes.c_int, [_CTX, ctypes.c_int, _D]),
    "gf_event_record": (ctypes.c_int, [_CTX, ctypes.c_int]),
    "gf_event_elapsed": (ctypes.c_int, [_CTX, ctypes.c_int, ctypes.c_int, _F]),
}

_lib = None


class GFError(RuntimeError):
    """A non-zero status from libgfhip."""

    def __init__(self, code, msg):
        super().__init__(f"libgfhip error {code}: {msg}")
        self.code = code


def load():
    """Load libgfhip.so (raises if it is absent: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libgfhip.so not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check_ildg(code):
    """Raise GFError with the ILDG layer's message for a non-zero status."""
    if code != GF_OK:
        msg = load().gf_ildg_last_error()
        raise GFError(code, msg.decode() if msg else "")
    return code


def check(ctx, code):
    if code != GF_OK:
        msg = load().gf_last_error(ctx)
        raise GFError(code, msg.decode() if msg else "")
    return code
