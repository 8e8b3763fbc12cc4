"""ctypes binding of the CPU oracle (oracle/_build/libgforacle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the oracle.  The product (libgfhip) never uses it.

All arrays are numpy float64 in the reference's Julia layouts:
  links   (4, V, 9) complex128, element (i, j) at i + 3*j (column-major 3x3),
  momenta (4, V, 8) float64, coefficients of sum_a p_a (i lambda_a / 2),
with site index s = x + NX*(y + NY*(z + NZ*t)).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libgforacle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "gfo_set_threads": (None, [ctypes.c_int]),
            "gfo_get_threads": (ctypes.c_int, []),
            "gfo_ta_coeffs": (None, [_D, _D]),
            "gfo_ta_matrix": (None, [_D, _D]),
            "gfo_exp_ta": (None, [_D, ctypes.c_double, _D]),
            "gfo_exp_antiherm": (None, [_D, _D]),
            "gfo_cdexpqdq": (None, [_D, _D, _D]),
            "gfo_plaquette_sum": (ctypes.c_double, [_I, _D]),
            "gfo_dsdu": (None, [_I, _D, ctypes.c_double, _D]),
            "gfo_force": (None, [_I, _D, ctypes.c_double, _D]),
            "gfo_kick": (None, [_I, _D, _D, ctypes.c_double]),
            "gfo_exp_update": (None, [_I, _D, _D, ctypes.c_double]),
            "gfo_momenta_dot": (ctypes.c_double, [_I, _D]),
            "gfo_md_update_unfused": (None, [_I, _D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "gfo_potential": (ctypes.c_double, [_I, _D, ctypes.c_double]),
            "gfo_flow_force": (None, [_I, _D, _D]),
            "gfo_flow": (None, [_I, _D, ctypes.c_double, ctypes.c_int, ctypes.c_int, _D]),
            "gfo_energy_density": (ctypes.c_double, [_I, _D]),
            "gfo_stout_forward": (None, [_I, _D, ctypes.c_double, _D, _D]),
            "gfo_stout_backward": (None, [_I, _D, ctypes.c_double, _D, _D]),
            "gfo_stout_smear": (None, [_I, _D, ctypes.c_int, ctypes.c_double, _D]),
            "gfo_stout_force": (None, [_I, _D, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D]),
            "gfo_stout_potential": (ctypes.c_double, [_I, _D, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
            "gfo_topological_charge_density": (ctypes.c_double, [_I, _D, ctypes.c_int, _D]),
            "gfo_energy_density_clover": (ctypes.c_double, [_I, _D]),
            "gfo_philox4x32_10": (None, [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                         ctypes.POINTER(ctypes.c_uint32)]),
            "gfo_gaussian_momenta": (None, [_I, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64, _D]),
            "gfo_hot_links": (None, [_I, ctypes.c_uint64, _D]),
            "gfo_heatbath_sweep": (ctypes.c_long, [_I, _D, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64,
                                                   ctypes.c_int, ctypes.c_int]),
            "gfo_kp_samples": (None, [ctypes.c_double, ctypes.c_long, ctypes.c_uint64, _D]),
            "gfo_md_trajectory": (None, [_I, _D, _D, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dims(dims):
    return (ctypes.c_int * 4)(*[int(d) for d in dims])


def _p(a):
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(_D)


def _links(U):
    U = np.ascontiguousarray(U, dtype=np.complex128)
    return U


def set_threads(n):
    lib().gfo_set_threads(int(n))


def get_threads():
    return lib().gfo_get_threads()


# --- single matrices: m is a (3,3) complex array (row i, col j) ---
def _m_in(m):
    return np.ascontiguousarray(np.asarray(m, dtype=np.complex128).T)  # column-major


def _m_out(buf):
    return buf.T.copy()


def ta_coeffs(m):
    out = np.zeros(8)
    lib().gfo_ta_coeffs(_p(_m_in(m)), _p(out))
    return out


def ta_matrix(m):
    out = np.zeros((3, 3), np.complex128)
    lib().gfo_ta_matrix(_p(_m_in(m)), _p(out))
    return _m_out(out)


def exp_ta(c8, t=1.0):
    out = np.zeros((3, 3), np.complex128)
    lib().gfo_exp_ta(_p(np.ascontiguousarray(c8, dtype=np.float64)), float(t), _p(out))
    return _m_out(out)


def exp_antiherm(q):
    out = np.zeros((3, 3), np.complex128)
    lib().gfo_exp_antiherm(_p(_m_in(q)), _p(out))
    return _m_out(out)


def cdexpqdq(c, q):
    out = np.zeros((3, 3), np.complex128)
    lib().gfo_cdexpqdq(_p(_m_in(c)), _p(_m_in(q)), _p(out))
    return _m_out(out)


# --- lattice fields ---
def plaquette_sum(dims, U):
    return lib().gfo_plaquette_sum(_dims(dims), _p(_links(U)))


def dsdu(dims, U, coeff):
    U = _links(U)
    out = np.zeros_like(U)
    lib().gfo_dsdu(_dims(dims), _p(U), float(coeff), _p(out))
    return out


def force(dims, U, coeff):
    U = _links(U)
    F = np.zeros((4, U.shape[1], 8))
    lib().gfo_force(_dims(dims), _p(U), float(coeff), _p(F))
    return F


def kick(dims, P, F, eps):
    lib().gfo_kick(_dims(dims), _p(P), _p(np.ascontiguousarray(F)), float(eps))


def exp_update(dims, U, P, eps):
    lib().gfo_exp_update(_dims(dims), _p(U), _p(np.ascontiguousarray(P)), float(eps))


def md_update_unfused(dims, U, P, coeff, eps_p, eps_u):
    """update_momenta!(eps_p) then update_gaugefields!(eps_u) in the reference's pass structure,
    in place on U and P (C-contiguous complex128 / float64)"""
    assert U.flags["C_CONTIGUOUS"] and P.flags["C_CONTIGUOUS"]
    lib().gfo_md_update_unfused(_dims(dims), _p(U), _p(P), float(coeff), float(eps_p), float(eps_u))


def momenta_dot(dims, P):
    return lib().gfo_momenta_dot(_dims(dims), _p(np.ascontiguousarray(P)))


def potential(dims, U, coeff):
    return lib().gfo_potential(_dims(dims), _p(_links(U)), float(coeff))


def flow_force(dims, U):
    U = _links(U)
    F = np.zeros((4, U.shape[1], 8))
    lib().gfo_flow_force(_dims(dims), _p(U), _p(F))
    return F


def flow(dims, U, eps, nsteps, every=0):
    nE = nsteps // every if every > 0 else 0
    E = np.zeros(max(nE, 1))
    lib().gfo_flow(_dims(dims), _p(U), float(eps), int(nsteps), int(every), _p(E))
    return E[:nE]


def energy_density(dims, U):
    return lib().gfo_energy_density(_dims(dims), _p(_links(U)))


TOPO_METHODS = {"plaquette": 0, "clover": 1, "improved": 2, "rectangle": 3}


def topological_charge_density(dims, U, method="plaquette"):
    """(Q, q) with q the per-site density in site order x + NX(y + NY(z + NZ t))"""
    q = np.zeros(int(np.prod(dims)))
    Q = lib().gfo_topological_charge_density(_dims(dims), _p(_links(U)), TOPO_METHODS[method], _p(q))
    return Q, q


def energy_density_clover(dims, U):
    return lib().gfo_energy_density_clover(_dims(dims), _p(_links(U)))


def stout_forward(dims, U, rho, want_C=False):
    U = _links(U)
    out = np.zeros_like(U)
    C = np.zeros_like(U) if want_C else None
    lib().gfo_stout_forward(_dims(dims), _p(U), float(rho), _p(out), _p(C) if want_C else None)
    return (out, C) if want_C else out


def stout_backward(dims, Uin, rho, Dout):
    Uin = _links(Uin)
    Din = np.zeros_like(Uin)
    lib().gfo_stout_backward(_dims(dims), _p(Uin), float(rho), _p(_links(Dout)), _p(Din))
    return Din


def stout_smear(dims, U, nlev, rho):
    U = _links(U)
    out = np.zeros_like(U)
    lib().gfo_stout_smear(_dims(dims), _p(U), int(nlev), float(rho), _p(out))
    return out


def stout_force(dims, U, nlev, rho, coeff):
    U = _links(U)
    F = np.zeros((4, U.shape[1], 8))
    lib().gfo_stout_force(_dims(dims), _p(U), int(nlev), float(rho), float(coeff), _p(F))
    return F


def stout_potential(dims, U, nlev, rho, coeff):
    return lib().gfo_stout_potential(_dims(dims), _p(_links(U)), int(nlev), float(rho), float(coeff))


# "2MN-merged": the Omelyan 2MN sequence with the adjacent kicks of consecutive steps applied as
# one kick (the same map; one force evaluation per step fewer)
INTEGRATORS = {"QPQ": 0, "PQP": 1, "2MN": 2, "2MN-merged": 3}


def md_trajectory(dims, U, P, coeff, steps, tau, integrator="QPQ", lam=0.1931833275037836,
                  nlev=0, rho=0.0):
    """In place on U (complex128 (4,V,9)) and P (float64 (4,V,8)); returns (H0, H1, dH)."""
    H = np.zeros(3)
    lib().gfo_md_trajectory(_dims(dims), _p(U), _p(P), float(coeff), int(nlev), float(rho),
                            INTEGRATORS[integrator], float(lam), int(steps), float(tau), _p(H))
    return tuple(H)


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().gfo_philox4x32_10(c, k, o)
    return tuple(o)


def gaussian_momenta(dims, seed, sweep=0, sigma=1.0):
    P = np.zeros((4, int(np.prod(dims)), 8))
    lib().gfo_gaussian_momenta(_dims(dims), float(sigma), int(seed), int(sweep), _p(P))
    return P


def hot_links(dims, seed):
    U = np.zeros((4, int(np.prod(dims)), 9), np.complex128)
    lib().gfo_hot_links(_dims(dims), int(seed), _p(U))
    return U


def heatbath_sweep(dims, U, beta, seed, sweep, itmax=100000, overrelax=False):
    """in place on U; returns the number of links a heatbath left unchanged"""
    assert U.flags["C_CONTIGUOUS"] and U.dtype == np.complex128
    return lib().gfo_heatbath_sweep(_dims(dims), _p(U), float(beta), int(seed), int(sweep), int(itmax),
                                    1 if overrelax else 0)


def kp_samples(k, n, seed):
    a0 = np.zeros(int(n))
    lib().gfo_kp_samples(float(k), int(n), int(seed), _p(a0))
    return a0
