"""Lattice: one device-resident SU(3) configuration with its momenta, behind libgfhip.

Host arrays use the reference's layouts (see generators.py): links (4, V, 9) complex128
with element (i, j) at i + 3*j; momenta and forces (4, V, 8) float64.
"""
import ctypes

import numpy as np

from . import _lib
from .generators import volume

_D = ctypes.POINTER(ctypes.c_double)

INTEGRATORS = {"QPQ": _lib.GF_QPQ, "PQP": _lib.GF_PQP, "2MN": _lib.GF_OMELYAN2MN}
# "f32force": links and momenta held and updated in FP64, every force evaluated in FP32 from the
# links rounded to float (GF_F32_FORCE)
PRECISIONS = {"f64": _lib.GF_F64, "f32": _lib.GF_F32, "f32force": _lib.GF_F32_FORCE}
#: Omelyan 2MN parameter (Omelyan, Mryglod, Folk, Comput. Phys. Commun. 151 (2003) 272)
OMELYAN_LAMBDA = 0.1931833275037836


def _dptr(a):
    return a.ctypes.data_as(_D)


class Lattice:
    """A 4D periodic SU(3) configuration on one GPU (one gf_ctx)."""

    def __init__(self, dims, precision="f64", device=0):
        self.dims = tuple(int(d) for d in dims)
        self.V = volume(self.dims)
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be 'f64', 'f32' or 'f32force'; got {precision!r}")
        self.precision = precision
        self._L = _lib.load()
        ctx = ctypes.c_void_p()
        d = (ctypes.c_int * 4)(*self.dims)
        rc = self._L.gf_create(d, PRECISIONS[precision], int(device), ctypes.byref(ctx))
        if rc != _lib.GF_OK:
            raise _lib.GFError(rc, f"gf_create({self.dims}, {precision}) failed")
        self._ctx = ctx
        self.nlev = 0

    # -- lifecycle --------------------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            self._L.gf_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        return _lib.check(self._ctx, rc)

    def sync(self):
        self._c(self._L.gf_sync(self._ctx))

    def device_bytes(self):
        n = ctypes.c_size_t()
        self._c(self._L.gf_info(self._ctx, None, None, ctypes.byref(n)))
        return n.value

    # -- fields -----------------------------------------------------------------------
    def _links_in(self, U):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        if U.shape != (4, self.V, 9):
            raise ValueError(f"links must have shape (4, {self.V}, 9); got {U.shape}")
        return U

    def _mom_in(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        if P.shape != (4, self.V, 8):
            raise ValueError(f"momenta must have shape (4, {self.V}, 8); got {P.shape}")
        return P

    def set_cold(self):
        self._c(self._L.gf_set_cold(self._ctx))

    def upload_links(self, U):
        U = self._links_in(U)
        self._c(self._L.gf_upload_links(self._ctx, _dptr(U)))

    def download_links(self):
        U = np.empty((4, self.V, 9), np.complex128)
        self._c(self._L.gf_download_links(self._ctx, _dptr(U)))
        return U

    def upload_momenta(self, P):
        P = self._mom_in(P)
        self._c(self._L.gf_upload_momenta(self._ctx, _dptr(P)))

    def download_momenta(self):
        P = np.empty((4, self.V, 8))
        self._c(self._L.gf_download_momenta(self._ctx, _dptr(P)))
        return P

    def clear_momenta(self):
        self._c(self._L.gf_clear_momenta(self._ctx))

    def download_force(self):
        F = np.empty((4, self.V, 8))
        self._c(self._L.gf_download_force(self._ctx, _dptr(F)))
        return F

    def save_links(self):
        self._c(self._L.gf_save_links(self._ctx))

    def restore_links(self):
        self._c(self._L.gf_restore_links(self._ctx))

    def load_ildg(self, path):
        """Read an ILDG file of this lattice's extents onto the device; returns its precision."""
        from .ildg import read_ildg

        U, prec = read_ildg(path, self.dims)
        self.upload_links(U)
        return prec

    def save_ildg(self, path, precision=64):
        """Write the device links to an ILDG file (precision 32 or 64 bits)."""
        from .ildg import write_ildg

        write_ildg(path, self.dims, self.download_links(), precision)

    def gaussian_momenta(self, seed, sweep=0, sigma=1.0):
        """P_a ~ N(0, sigma^2) generated on the device (Philox4x32-10 stream (seed, sweep))"""
        self._c(self._L.gf_gaussian_momenta(self._ctx, ctypes.c_double(sigma), ctypes.c_ulonglong(seed),
                                            ctypes.c_ulonglong(sweep)))

    def hot_start(self, seed):
        """random SU(3) links generated on the device (Philox4x32-10 stream seed)"""
        self._c(self._L.gf_hot_start(self._ctx, ctypes.c_ulonglong(seed)))

    def heatbath(self, beta, seed, sweep, itmax=100000):
        """one checkerboard heatbath sweep (plaquette action at beta); returns the number of
        links left unchanged because no proposal was accepted within itmax tries"""
        f = ctypes.c_longlong()
        self._c(self._L.gf_heatbath(self._ctx, ctypes.c_double(beta), ctypes.c_ulonglong(seed),
                                    ctypes.c_ulonglong(sweep), int(itmax), ctypes.byref(f)))
        return f.value

    def overrelaxation(self, seed, sweep):
        self._c(self._L.gf_overrelaxation(self._ctx, ctypes.c_ulonglong(seed), ctypes.c_ulonglong(sweep)))

    def links_from_algebra(self, omega):
        """U = exp(i sum_a omega_a lambda_a / 2) per link (the configs' seeded start)."""
        self.set_cold()
        self.upload_momenta(omega)
        self.exp_update(1.0)
        self.clear_momenta()

    # -- measurements -----------------------------------------------------------------
    def _scalar(self, fn, *args):
        out = ctypes.c_double()
        self._c(fn(self._ctx, *args, ctypes.byref(out)))
        return out.value

    def plaquette_sum(self):
        return self._scalar(self._L.gf_plaquette)

    def plaquette(self):
        return self.plaquette_sum() / (6 * self.V * 3)

    def action(self, coeff):
        return self._scalar(self._L.gf_action, ctypes.c_double(coeff))

    def action_untraced(self, coeff):
        """coeff sum_{mu<nu} (P_munu + P_munu^dag) per site, (V, 3, 3) complex [i, j]"""
        ref = np.empty((self.V, 3, 3), np.complex128, order="C")
        self._c(self._L.gf_action_untraced(self._ctx, ctypes.c_double(coeff), _dptr(ref)))
        # the host layout is column-major per site: element (i, j) at 3 j + i
        return ref.transpose(0, 2, 1).copy()

    def momenta_dot(self):
        return self._scalar(self._L.gf_momenta_dot)

    def hamiltonian(self, coeff):
        return self._scalar(self._L.gf_hamiltonian, ctypes.c_double(coeff))

    def energy_density(self):
        return self._scalar(self._L.gf_energy_density)

    # -- field strengths --------------------------------------------------------------
    TOPO_METHODS = {"plaquette": _lib.GF_TOPO_PLAQUETTE, "clover": _lib.GF_TOPO_CLOVER,
                    "improved": _lib.GF_TOPO_IMPROVED, "rectangle": _lib.GF_TOPO_RECTANGLE}

    def _topo(self, method):
        if method not in self.TOPO_METHODS:
            raise ValueError(f"method must be one of {list(self.TOPO_METHODS)}; got {method!r}")
        return self.TOPO_METHODS[method]

    def topological_charge(self, method="plaquette"):
        out = ctypes.c_double()
        self._c(self._L.gf_topological_charge(self._ctx, self._topo(method), ctypes.byref(out)))
        return out.value

    def topological_charge_density(self, method="plaquette"):
        """per-site density in site order x + NX*(y + NY*(z + NZ*t))"""
        q = np.empty(self.V)
        self._c(self._L.gf_topological_charge_density(self._ctx, self._topo(method), _dptr(q)))
        return q

    def energy_density_clover(self):
        return self._scalar(self._L.gf_energy_density_clover)

    def flow_measure(self, eps, nsteps, every):
        """rows {t, plaquette, E_plaq, E_clover, Q_plaq, Q_clover, Q_improved} every `every` steps"""
        n = nsteps // every if every > 0 else 0
        out = np.zeros((max(n, 1), _lib.GF_FLOW_NOBS))
        self._c(self._L.gf_flow_measure(self._ctx, ctypes.c_double(eps), int(nsteps), int(every), _dptr(out)))
        return out[:n]

    # -- derivative, force, updates -----------------------------------------------------
    def dsdu(self, coeff, download=True):
        out = np.empty((4, self.V, 9), np.complex128) if download else None
        self._c(self._L.gf_dsdu(self._ctx, ctypes.c_double(coeff), _dptr(out) if download else None))
        return out

    def force(self, coeff):
        self._c(self._L.gf_force(self._ctx, ctypes.c_double(coeff)))

    def force_add(self, coeff):
        """force field F += F(coeff) (on the smeared links when stout levels are configured)"""
        self._c(self._L.gf_stout_force_add(self._ctx, ctypes.c_double(coeff)))

    def force_kick(self, coeff, eps):
        self._c(self._L.gf_force_kick(self._ctx, ctypes.c_double(coeff), ctypes.c_double(eps)))

    def exp_update(self, eps):
        self._c(self._L.gf_exp_update(self._ctx, ctypes.c_double(eps)))

    def md_update(self, coeff, eps_p, eps_u):
        self._c(self._L.gf_md_update(self._ctx, ctypes.c_double(coeff), ctypes.c_double(eps_p),
                                     ctypes.c_double(eps_u)))

    def md_trajectory(self, coeff, steps, tau, integrator="QPQ", lam=OMELYAN_LAMBDA, fused=False,
                      diagnostics=True):
        if integrator not in INTEGRATORS:
            raise ValueError(f"integrator must be one of {list(INTEGRATORS)}; got {integrator!r}")
        code = INTEGRATORS[integrator] | (_lib.GF_FUSED if fused else 0)
        H = np.zeros(3)
        self._c(self._L.gf_md_trajectory(self._ctx, ctypes.c_double(coeff), code, ctypes.c_double(lam),
                                         int(steps), ctypes.c_double(tau), _dptr(H) if diagnostics else None))
        return tuple(H) if diagnostics else None

    # -- flow ---------------------------------------------------------------------------
    def flow(self, eps, nsteps, every=0):
        n = nsteps // every if every > 0 else 0
        E = np.zeros(max(n, 1))
        self._c(self._L.gf_flow(self._ctx, ctypes.c_double(eps), int(nsteps), int(every),
                                _dptr(E) if every > 0 else None))
        return E[:n]

    # -- stout ----------------------------------------------------------------------------
    def stout_config(self, nlev, rho, mode="store"):
        """mode: "store" (keep C per level), "recompute", or "auto" (store if it fits)"""
        m = {"store": _lib.GF_STOUT_STORE, "recompute": _lib.GF_STOUT_RECOMPUTE, "auto": _lib.GF_STOUT_AUTO}[mode]
        self._c(self._L.gf_stout_config(self._ctx, int(nlev), ctypes.c_double(rho), m))
        self.nlev = int(nlev)

    def stout_mode(self):
        m = ctypes.c_int()
        self._c(self._L.gf_stout_mode(self._ctx, ctypes.byref(m)))
        return {_lib.GF_STOUT_STORE: "store", _lib.GF_STOUT_RECOMPUTE: "recompute"}[m.value]

    def stout_smear(self, download=True):
        out = np.empty((4, self.V, 9), np.complex128) if download else None
        self._c(self._L.gf_stout_smear(self._ctx, _dptr(out) if download else None))
        return out

    def stout_backward(self, dout):
        dout = self._links_in(dout)
        din = np.empty_like(dout)
        self._c(self._L.gf_stout_backward(self._ctx, _dptr(dout), _dptr(din)))
        return din

    def stout_force(self, coeff):
        self._c(self._L.gf_stout_force(self._ctx, ctypes.c_double(coeff)))

    # -- timing ---------------------------------------------------------------------------
    def event_record(self, slot):
        self._c(self._L.gf_event_record(self._ctx, int(slot)))

    def event_elapsed_ms(self, a, b):
        ms = ctypes.c_float()
        self._c(self._L.gf_event_elapsed(self._ctx, int(a), int(b), ctypes.byref(ms)))
        return ms.value

    def stream_copy_rate(self, reps=20):
        """bytes/s (read + written) of the library's streaming copy of the link field"""
        r = ctypes.c_double()
        self._c(self._L.gf_stream_copy(self._ctx, int(reps), ctypes.byref(r)))
        return r.value

    def round_to_float(self, links=False, momenta=False, project=True):
        """FP64 context: round link rows 0-1 (SU(3)-projected in double first if project) and/or
        momenta to the nearest float in place (the FP32-bulk storage rounding)"""
        from ._lib import GF_ROUND_LINKS, GF_ROUND_MOMENTA, GF_ROUND_PROJECT

        f = (GF_ROUND_LINKS if links else 0) | (GF_ROUND_MOMENTA if momenta else 0)
        f |= GF_ROUND_PROJECT if (links and project) else 0
        self._c(self._L.gf_round_to_float(self._ctx, f))
