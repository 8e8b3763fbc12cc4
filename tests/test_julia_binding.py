"""The Julia binding (gaugefields.jl_amd/julia/GaugefieldsHIP.jl) against the C ABI.

The image has no Julia, so the binding is checked by reading it:

* every ccall names a symbol include/gfhip.h declares, with the same number of arguments and
  argument types that match the prototype (CPU);
* each method of the binding has a Python replay below that makes the same libgfhip calls with
  the same arguments in the same order; the replay's recorded call sequence is compared with
  the sequence read from the Julia method (CPU), and the replays run against a device context
  and are compared with the oracle (GPU);
* the cotangent conversion of the AD rules (ChainRules/Enzyme convention <-> the library's
  dS/dU convention) is checked against central finite differences of a real loss (CPU, oracle).
"""
import ctypes

import numpy as np
import pytest
from scipy.linalg import expm

import julia_ccalls as jc
import lattice_tools as lt

gen = lt.gen
AUTO, STORE, RECOMPUTE = 2, 0, 1
LAM = 0.1931833275037836
GF_FUSED = 0x100
TOPO = {"plaquette": 0, "clover": 1, "improved": 2}


def test_every_ccall_matches_a_header_prototype():
    protos = jc.header_prototypes()
    calls = jc.julia_ccalls()
    assert len(calls) > 30
    for sym, ret, types, args in calls:
        assert sym in protos, sym
        pret, pkinds = protos[sym]
        assert {"Cint": "int", "Cstring": "cstr"}[ret] == pret, (sym, ret)
        assert [jc.julia_kind(t) for t in types] == pkinds, (sym, types, pkinds)
        assert len(args) == len(types), (sym, args, types)


def test_parser_reads_nesting_strings_and_comments():
    text = '''
function _a(c)
    ccall((:gf_sync, libgfhip), Cint, (Ptr{Cvoid},), c)  # gf_destroy( in a comment
end
function outer(c)
    msg = "ccall((:gf_info, libgfhip)"
    _b(c, _a(c))
end
_b(c, x) = ccall((:gf_plaquette, libgfhip), Cint, (Ptr{Cvoid}, Ref{Float64}), c, x)
'''
    seqs = jc.method_sequences(text)
    assert seqs["function outer(c)"] == ["gf_sync", "gf_plaquette"]


# ----------------------------------------------------------------------------- replays


class Rec:
    """libgfhip seen through a recorder: every gf_ call is appended to .calls and, when live,
    made (otherwise it returns GF_OK without running)."""

    def __init__(self, live):
        from gaugefields_jl_amd import _lib

        self.lib = _lib.load()
        self.live = live
        self.calls = []

    def __getattr__(self, name):
        f = getattr(self.lib, name)

        def call(*args):
            self.calls.append(name)
            return f(*args) if self.live else 0

        return call


class Ctx:
    """the Julia HIPContext: pointer, extents"""

    def __init__(self, ptr, dims):
        self.ptr = ptr
        self.dims = tuple(dims)
        self.V = int(np.prod(dims))


def _ok(rc):
    assert rc == 0, rc


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _dims(c):
    return (ctypes.c_int * 4)(*c.dims)


def dag(F):
    """per-site Hermitian conjugate of a (..., 9) field in the link layout (i + 3 j)"""
    T = F.reshape(F.shape[:-1] + (3, 3))
    return np.conj(np.swapaxes(T, -1, -2)).reshape(F.shape)


# one function per Julia method: the same calls, arguments and order


def jl_calculate_Plaquette(R, c):  # calculate_Plaquette(U::HIPLinks)
    s = ctypes.c_double(0)
    _ok(R.gf_plaquette(c.ptr, ctypes.byref(s)))
    return s.value


def jl_measure_plaquette(R, c):  # measure_plaquette(U::HIPLinks; normalize=true)
    return jl_calculate_Plaquette(R, c) / (6 * c.V * 3)


def jl_evaluate_GaugeAction(R, c, coeff):
    _ok(R.gf_stout_config(c.ptr, 0, 0.0, AUTO))
    s = ctypes.c_double(0)
    _ok(R.gf_action(c.ptr, coeff, ctypes.byref(s)))
    return s.value


def jl_calc_dSdU(R, c, coeff, mu):  # calc_dSdUμ!(dSdUμ, action, μ, U)
    _ok(R.gf_stout_config(c.ptr, 0, 0.0, AUTO))
    buf = np.empty(36 * c.V, np.complex128)
    _ok(R.gf_dsdu(c.ptr, coeff, _d(buf)))
    return buf.reshape(4, c.V, 9)[mu - 1]


def jl_gauss_distribution(R, c, sigma, seed, sweep):
    _ok(R.gf_gaussian_momenta(c.ptr, sigma, seed, sweep))


def jl_clear_momenta(R, c):  # clear_U!(P::HIPMomenta)
    _ok(R.gf_clear_momenta(c.ptr))


def jl_pp(R, c):  # Base.:*(P, P)
    s = ctypes.c_double(0)
    _ok(R.gf_momenta_dot(c.ptr, ctypes.byref(s)))
    return s.value


def jl_substitute_momenta(R, c, P):  # substitute_U!(P::HIPMomenta, Pref)
    P = np.ascontiguousarray(P, np.float64)
    _ok(R.gf_upload_momenta(c.ptr, _d(P)))


def jl_Array_momenta(R, c, slot):  # Base.Array(P::HIPMomenta)
    buf = np.empty(32 * c.V)
    if slot == "momenta":
        _ok(R.gf_download_momenta(c.ptr, _d(buf)))
    else:
        _ok(R.gf_download_force(c.ptr, _d(buf)))
    return buf.reshape(4, c.V, 8)


def jl_substitute_links(R, c, U):  # substitute_U!(U::HIPLinks, Uref) with host arrays
    U = np.ascontiguousarray(U, np.complex128)
    _ok(R.gf_upload_links(c.ptr, _d(U)))


def jl_Array_links(R, c):  # Base.Array(U::HIPLinks)
    buf = np.empty(36 * c.V, np.complex128)
    _ok(R.gf_download_links(c.ptr, _d(buf)))
    return buf.reshape(4, c.V, 9)


def jl_md_potential(R, c, coeff):
    _ok(R.gf_stout_config(c.ptr, 0, 0.0, AUTO))
    H, s = ctypes.c_double(0), ctypes.c_double(0)
    _ok(R.gf_hamiltonian(c.ptr, coeff, ctypes.byref(H)))
    _ok(R.gf_momenta_dot(c.ptr, ctypes.byref(s)))
    return H.value - s.value / 2


def jl_md_force(R, c, coeff):
    _ok(R.gf_stout_config(c.ptr, 0, 0.0, AUTO))
    _ok(R.gf_stout_force(c.ptr, coeff))


def jl_evaluate_GaugeAction_untraced(R, c, coeff):  # (3, 3, NX, NY, NZ, NT) column-major
    out = np.empty(9 * c.V, np.complex128)
    _ok(R.gf_action_untraced(c.ptr, coeff, _d(out)))
    return out.reshape(c.V, 3, 3).transpose(0, 2, 1)  # site-major (V, i, j)


def jl_md_potential_set(R, c, coeffs):  # md_potential(::MDActionSet) of thin plaquette members
    return sum(jl_md_potential(R, c, k) for k in coeffs)


def jl_md_force_set(R, c, coeffs):  # md_force!(force, ::MDActionSet, U, ws[, group]), thin members
    for k, coeff in enumerate(coeffs):
        _ok(R.gf_stout_config(c.ptr, 0, 0.0, AUTO))
        _ok(R.gf_stout_force(c.ptr, coeff) if k == 0 else R.gf_stout_force_add(c.ptr, coeff))


def jl_update_momenta_set(R, c, coeffs, eps):  # update_momenta!(P, U, eps, driver, group)
    for coeff in coeffs:
        jl_update_momenta(R, c, coeff, eps)


def jl_md_hamiltonian(R, c, coeff):
    _ok(R.gf_stout_config(c.ptr, 0, 0.0, AUTO))
    H = ctypes.c_double(0)
    _ok(R.gf_hamiltonian(c.ptr, coeff, ctypes.byref(H)))
    return H.value


def jl_update_gaugefields(R, c, eps):
    _ok(R.gf_exp_update(c.ptr, eps))


def jl_update_momenta(R, c, coeff, eps):
    _ok(R.gf_stout_config(c.ptr, 0, 0.0, AUTO))
    _ok(R.gf_stout_force_kick(c.ptr, coeff, eps))


def jl_md_step_omelyan(R, c, coeff, eps, lam=LAM):  # md_step!(::OmelyanPQP, U, P, ε, driver)
    jl_update_momenta(R, c, coeff, lam * eps)
    jl_update_gaugefields(R, c, eps / 2)
    jl_update_momenta(R, c, coeff, (1 - 2 * lam) * eps)
    jl_update_gaugefields(R, c, eps / 2)
    jl_update_momenta(R, c, coeff, lam * eps)


def jl_md_trajectory(R, c, coeff, code, lam, steps, tau, fused=False):  # native integrators
    _ok(R.gf_stout_config(c.ptr, 0, 0.0, AUTO))
    H = np.zeros(3)
    _ok(R.gf_md_trajectory(c.ptr, coeff, code | (GF_FUSED if fused else 0), lam, steps, tau, _d(H)))
    return H


def jl_flow(R, c, eps, nflow):  # flow!(U, g::Gradientflow)
    E = np.zeros(1)
    _ok(R.gf_flow(c.ptr, eps, nflow, 0, _d(E)))


def jl_calc_smearedU(R, c, nlev, rho, mode):
    _ok(R.gf_stout_config(c.ptr, nlev, rho, mode))
    host = np.empty(36 * c.V, np.complex128)
    _ok(R.gf_stout_smear(c.ptr, _d(host)))
    return host.reshape(4, c.V, 9)


def _stout_backward(R, c, D):
    dout = np.ascontiguousarray(D, np.complex128)
    din = np.empty(36 * c.V, np.complex128)
    _ok(R.gf_stout_backward(c.ptr, _d(dout), _d(din)))
    return din.reshape(4, c.V, 9)


def jl_back_prop(R, c, nlev, rho, mode, D):  # back_prop(δL, nn, Uout_multi, Uin)
    _ok(R.gf_stout_config(c.ptr, nlev, rho, mode))
    _ok(R.gf_stout_smear(c.ptr, None))
    return _stout_backward(R, c, D)


def jl_stout_smear(R, c, nlev, rho, mode, links):  # the AD map
    jl_substitute_links(R, c, links)
    return jl_calc_smearedU(R, c, nlev, rho, mode)


def jl_stout_pullback(R, c, nlev, rho, mode, links, Ubar):
    jl_substitute_links(R, c, links)
    _ok(R.gf_stout_config(c.ptr, nlev, rho, mode))
    _ok(R.gf_stout_smear(c.ptr, None))
    D = dag(Ubar) / 2
    return 2 * dag(_stout_backward(R, c, D))


def jl_topological_charge(R, c, method):
    Q = ctypes.c_double(0)
    _ok(R.gf_topological_charge(c.ptr, TOPO[method], ctypes.byref(Q)))
    return Q.value


def jl_topological_charge_density(R, c, method):
    q = np.empty(c.V)
    _ok(R.gf_topological_charge_density(c.ptr, TOPO[method], _d(q)))
    return q


def jl_calculate_energy_density(R, c):  # clover
    E = ctypes.c_double(0)
    _ok(R.gf_energy_density_clover(c.ptr, ctypes.byref(E)))
    return E.value


def jl_energy_density(R, c):  # plaquette
    E = ctypes.c_double(0)
    _ok(R.gf_energy_density(c.ptr, ctypes.byref(E)))
    return E.value


def jl_flow_measure(R, c, eps, nsteps, every):
    out = np.empty(7 * (nsteps // every))
    _ok(R.gf_flow_measure(c.ptr, eps, nsteps, every, _d(out)))
    return out.reshape(nsteps // every, 7)


def jl_backup(R, c):  # substitute_U!(b::HIPBackup, U)
    _ok(R.gf_save_links(c.ptr))


def jl_restore(R, c):  # substitute_U!(U, b::HIPBackup)
    _ok(R.gf_restore_links(c.ptr))


def jl_heatbath(R, c, beta, seed, sweep, itmax):
    fails = ctypes.c_longlong(0)
    _ok(R.gf_heatbath(c.ptr, beta, seed, sweep, min(itmax, 65534), ctypes.byref(fails)))
    return fails.value


def jl_overrelaxation(R, c, seed, sweep):
    _ok(R.gf_overrelaxation(c.ptr, seed, sweep))


def jl_hot_start(R, c, seed):  # gauge_configuration(...; start=:hot) after gf_create
    _ok(R.gf_hot_start(c.ptr, seed))


def jl_load_gaugefield(R, c, path):  # load_gaugefield!(U, 1, ildg, L, 3)
    buf = np.empty(36 * c.V, np.complex128)
    prec = ctypes.c_int(0)
    _ok(R.gf_ildg_read(path.encode(), _dims(c), _d(buf), ctypes.byref(prec)))
    _ok(R.gf_upload_links(c.ptr, _d(buf)))


def jl_save_binarydata(R, c, path, precision=64):
    buf = np.empty(36 * c.V, np.complex128)
    _ok(R.gf_download_links(c.ptr, _d(buf)))
    _ok(R.gf_ildg_write(path.encode(), _dims(c), _d(buf), precision))


DIMS = (4, 4, 4, 8)
V = int(np.prod(DIMS))
ZL = np.zeros((4, V, 9), np.complex128)
ZP = np.zeros((4, V, 8))

# (selector of the Julia definition, replay, arguments, the sequence must be equal (True) or a
# subsequence of the Julia one (False: the method has branches the replay does not take))
REPLAYS = [
    ("function calculate_Plaquette(U::HIPLinks)", jl_calculate_Plaquette, (), True),
    ("function measure_plaquette(", jl_measure_plaquette, (), True),
    ("function evaluate_GaugeAction(", jl_evaluate_GaugeAction, (3.0,), True),
    ("function calc_dSdUμ!(", jl_calc_dSdU, (3.0, 2), True),
    ("function evaluate_GaugeAction_untraced!(", jl_evaluate_GaugeAction_untraced, (3.0,), True),
    ("function evaluate_GaugeAction_untraced(action", jl_evaluate_GaugeAction_untraced, (3.0,), True),
    ("function gauss_distribution!(", jl_gauss_distribution, (1.0, 5, 1), True),
    ("function Base.:*(", jl_pp, (), True),
    ("function clear_U!(P::HIPMomenta)", jl_clear_momenta, (), True),
    ("function substitute_U!(P::HIPMomenta", jl_substitute_momenta, (ZP,), True),
    ("function Base.Array(P::HIPMomenta)", jl_Array_momenta, ("momenta",), False),
    ("function substitute_U!(U::HIPLinks, Uref", jl_substitute_links, (ZL,), False),
    ("Base.Array(U::HIPLinks)", jl_Array_links, (), True),
    ("function substitute_U!(Uhost::AbstractVector", jl_Array_links, (), True),
    ("function md_potential(action::Union", jl_md_potential, (3.0,), True),
    ("function md_potential(actions::MDActionSet", jl_md_potential_set, ((3.0,),), False),
    ("function md_force!(force::HIPMomenta, action::Union", jl_md_force, (3.0,), True),
    ("function md_force!(force::HIPMomenta, actions::MDActionSet, U::HIPLinks, workspace)", jl_md_force_set, ((3.0,),), False),
    ("workspace, group::MDForceGroup)", jl_md_force_set, ((1.5,),), False),
    ("function md_hamiltonian(", jl_md_hamiltonian, (3.0,), False),
    ("function update_gaugefields!(", jl_update_gaugefields, (0.02,), True),
    ("update_momenta!(P::HIPMomenta, U::HIPLinks, step_size, driver::MDDriver)", jl_update_momenta, (3.0, 0.02), True),
    ("driver::MDDriver, group::MDForceGroup)", jl_update_momenta_set, ((0.75,), 0.02), True),
    ("function md_step!(integ::OmelyanPQP", jl_md_step_omelyan, (3.0, 0.1), True),
    ("function md_trajectory!(", jl_md_trajectory, (3.0, 0, LAM, 4, 0.4), False),
    ("function flow!(U::HIPLinks, g::Gradientflow)", jl_flow, (0.01, 3), True),
    ("function calc_smearedU(", jl_calc_smearedU, (2, 0.1, AUTO), True),
    ("function back_prop(", jl_back_prop, (2, 0.1, AUTO, ZL), True),
    ("function stout_smear(", jl_stout_smear, (2, 0.1, AUTO, ZL), True),
    ("function stout_pullback(", jl_stout_pullback, (2, 0.1, AUTO, ZL, ZL), True),
    ("function topological_charge(", jl_topological_charge, ("clover",), True),
    ("function topological_charge_density(", jl_topological_charge_density, ("improved",), True),
    ("function calculate_energy_density(", jl_calculate_energy_density, (), True),
    ("function energy_density(", jl_energy_density, (), True),
    ("function flow_measure!(", jl_flow_measure, (0.01, 4, 2), True),
    ("function substitute_U!(b::HIPBackup", jl_backup, (), True),
    ("function substitute_U!(U::HIPLinks, b::HIPBackup)", jl_restore, (), True),
    ("function heatbath!(", jl_heatbath, (6.0, 3, 0, 100000), True),
    ("function overrelaxation!(", jl_overrelaxation, (3, 0), True),
    ("function gauge_configuration(", jl_hot_start, (7,), True),
    ("function load_gaugefield!(", jl_load_gaugefield, ("x.ildg",), True),
    ("function save_binarydata(", jl_save_binarydata, ("x.ildg",), True),
]


def _is_subsequence(a, b):
    it = iter(b)
    return all(x in it for x in a)


@pytest.mark.parametrize("selector,replay,args,exact", REPLAYS, ids=[r[1].__name__ for r in REPLAYS])
def test_replay_makes_the_julia_method_calls(selector, replay, args, exact):
    R = Rec(live=False)
    replay(R, Ctx(None, DIMS), *args)
    want = jc.sequence_of(selector)
    assert R.calls, selector
    if exact:
        assert R.calls == want, (selector, R.calls, want)
    else:
        assert _is_subsequence(R.calls, want), (selector, R.calls, want)


def test_every_julia_method_that_calls_the_library_has_a_replay():
    covered = {sel for sel, *_ in REPLAYS}
    internal = ("function _", "_configure!(", "gaussian_momenta!(", "function stout_smear!(",
                "function flow!(U::HIPLinks, eps", "function forward!(", "function layer_pullback!(")
    for sig, seq in jc.method_sequences().items():
        if not seq or sig.startswith(internal) or sig.lstrip().startswith("_"):
            continue
        assert any(sel in sig for sel in covered), f"no replay for {sig}"


def test_stout_pullback_convention_against_finite_differences(oracle):
    """L(U) = Re sum conj(W) . S(U) over every link element, S the 2-level stout map: the
    cotangent of the smeared links is W; the binding's conversion D = W^dag/2 -> backward chain
    -> 2 D_bare^dag must give dL along U -> exp(h i T_a) U as Re sum conj(Ubar) . (i T_a U)."""
    dims = (4, 4, 4, 4)
    U = gen.cold_links(dims)
    oracle.exp_update(dims, U, gen.algebra_coefficients(dims, 0.4, 11), 1.0)
    rng = np.random.default_rng(12)
    W = rng.normal(size=U.shape) + 1j * rng.normal(size=U.shape)
    rho = 0.1
    levels = [U]
    for _ in range(2):
        levels.append(oracle.stout_forward(dims, levels[-1], rho))
    D = dag(W) / 2
    for lv in (1, 0):  # the backward chain through the levels, top first
        D = oracle.stout_backward(dims, levels[lv], rho, D)
    Ubar = 2 * dag(D)

    def loss(Ux):
        return float(np.real(np.sum(np.conj(W) * oracle.stout_smear(dims, Ux, 2, rho))))

    M = gen.links_to_matrices(U)
    h = 1e-5
    for mu, s, a in [(0, 3, 2), (2, 100, 7), (3, 255, 0), (1, 17, 4)]:
        vals = []
        for sg in (1, -1):
            M2 = M.copy()
            M2[mu, s] = expm(sg * h * 1j * lt.T_GEN[a]) @ M[mu, s]
            vals.append(loss(gen.matrices_to_links(M2)))
        fd = (vals[0] - vals[1]) / (2 * h)
        dU = gen.matrices_to_links((1j * lt.T_GEN[a] @ M[mu, s])[None, None])[0, 0]
        pred = float(np.real(np.sum(np.conj(Ubar[mu, s]) * dU)))
        assert abs(fd - pred) < 1e-7 * max(1.0, abs(pred)), (mu, s, a, fd, pred)


@pytest.mark.gpu
def test_replays_against_the_oracle(gf, oracle, tmp_path):
    from gaugefields_jl_amd import _lib

    L = _lib.load()
    ptr = ctypes.c_void_p()
    _ok(L.gf_create((ctypes.c_int * 4)(*DIMS), _lib.GF_F64, 0, ctypes.byref(ptr)))
    c = Ctx(ptr, DIMS)
    R = Rec(live=True)
    try:
        dims, coeff = DIMS, 3.0
        U = gen.cold_links(dims)
        oracle.exp_update(dims, U, gen.algebra_coefficients(dims, 0.3, 21), 1.0)
        jl_substitute_links(R, c, U)
        Ud = jl_Array_links(R, c)  # the third row rebuilt from the two stored rows
        assert np.abs(Ud - U).max() < 1e-15
        p = oracle.plaquette_sum(dims, U)
        assert abs(jl_calculate_Plaquette(R, c) - p) < 1e-12 * abs(p)
        assert abs(jl_measure_plaquette(R, c) - p / (18 * V)) < 1e-12
        assert abs(jl_evaluate_GaugeAction(R, c, coeff) - 2 * coeff * p) < 1e-12 * abs(2 * coeff * p)
        Mu = jl_evaluate_GaugeAction_untraced(R, c, coeff)
        Mref = lt.untraced_plaquette_action(U, dims, coeff)
        assert np.abs(Mu - Mref).max() < 1e-12 * np.abs(Mref).max()
        Dref = oracle.dsdu(dims, U, coeff)
        assert np.abs(jl_calc_dSdU(R, c, coeff, 2) - Dref[1]).max() < 1e-12 * np.abs(Dref).max()
        # momenta: device Philox stream = oracle stream
        jl_gauss_distribution(R, c, 1.0, 5, 1)
        P = jl_Array_momenta(R, c, "momenta")
        assert np.abs(P - oracle.gaussian_momenta(dims, 5, 1, 1.0)).max() < 1e-12
        assert abs(jl_pp(R, c) - oracle.momenta_dot(dims, P)) < 1e-12 * oracle.momenta_dot(dims, P)
        Vref = oracle.potential(dims, U, coeff)
        assert abs(jl_md_potential(R, c, coeff) - Vref) < 1e-12 * abs(Vref)
        Href = Vref + 0.5 * oracle.momenta_dot(dims, P)
        assert abs(jl_md_hamiltonian(R, c, coeff) - Href) < 1e-12 * abs(Href)
        jl_md_force(R, c, coeff)
        F = oracle.force(dims, U, coeff)
        assert np.abs(jl_Array_momenta(R, c, "force") - F).max() < 1e-12 * np.abs(F).max()
        # an MDActionSet of two plaquette members: potential, force and a group's kick sum theirs
        Vs = oracle.potential(dims, U, coeff) + oracle.potential(dims, U, -0.5)
        assert abs(jl_md_potential_set(R, c, (coeff, -0.5)) - Vs) < 1e-12 * abs(Vs)
        jl_md_force_set(R, c, (coeff, -0.5))
        Fs = F + oracle.force(dims, U, -0.5)
        assert np.abs(jl_Array_momenta(R, c, "force") - Fs).max() < 1e-12 * np.abs(Fs).max()
        jl_update_momenta_set(R, c, (0.75, 0.25), 0.02)
        Pk = np.ascontiguousarray(P.copy())
        oracle.kick(dims, Pk, oracle.force(dims, U, 0.75), 0.02)
        oracle.kick(dims, Pk, oracle.force(dims, U, 0.25), 0.02)
        assert np.abs(jl_Array_momenta(R, c, "momenta") - Pk).max() < 1e-12
        jl_substitute_momenta(R, c, P)
        # one Omelyan step through the per-call methods = the oracle's 2MN trajectory of one step
        jl_backup(R, c)
        jl_md_step_omelyan(R, c, coeff, 0.1)
        Uo, Po = U.copy(), P.copy()
        oracle.md_trajectory(dims, Uo, Po, coeff, 1, 0.1, "2MN")
        assert np.abs(jl_Array_links(R, c) - Uo).max() < 1e-12
        assert np.abs(jl_Array_momenta(R, c, "momenta") - Po).max() < 1e-11
        jl_restore(R, c)
        assert np.array_equal(jl_Array_links(R, c), Ud)
        # native trajectory
        jl_clear_momenta(R, c)
        assert jl_pp(R, c) == 0
        jl_substitute_momenta(R, c, P)
        H = jl_md_trajectory(R, c, coeff, 0, LAM, 4, 0.4)
        Uo, Po = U.copy(), P.copy()
        Hr = oracle.md_trajectory(dims, Uo, Po, coeff, 4, 0.4, "QPQ")
        assert abs(H[2] - Hr[2]) < 1e-12 * abs(Hr[0]), (H, Hr)
        assert np.abs(jl_Array_links(R, c) - Uo).max() < 1e-12
        # flow
        jl_substitute_links(R, c, U)
        jl_flow(R, c, 0.01, 3)
        Uf = U.copy()
        oracle.flow(dims, Uf, 0.01, 3)
        assert np.abs(jl_Array_links(R, c) - Uf).max() < 1e-12
        # stout: smear, back_prop, the AD map and its pullback
        jl_substitute_links(R, c, U)
        S = jl_calc_smearedU(R, c, 2, 0.1, AUTO)
        assert np.abs(S - oracle.stout_smear(dims, U, 2, 0.1)).max() < 1e-13
        rng = np.random.default_rng(22)
        D = rng.normal(size=U.shape) + 1j * rng.normal(size=U.shape)
        levels = [U, oracle.stout_forward(dims, U, 0.1)]
        Dref = oracle.stout_backward(dims, levels[0], 0.1, oracle.stout_backward(dims, levels[1], 0.1, D))
        assert np.abs(jl_back_prop(R, c, 2, 0.1, AUTO, D) - Dref).max() < 1e-12 * np.abs(Dref).max()
        assert np.abs(jl_stout_smear(R, c, 2, 0.1, AUTO, U) - S).max() == 0
        W = rng.normal(size=U.shape) + 1j * rng.normal(size=U.shape)
        Ubar = jl_stout_pullback(R, c, 2, 0.1, AUTO, U, W)
        Dw = oracle.stout_backward(dims, levels[0], 0.1, oracle.stout_backward(dims, levels[1], 0.1, dag(W) / 2))
        assert np.abs(Ubar - 2 * dag(Dw)).max() < 1e-12 * np.abs(Ubar).max()
        # measurements, heatbath, files
        jl_substitute_links(R, c, U)
        qref = oracle.topological_charge_density(dims, U, "clover")[0]
        assert abs(jl_topological_charge(R, c, "clover") - qref) < 1e-12
        Qi, qi = oracle.topological_charge_density(dims, U, "improved")
        assert np.abs(jl_topological_charge_density(R, c, "improved") - qi).max() < 1e-12
        Ec = oracle.energy_density_clover(dims, U)
        assert abs(jl_calculate_energy_density(R, c) - Ec) < 1e-12 * abs(Ec)
        Ep = oracle.energy_density(dims, U)
        assert abs(jl_energy_density(R, c) - Ep) < 1e-11
        rows = jl_flow_measure(R, c, 0.01, 4, 2)
        Uf2 = U.copy()
        oracle.flow(dims, Uf2, 0.01, 2)
        assert rows.shape == (2, 7) and abs(rows[0, 0] - 0.02) < 1e-15
        assert abs(rows[0, 2] - oracle.energy_density(dims, Uf2)) < 1e-11
        jl_substitute_links(R, c, U)
        Uh = U.copy()
        fails = jl_heatbath(R, c, 6.0, 3, 0, 100000)
        oracle.heatbath_sweep(dims, Uh, 6.0, 3, 0, 65534)
        assert fails == 0
        assert np.abs(jl_Array_links(R, c) - Uh).max() < 1e-10
        path = str(tmp_path / "conf.ildg")
        jl_save_binarydata(R, c, path, 64)
        jl_hot_start(R, c, 7)
        jl_load_gaugefield(R, c, path)
        assert np.abs(jl_Array_links(R, c) - Uh).max() < 1e-10
    finally:
        L.gf_destroy(ptr)
