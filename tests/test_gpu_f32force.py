"""The FP32-force form (GF_F32_FORCE, precision "f32force"): links and momenta held and updated in
FP64, every force -- staples, stout chain, TA projection -- evaluated in FP32 from the links
rounded to float.  Checked against the FP64 oracle: the state, its link updates, the Hamiltonian
and the plaquette to the FP64 bars; the forces to the FP32-bulk bars against the oracle's force of
the same float-rounded links; whole trajectories to FP32-force accuracy.
"""
import numpy as np
import pytest

import lattice_tools as lt

pytestmark = pytest.mark.gpu

gen = lt.gen
D4 = (4, 4, 4, 4)
COEFF = 3.0  # beta = 6


@pytest.fixture(scope="module")
def Lat(gf):
    made = []

    def make(dims=D4, precision="f32force"):
        L = gf.Lattice(dims, precision=precision)
        made.append(L)
        return L

    yield make
    for L in made:
        L.close()


def alg_links(oracle, dims, sigma, seed):
    U = gen.cold_links(dims)
    oracle.exp_update(dims, U, gen.algebra_coefficients(dims, sigma, seed), 1.0)
    return U


def maxdiff(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max())


def rounded(Lat, dims, U):
    """the links as an FP32 context holds them (rows 0 and 1 rounded to float)"""
    L32 = Lat(dims, "f32")
    L32.upload_links(U)
    return L32.download_links()


def test_state_is_fp64(Lat, oracle):
    print("[f32force] state", flush=True)
    L = Lat()
    U = alg_links(oracle, D4, 0.3, 71)
    P = gen.gaussian_momenta(D4, 72)
    L.upload_links(U)
    L.upload_momenta(P)
    M, Md = gen.links_to_matrices(U), gen.links_to_matrices(L.download_links())
    assert np.array_equal(Md[..., :2, :], M[..., :2, :])
    assert np.array_equal(L.download_momenta(), P)
    ref = oracle.plaquette_sum(D4, U)
    assert abs(L.plaquette_sum() - ref) < 1e-12 * abs(ref)
    H = L.hamiltonian(COEFF)
    Href = oracle.potential(D4, U, COEFF) + 0.5 * oracle.momenta_dot(D4, P)
    assert abs(H - Href) < 1e-12 * abs(Href)
    # Metropolis backup: restored exactly
    L.save_links()
    L.md_trajectory(COEFF, 2, 0.2, "QPQ")
    L.restore_links()
    Mr = gen.links_to_matrices(L.download_links())
    assert np.array_equal(Mr[..., :2, :], M[..., :2, :])


@pytest.mark.parametrize("nlev", [0, 3])
def test_force_is_fp32_of_rounded_links(Lat, oracle, nlev):
    print(f"[f32force] force nlev {nlev}", flush=True)
    L = Lat()
    U = alg_links(oracle, D4, 0.3, 73)
    L.upload_links(U)
    Ur = rounded(Lat, D4, U)
    if nlev:
        L.stout_config(nlev, 0.1, "store")
        F = oracle.stout_force(D4, Ur, nlev, 0.1, COEFF)
    else:
        F = oracle.force(D4, Ur, COEFF)
    L.force(COEFF)
    # the FP32-bulk bars for the same float-rounded links
    assert maxdiff(L.download_force(), F) < (2e-5 if nlev else 2e-6) * np.abs(F).max()


def test_kick_and_update_in_fp64(Lat, oracle):
    # P += eps F in double from the FP32 force; U <- exp(eps P) U in double, as the FP64 oracle
    print("[f32force] kick and update", flush=True)
    L = Lat()
    U = alg_links(oracle, D4, 0.3, 75)
    P = gen.gaussian_momenta(D4, 76)
    L.upload_links(U)
    L.upload_momenta(P)
    L.force(COEFF)
    F = L.download_force()
    L.force_kick(COEFF, 0.05)
    P1 = P + 0.05 * F
    assert maxdiff(L.download_momenta(), P1) < 1e-15 * np.abs(P1).max()
    L.exp_update(0.07)
    oracle.exp_update(D4, U, L.download_momenta(), 0.07)
    assert maxdiff(L.download_links(), U) < 1e-14


@pytest.mark.parametrize("integ", ["QPQ", "PQP", "2MN"])
@pytest.mark.parametrize("fused", [False, True])
def test_trajectory_against_fp64_oracle(Lat, oracle, integ, fused):
    print(f"[f32force] trajectory {integ} fused={fused}", flush=True)
    U = alg_links(oracle, D4, 0.3, 77)
    P = gen.gaussian_momenta(D4, 78)
    L = Lat()
    L.upload_links(U)
    L.upload_momenta(P)
    H = L.md_trajectory(COEFF, 6, 0.6, integ, fused=fused)
    Uo, Po = U.copy(), P.copy()
    Href = oracle.md_trajectory(D4, Uo, Po, COEFF, 6, 0.6, integ)
    assert abs(H[0] - Href[0]) < 1e-12 * abs(Href[0])
    assert abs(H[2] - Href[2]) < 1e-4
    assert maxdiff(L.download_links(), Uo) < 1e-5 and maxdiff(L.download_momenta(), Po) < 1e-4


# the plain force in its other two stencil forms (LDS-shared pairs, per-thread staples), with the
# FP64 kick and link update of the streaming state kernel
@pytest.mark.parametrize("dims", [(12, 4, 4, 4), (2, 4, 2, 6)])
def test_trajectory_other_stencil_forms(Lat, oracle, dims):
    print(f"[f32force] trajectory {dims}", flush=True)
    U = alg_links(oracle, dims, 0.3, 83)
    P = gen.gaussian_momenta(dims, 84)
    L = Lat(dims)
    L.upload_links(U)
    L.upload_momenta(P)
    H = L.md_trajectory(COEFF, 6, 0.6, "2MN", fused=True)
    Uo, Po = U.copy(), P.copy()
    Href = oracle.md_trajectory(dims, Uo, Po, COEFF, 6, 0.6, "2MN")
    assert abs(H[0] - Href[0]) < 1e-12 * abs(Href[0])
    assert abs(H[2] - Href[2]) < 1e-4
    assert maxdiff(L.download_links(), Uo) < 1e-5 and maxdiff(L.download_momenta(), Po) < 1e-4


# the three stencil forms of the fused stout epilogue: whole x-rows (4^4), LDS-shared pairs
# (12,4,4,4), per-thread staples (2,4,2,6: V not a multiple of 64)
@pytest.mark.parametrize("dims", [D4, (12, 4, 4, 4), (2, 4, 2, 6)])
def test_stout_trajectory_and_reversibility(Lat, oracle, dims):
    print(f"[f32force] stout trajectory {dims}", flush=True)
    U = alg_links(oracle, dims, 0.3, 79)
    P = gen.gaussian_momenta(dims, 80)
    L = Lat(dims)
    L.upload_links(U)
    L.upload_momenta(P)
    L.stout_config(3, 0.1, "store")
    H = L.md_trajectory(COEFF, 4, 0.2, "2MN", fused=True)
    Href = oracle.md_trajectory(dims, U.copy(), P.copy(), COEFF, 4, 0.2, "2MN", nlev=3, rho=0.1)
    assert abs(H[0] - Href[0]) < 1e-12 * abs(Href[0])
    assert abs(H[2] - Href[2]) < 1e-4
    # forward and back: the FP64 state returns to the start up to the FP32 forces' rounding
    L.md_trajectory(COEFF, 4, -0.2, "2MN", fused=True)
    assert maxdiff(L.download_links(), U) < 1e-6
    assert maxdiff(L.download_momenta(), P) < 1e-5


@pytest.mark.parametrize("nlev", [0, 3])
def test_fused_update_equals_kick_then_update(Lat, oracle, nlev):
    # the fused kernel (force, kick and link update in one launch for the stout path) does the
    # same double-precision kick and update as the separate calls: identical results
    print(f"[f32force] fused md_update nlev {nlev}", flush=True)
    U = alg_links(oracle, D4, 0.3, 81)
    P = gen.gaussian_momenta(D4, 82)
    out = []
    for fused in (True, False):
        L = Lat()
        L.upload_links(U)
        L.upload_momenta(P)
        if nlev:
            L.stout_config(nlev, 0.1, "store")
        for _ in range(2):
            if fused:
                L.md_update(COEFF, 0.05, 0.04)
            else:
                L.force_kick(COEFF, 0.05)
                L.exp_update(0.04)
        out.append((L.download_links(), L.download_momenta(), L.plaquette_sum()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_cold_start(Lat):
    L = Lat()
    H = L.md_trajectory(COEFF, 3, 1.0, "QPQ")
    assert H[2] == 0.0 and L.plaquette() == 1.0


def test_unavailable_operations(Lat, gf):
    L = Lat()
    from gaugefields_jl_amd import _lib

    for call in (lambda: L.flow(0.01, 1), lambda: L.heatbath(6.0, 1, 0), lambda: L.topological_charge("clover"),
                 lambda: L.stout_backward(np.zeros((4, 256, 9), complex))):
        with pytest.raises(_lib.GFError) as e:
            call()
        assert e.value.code == _lib.GF_ESTATE


def test_gauge_configuration_force_eltype(gf):
    from gaugefields_jl_amd import api

    U = api.gauge_configuration((4, 4, 4, 4), start="hot", seed=5, force_eltype=np.complex64)
    assert U.lattice.precision == "f32force" and U.eltype == np.complex128
    with pytest.raises(api.ArgumentError):
        api.gauge_configuration((4, 4, 4, 4), eltype=np.complex64, force_eltype=np.complex128)


def test_api_smeared_potential_and_trajectory(gf):
    # the reference-style API on an FP32-force configuration: the smeared potential is the FP64
    # one of the same links, and a stout trajectory runs through md_trajectory
    from gaugefields_jl_amd import api

    print("[f32force] api smeared potential", flush=True)
    U = api.gauge_configuration(D4, start="hot", seed=91, force_eltype=np.complex64)
    U64 = api.gauge_configuration(D4)
    U64.lattice.upload_links(U.to_host())

    def smeared(W):
        a = api.GaugeAction(W)
        plaq = api.make_loops_fromname("plaquette")
        a.push(2.5, plaq + api.dagger(plaq))
        return api.SmearedGaugeAction(a, api.stout_smearing(W, rho=0.1, levels=2, mode="store"))

    s32, s64 = smeared(U), smeared(U64)
    v32, v64 = api.md_potential(s32, U), api.md_potential(s64, U64)
    assert abs(v32 - v64) < 1e-12 * abs(v64)
    P = api.gaussian_momenta(U, seed=92)
    r = api.md_trajectory(U, P, api.md_driver(U, s32, steps=3, trajectory_length=0.3,
                                              integrator=api.OmelyanPQP()))
    assert np.isfinite(r["delta_hamiltonian"])
