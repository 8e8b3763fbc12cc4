"""Pins of the CPU oracle (oracle/gf_oracle.c) -- CPU only.

Each oracle function is checked against something it does not itself restate: the
reference's golden values, closed forms, scipy routines, a numpy whole-lattice brute
force, finite differences of the potential, and gauge symmetry.  The layout of each test
names the reference file:line whose behaviour it pins.
"""
import numpy as np
import pytest
from scipy.linalg import expm, expm_frechet

import lattice_tools as lt

gen = lt.gen
D4 = (4, 4, 4, 4)
BETA = 6.0
COEFF = BETA / 2  # push!(action, beta/2, plaquette ∪ plaquette^dag)


def hot(dims, seed):
    return gen.hot_start_links(dims, seed)


def alg_links(oracle, dims, sigma, seed):
    U = gen.cold_links(dims)
    oracle.exp_update(dims, U, gen.algebra_coefficients(dims, sigma, seed), 1.0)
    return U


# ---------------------------------------------------------------- plaquette


def test_cold_plaquette_is_one(oracle):
    # test/init.jl cold testsets; test/high_level_api.jl:48-49
    for dims in [D4, (2, 2, 2, 2), (4, 2, 3, 2)]:
        U = gen.cold_links(dims)
        V = np.prod(dims)
        assert oracle.plaquette_sum(dims, U) / (6 * V * 3) == 1.0


def test_instanton_golden(oracle):
    # test/init.jl:351-371 calls Oneinstanton(NC, NX, NY, NZ, NT, Nwing) against the
    # signature Oneinstanton(NC, NDW, NN...) (AbstractGaugefields.jl:998), so its lattice
    # is 4x4x4x1 (halo 4): SU(2) plaquette 0.9796864531099871.  SU(3) embedding gives
    # (2p + 1)/3 (test/sun_embedded_instanton.jl:145).
    p2_ref = 0.9796864531099871
    dims = (4, 4, 4, 1)
    V = 64
    M2 = lt.instanton_links(dims, nc=2)
    p2 = lt.plaquette_field_sum(M2.reshape(4, 1, 4, 4, 4, 2, 2), dims) / (6 * V * 2)
    assert abs(p2 - p2_ref) / p2_ref < 1e-14
    for block in [(0, 1), (0, 2), (1, 2)]:
        U = gen.matrices_to_links(lt.instanton_links(dims, nc=3, block=block))
        p3 = oracle.plaquette_sum(dims, U) / (6 * V * 3)
        assert abs(p3 - (2 * p2_ref + 1) / 3) < 1e-14


@pytest.mark.parametrize("dims", [D4, (2, 2, 2, 2), (4, 2, 3, 2), (2, 4, 2, 6)])
def test_plaquette_vs_numpy_bruteforce(oracle, dims):
    U = hot(dims, 7)
    ref = lt.plaquette_field_sum(U, dims)
    assert abs(oracle.plaquette_sum(dims, U) - ref) < 1e-12 * abs(ref) + 1e-12


def test_plaquette_gauge_invariant(oracle):
    rng = np.random.default_rng(3)
    U = alg_links(oracle, D4, 0.3, 5)
    g = lt.random_su3(256, rng)
    Ug = lt.gauge_transform(U, D4, g)
    a, b = oracle.plaquette_sum(D4, U), oracle.plaquette_sum(D4, Ug)
    assert abs(a - b) < 1e-13 * abs(a)


def test_potential_normalization(oracle):
    # md_potential = -real(evaluate_GaugeAction)/NC, evaluate = 2*coeff*sum Re tr P
    # (molecular_dynamics.jl:247-249, test/enzyme_md.jl:17-40)
    U = hot(D4, 9)
    M = lt.field(U, D4)
    val = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            P = M[mu] @ lt.sh(M[nu], mu) @ lt.dg(lt.sh(M[mu], nu)) @ lt.dg(M[nu])
            Pr = M[nu] @ lt.sh(M[mu], nu) @ lt.dg(lt.sh(M[nu], mu)) @ lt.dg(M[mu])
            val += np.trace(P, axis1=-2, axis2=-1).real.sum() + np.trace(Pr, axis1=-2, axis2=-1).real.sum()
    assert abs(oracle.potential(D4, U, 2.75) - (-2.75 * val / 3)) < 1e-12 * abs(val)


# ---------------------------------------------------------------- single-matrix primitives


def test_ta_projection_closed_form(oracle):
    # Traceless_antihermitian_add! (TA_gaugefields_4D_serial.jl:181-269): coefficients of
    # TA(M) = (M - M^dag)/2 - tr(M - M^dag)/6 in the basis i lambda_a / 2
    rng = np.random.default_rng(1)
    for _ in range(20):
        M = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        c = oracle.ta_coeffs(M)
        assert np.abs(lt.coeffs_to_matrix(c) - lt.ta_matrix(M)).max() < 1e-14
        assert np.abs(oracle.ta_matrix(M) - lt.ta_matrix(M)).max() < 1e-15


@pytest.mark.parametrize("scale", [3.0, 1.0, 0.3, 0.05, 1e-3, 1e-7, 0.0])
def test_exp_vs_scipy(oracle, scale):
    rng = np.random.default_rng(int(scale * 1e6) + 1)
    for trial in range(40):
        c = scale * rng.standard_normal(8)
        if trial % 2:
            c = -c
        t = 0.7
        E = oracle.exp_ta(c, t)
        ref = expm(t * lt.coeffs_to_matrix(c))
        assert np.abs(E - ref).max() < 2e-15 * max(1.0, scale * 4)
        assert np.abs(E.conj().T @ E - np.eye(3)).max() < 5e-15
        assert abs(np.linalg.det(E) - 1) < 5e-15


def test_exp_degenerate_spectra(oracle):
    # two equal eigenvalues of either sign of det Q (9u^2 = w^2 without the c0 symmetry)
    for c8 in [0.3, -0.3, 1.7, -1.7]:
        c = np.zeros(8)
        c[7] = c8
        assert np.abs(oracle.exp_ta(c, 1.0) - expm(lt.coeffs_to_matrix(c))).max() < 2e-15
    c = np.zeros(8)
    c[2] = 0.4
    assert np.abs(oracle.exp_ta(c, 1.0) - expm(lt.coeffs_to_matrix(c))).max() < 2e-15


@pytest.mark.parametrize("scale", [2.0, 0.6, 0.3, 0.15, 0.05, 1e-3, 1e-6])
def test_cdexpqdq_vs_expm_frechet(oracle, scale):
    # CdexpQdQ! (stout_fast.jl:888-946): d tr(C e^Q) = tr(L dQ) along su(3)
    rng = np.random.default_rng(int(scale * 1e4) + 2)
    for trial in range(30):
        c = scale * rng.standard_normal(8)
        if trial % 2:
            c = -c
        Q = lt.coeffs_to_matrix(c)
        C = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        L = oracle.cdexpqdq(C, Q)
        for a in range(8):
            E = 1j * lt.T_GEN[a]
            ref = np.trace(C @ expm_frechet(Q, E, compute_expm=False))
            assert abs(np.trace(L @ E) - ref) < 3e-14 * max(1.0, np.abs(C).max())


def test_exp_antiherm(oracle):
    rng = np.random.default_rng(4)
    Q = lt.coeffs_to_matrix(rng.standard_normal(8))
    assert np.abs(oracle.exp_antiherm(Q) - expm(Q)).max() < 3e-15


# ---------------------------------------------------------------- derivative and force


def test_dsdu_closes_plaquettes(oracle):
    # calc_dSdUμ! (GaugeActions.jl:95-123): sum_{x,mu} Re tr(U_mu dSdU_mu) = 4 coeff sum Re tr P
    U = hot(D4, 21)
    D = oracle.dsdu(D4, U, COEFF)
    MU, MD = gen.links_to_matrices(U), gen.links_to_matrices(D)
    tot = np.trace(MU @ MD, axis1=-2, axis2=-1).real.sum()
    assert abs(tot - 4 * COEFF * oracle.plaquette_sum(D4, U)) < 1e-11 * abs(tot)


def test_dsdu_vs_numpy(oracle):
    dims = (2, 4, 2, 4)
    U = hot(dims, 22)
    M = lt.field(U, dims)
    D = lt.field(oracle.dsdu(dims, U, 1.3), dims)
    for mu in range(4):
        ref = 0
        for nu in range(4):
            if nu == mu:
                continue
            ref = ref + lt.sh(M[nu], mu) @ lt.dg(lt.sh(M[mu], nu)) @ lt.dg(M[nu])
            Mnm = lt.sh(M[nu], nu, -1)
            ref = ref + lt.dg(lt.sh(Mnm, mu)) @ lt.dg(lt.sh(M[mu], nu, -1)) @ Mnm
        assert np.abs(D[mu] - 1.3 * ref).max() < 1e-13


def _fd_force(pot, U, mu, s, a, h=1e-5):
    return -(pot(lt.perturb_link(U, mu, s, a, h)) - pot(lt.perturb_link(U, mu, s, a, -h))) / (2 * h)


def test_force_vs_finite_difference(oracle):
    # md_force! (molecular_dynamics.jl:251-267) is dp/dtau = -dV/domega_a along
    # U -> exp(omega i T_a) U; central differences h=1e-5 on 64 random (link, generator)
    U = alg_links(oracle, D4, 0.3, 31)
    F = oracle.force(D4, U, COEFF)
    rng = np.random.default_rng(5)
    pot = lambda W: oracle.potential(D4, W, COEFF)  # noqa: E731
    for _ in range(64):
        mu, s, a = rng.integers(4), rng.integers(256), rng.integers(8)
        fd = _fd_force(pot, U, mu, s, a)
        assert abs(F[mu, s, a] - fd) < 1e-7 * max(1.0, abs(fd))


def test_force_gauge_covariant(oracle):
    rng = np.random.default_rng(6)
    U = hot(D4, 32)
    g = lt.random_su3(256, rng)
    F = lt.coeffs_to_matrix(oracle.force(D4, U, COEFF))
    Fg = lt.coeffs_to_matrix(oracle.force(D4, lt.gauge_transform(U, D4, g), COEFF))
    assert np.abs(Fg - g[None] @ F @ lt.dg(g)[None]).max() < 1e-13


def test_cold_force_zero(oracle):
    F = oracle.force(D4, gen.cold_links(D4), COEFF)
    assert np.abs(F).max() == 0.0


def test_exp_update_and_kick(oracle):
    U = hot(D4, 41)
    P = gen.gaussian_momenta(D4, 42)
    U2 = U.copy()
    oracle.exp_update(D4, U2, P, 0.02)
    MU, M2 = gen.links_to_matrices(U), gen.links_to_matrices(U2)
    for mu, s in [(0, 0), (3, 255), (2, 77)]:
        ref = expm(0.02 * lt.coeffs_to_matrix(P[mu, s])) @ MU[mu, s]
        assert np.abs(M2[mu, s] - ref).max() < 1e-15
    F = gen.gaussian_momenta(D4, 43)
    P2 = P.copy()
    oracle.kick(D4, P2, F, 0.3)
    assert np.abs(P2 - (P + 0.3 * F)).max() == 0.0
    assert abs(oracle.momenta_dot(D4, P) - np.sum(P * P)) < 1e-10


@pytest.mark.parametrize("dims", [D4, (4, 2, 3, 6)])
def test_unfused_md_update_is_force_kick_exp(oracle, dims):
    """the reference-structured CPU baseline (104 whole-lattice passes through temporaries) is
    the same map as the finite-difference-pinned force, the kick and the expm-pinned update"""
    U = hot(dims, 44)
    P = gen.gaussian_momenta(dims, 45)
    U1, P1 = U.copy(), P.copy()
    oracle.md_update_unfused(dims, U1, P1, COEFF, 0.02, 0.03)
    U2, P2 = U.copy(), P.copy()
    oracle.kick(dims, P2, oracle.force(dims, U2, COEFF), 0.02)
    oracle.exp_update(dims, U2, P2, 0.03)
    assert np.abs(P1 - P2).max() < 1e-14
    assert np.abs(U1 - U2).max() < 1e-14


# ---------------------------------------------------------------- Wilson flow


def test_flow_force_is_scaled_md_force(oracle):
    # add_force!(plaqonly) (AbstractGaugefields.jl:2717-2762): TA(U V^dag) with V the
    # loops_staple_prime sum equals -(NC/coeff) times the FD-pinned MD force
    U = hot(D4, 51)
    assert np.abs(oracle.flow_force(D4, U) + 3.0 / COEFF * oracle.force(D4, U, COEFF)).max() < 1e-13


def test_flow_cold_stays_cold_and_energy_decreases(oracle):
    U = gen.cold_links(D4)
    E = oracle.flow(D4, U, 0.01, 5, 1)
    assert np.all(E == 0.0)
    U = alg_links(oracle, D4, 0.4, 52)
    E0 = oracle.energy_density(D4, U)
    E = oracle.flow(D4, U, 0.01, 40, 5)
    assert E0 > E[0] and np.all(np.diff(E) < 0)


def test_flow_rk3_third_order(oracle):
    # gradientflow.jl:171-238 is Luscher's 3rd-order RK: halving eps cuts the global
    # error ~8x; compare against a 4x finer run
    U0 = alg_links(oracle, D4, 0.4, 53)
    runs = {}
    for n in (10, 20, 80):
        U = U0.copy()
        oracle.flow(D4, U, 0.1 / n, n, 0)
        runs[n] = U
    e10 = np.abs(runs[10] - runs[80]).max()
    e20 = np.abs(runs[20] - runs[80]).max()
    assert 5.0 < e10 / e20 < 10.5


def test_flow_gauge_covariant(oracle):
    rng = np.random.default_rng(7)
    U = alg_links(oracle, D4, 0.4, 54)
    g = lt.random_su3(256, rng)
    Ug = lt.gauge_transform(U, D4, g)
    oracle.flow(D4, U, 0.02, 3, 0)
    oracle.flow(D4, Ug, 0.02, 3, 0)
    assert np.abs(Ug - lt.gauge_transform(U, D4, g)).max() < 1e-13


# ---------------------------------------------------------------- stout


def test_stout_rho_zero_identity(oracle):
    U = hot(D4, 61)
    assert np.abs(oracle.stout_forward(D4, U, 0.0) - U).max() == 0.0


def test_stout_forward_vs_numpy(oracle):
    dims = (2, 4, 2, 4)
    U = hot(dims, 62)
    M = lt.field(U, dims)
    out = lt.field(oracle.stout_forward(dims, U, 0.1), dims)
    for mu in range(4):
        C = 0
        for nu in range(4):
            if nu == mu:
                continue
            C = C + M[nu] @ lt.sh(M[mu], nu) @ lt.dg(lt.sh(M[nu], mu))
            Mnm = lt.sh(M[nu], nu, -1)
            C = C + lt.dg(Mnm) @ lt.sh(M[mu], nu, -1) @ lt.sh(Mnm, mu)
        Q = lt.ta_matrix(0.1 * C @ lt.dg(M[mu]))
        flat = Q.reshape(-1, 3, 3)
        E = np.array([expm(q) for q in flat]).reshape(Q.shape)
        assert np.abs(out[mu] - E @ M[mu]).max() < 1e-14


def test_stout_forward_gauge_covariant(oracle):
    rng = np.random.default_rng(8)
    U = hot(D4, 63)
    g = lt.random_su3(256, rng)
    a = oracle.stout_smear(D4, lt.gauge_transform(U, D4, g), 3, 0.12)
    b = lt.gauge_transform(oracle.stout_smear(D4, U, 3, 0.12), D4, g)
    assert np.abs(a - b).max() < 1e-13


@pytest.mark.parametrize("nlev,rho,sigma", [(1, 0.1, 0.3), (3, 0.1, 0.3), (2, 0.12, 1.0)])
def test_stout_force_vs_finite_difference(oracle, nlev, rho, sigma):
    # stout HMC force (test/HMCstout_test_nowing.jl:99-118) = -dV_smeared/domega
    U = alg_links(oracle, D4, sigma, 70 + nlev)
    F = oracle.stout_force(D4, U, nlev, rho, COEFF)
    rng = np.random.default_rng(9 + nlev)
    pot = lambda W: oracle.stout_potential(D4, W, nlev, rho, COEFF)  # noqa: E731
    for _ in range(64 if nlev == 3 else 24):
        mu, s, a = rng.integers(4), rng.integers(256), rng.integers(8)
        fd = _fd_force(pot, U, mu, s, a)
        assert abs(F[mu, s, a] - fd) < 2e-7 * max(1.0, abs(fd))


def test_stout_zero_levels_is_plain_force(oracle):
    U = hot(D4, 81)
    assert np.abs(oracle.stout_force(D4, U, 0, 0.1, COEFF) - oracle.force(D4, U, COEFF)).max() < 1e-15


# ---------------------------------------------------------------- MD


def test_md_cold_start(oracle):
    # test/md_driver.jl:258-282: cold links, zero momenta -> dH == 0, plaquette stays 1
    U = gen.cold_links(D4)
    P = np.zeros((4, 256, 8))
    H = oracle.md_trajectory(D4, U, P, COEFF, 5, 1.0, "QPQ")
    assert H[2] == 0.0
    assert oracle.plaquette_sum(D4, U) / (18 * 256) == 1.0


@pytest.mark.parametrize("integ", ["QPQ", "PQP", "2MN"])
def test_md_reversibility(oracle, integ):
    # test/md_driver.jl:328-368: forward then trajectory_length -> -tau returns, < 2e-12
    U0 = hot(D4, 91)
    P0 = gen.gaussian_momenta(D4, 92)
    U, P = U0.copy(), P0.copy()
    oracle.md_trajectory(D4, U, P, COEFF, 5, 0.2, integ)
    oracle.md_trajectory(D4, U, P, COEFF, 5, -0.2, integ)
    assert np.abs(U - U0).max() < 2e-12 and np.abs(P - P0).max() < 2e-12


def test_md_dH_scales_as_eps_squared(oracle):
    U0 = alg_links(oracle, D4, 0.3, 93)
    P0 = gen.gaussian_momenta(D4, 94)
    dH = []
    for steps in (10, 20):
        U, P = U0.copy(), P0.copy()
        dH.append(oracle.md_trajectory(D4, U, P, COEFF, steps, 0.5, "QPQ")[2])
    assert 3.0 < dH[0] / dH[1] < 5.0


@pytest.mark.parametrize("nlev", [0, 2])
def test_md_2mn_merged_kicks_is_the_same_map(oracle, nlev):
    # adjacent kicks of consecutive 2MN steps act on the same links: merging them changes the
    # result by rounding only, while a dropped or doubled kick would move P by O(eps)
    U0 = alg_links(oracle, D4, 0.3, 97)
    P0 = gen.gaussian_momenta(D4, 98)
    out = {}
    for integ in ("2MN", "2MN-merged"):
        U, P = U0.copy(), P0.copy()
        H = oracle.md_trajectory(D4, U, P, COEFF, 4, 0.5, integ, nlev=nlev, rho=0.1)
        out[integ] = (U, P, H)
    (Ua, Pa, Ha), (Ub, Pb, Hb) = out["2MN"], out["2MN-merged"]
    assert np.abs(Ua - Ub).max() < 1e-13 and np.abs(Pa - Pb).max() < 1e-12
    assert abs(Ha[2] - Hb[2]) < 1e-12 * abs(Ha[0])
    assert np.abs(Pa - P0).max() > 1e-2


def test_config1_chain_accepts(oracle):
    # the plain-HMC workload after its thermalization: the Metropolis trajectories that the GPU
    # test compares are mostly accepted, so the accept-and-evolve path is what gets compared
    recs = lt.config1_chain(oracle)
    dH = [r["H"][2] for r in recs]
    assert dH[0] > 5.0  # hot start: far from equilibrium
    n_acc = sum(r["accepted"] for r in recs if r["metropolis"])
    assert n_acc >= 10, (n_acc, dH)
    assert all(abs(d) < 2.0 for d in dH[10:]), dH


def test_md_stout_reversibility(oracle):
    U0 = alg_links(oracle, D4, 0.3, 95)
    P0 = gen.gaussian_momenta(D4, 96)
    U, P = U0.copy(), P0.copy()
    H = oracle.md_trajectory(D4, U, P, COEFF, 3, 0.1, "2MN", nlev=2, rho=0.1)
    oracle.md_trajectory(D4, U, P, COEFF, 3, -0.1, "2MN", nlev=2, rho=0.1)
    assert np.abs(U - U0).max() < 2e-12
    assert abs(H[2]) < 0.5


# ---------------------------------------------------------------- field strengths
TOPO = ["plaquette", "clover", "rectangle", "improved"]


def test_topological_charge_cold_is_zero(oracle):
    # test/sun_embedded_instanton.jl:316-336: all densities vanish on a cold start
    U = lt.gen.cold_links(D4)
    for m in TOPO:
        Q, q = oracle.topological_charge_density(D4, U, m)
        assert Q == 0.0 and np.all(q == 0.0)
    assert oracle.energy_density_clover(D4, U) == 0.0


@pytest.mark.parametrize("dims,n", [((4, 4, 4, 4), (1, 1)), ((4, 6, 6, 4), (2, -1)), ((8, 4, 4, 8), (1, 3))])
def test_field_strength_constant_abelian_closed_form(oracle, dims, n):
    # constant flux t through every (0,1) plaquette and t' through every (2,3) plaquette, as
    # diag(e^{it}, e^{-it}, 1): the clover and plaquette F_01 = 4 or 1 times diag(i sin t,
    # -i sin t, 0), so q = sin t sin t' / (2 pi^2) per site; the 2x1 rectangles carry 2t;
    # E_clover = 2 (sin^2 t + sin^2 t'), E_plaq = 4 (1 - cos t) + 4 (1 - cos t')
    U, t, tp = lt.abelian_flux_links(dims, *n)
    V = int(np.prod(dims))
    qc = np.sin(t) * np.sin(tp) / (2 * np.pi ** 2)
    qr = np.sin(2 * t) * np.sin(2 * tp) / np.pi ** 2
    expect = {"plaquette": qc, "clover": qc, "rectangle": qr, "improved": 5 / 3 * qc - 1 / 12 * qr}
    for m in TOPO:
        Q, q = oracle.topological_charge_density(dims, U, m)
        assert np.allclose(q, expect[m], rtol=1e-12, atol=1e-15)
        assert Q == pytest.approx(V * expect[m], rel=1e-12)
    assert oracle.energy_density_clover(dims, U) == pytest.approx(2 * (np.sin(t) ** 2 + np.sin(tp) ** 2), rel=1e-12)
    assert oracle.energy_density(dims, U) == pytest.approx(4 * (2 - np.cos(t) - np.cos(tp)), rel=1e-12)


def test_abelian_charge_tends_to_flux_product(oracle):
    # small flux quanta on a larger torus: the improved charge approaches its continuum value
    # 2 n01 n23 (two U(1) charges of the diag(1, -1, 0) embedding)
    dims = (12, 12, 12, 12)
    U, _, _ = lt.abelian_flux_links(dims, 1, 1)
    for m in ("clover", "improved"):
        Q, _ = oracle.topological_charge_density(dims, U, m)
        assert Q == pytest.approx(2.0, abs=2e-3)


@pytest.mark.parametrize("dims", [(2, 2, 2, 2), (3, 2, 2, 3), (4, 3, 2, 3)])
def test_topological_density_vs_numpy(oracle, dims):
    U = hot(dims, 21)
    for m in TOPO:
        Q, q = oracle.topological_charge_density(dims, U, m)
        qn = lt.topo_density_numpy(U, dims, m)
        assert np.abs(q - qn).max() < 1e-13
        assert Q == pytest.approx(qn.sum(), rel=1e-11, abs=1e-13)


def test_topological_density_gauge_invariant(oracle):
    rng = np.random.default_rng(8)
    U = alg_links(oracle, D4, 0.4, 9)
    Ug = lt.gauge_transform(U, D4, lt.random_su3(256, rng))
    for m in TOPO:
        q1 = oracle.topological_charge_density(D4, U, m)[1]
        q2 = oracle.topological_charge_density(D4, Ug, m)[1]
        assert np.abs(q1 - q2).max() < 1e-13
    assert oracle.energy_density_clover(D4, Ug) == pytest.approx(oracle.energy_density_clover(D4, U), rel=1e-13)


def test_clover_charge_odd_under_reflection(oracle):
    # x0 -> -x0: U'_0(x) = U_0(Rx - 0)^dag, U'_k(x) = U_k(Rx); the clover and improved
    # charges change sign, the clover energy does not
    dims = (4, 4, 2, 3)
    U = alg_links(oracle, dims, 0.5, 12)
    M = lt.field(U, dims)  # (4, NT, NZ, NY, NX, 3, 3)
    R = np.flip(M, axis=4)
    R = np.roll(R, 1, axis=4)  # index x holds M[(-x) mod NX]
    out = R.copy()
    out[0] = lt.dg(np.roll(R[0], -1, axis=3))  # U_0((-x-1) mod NX)^dag
    Ur = lt.unfield(out)
    for m in ("clover", "improved"):
        assert oracle.topological_charge_density(dims, Ur, m)[0] == pytest.approx(
            -oracle.topological_charge_density(dims, U, m)[0], rel=1e-12, abs=1e-14)
    assert oracle.energy_density_clover(dims, Ur) == pytest.approx(oracle.energy_density_clover(dims, U), rel=1e-12)


def test_instanton_charge_independent_of_embedding(oracle):
    # test/sun_embedded_instanton.jl:376-391: the SU(2) instanton embedded at blocks (1,2),
    # (1,3), (2,3) has one charge; |Q| > 0.1 (:341)
    vals = []
    for blk in [(0, 1), (0, 2), (1, 2)]:
        U = lt.gen.matrices_to_links(lt.instanton_links(D4, 3, blk))
        vals.append([oracle.topological_charge_density(D4, U, m)[0] for m in ("plaquette", "clover", "improved")])
    assert np.allclose(vals, vals[0], rtol=1e-12, atol=1e-14)
    assert abs(vals[0][0]) > 0.1


# ---------------------------------------------------------------- random fields
def test_philox_known_answers(oracle):
    # Philox4x32-10 known-answer vectors of the Random123 distribution (kat_vectors)
    assert oracle.philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert oracle.philox4x32_10((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert oracle.philox4x32_10((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == (
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_oracle_random_fields(oracle):
    dims = (6, 6, 6, 6)
    P = oracle.gaussian_momenta(dims, 77, 0)
    n = P.size
    assert abs(P.mean()) < 5 / np.sqrt(n) and abs(P.var() - 1) < 5 * np.sqrt(2 / n)
    assert np.array_equal(P, oracle.gaussian_momenta(dims, 77, 0))
    assert not np.array_equal(P, oracle.gaussian_momenta(dims, 77, 1))
    assert np.allclose(oracle.gaussian_momenta(dims, 77, 0, sigma=2.0), 2.0 * P, rtol=0, atol=1e-15)
    U = oracle.hot_links(dims, 5)
    M = lt.gen.links_to_matrices(U)
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14


# ---------------------------------------------------------------- heatbath, overrelaxation
@pytest.mark.parametrize("k", [0.5, 2.0, 8.0])
def test_kennedy_pendleton_distribution(oracle, k):
    # accepted a0 follow p(a0) ~ sqrt(1 - a0^2) exp(k a0) on [-1, 1] (Kennedy & Pendleton,
    # Phys. Lett. B 156 (1985) 393; _su2_update_kp_core!, heatbath/portable/kernels.jl:63-150)
    from scipy.integrate import quad

    a = oracle.kp_samples(k, 400000, 7)
    assert not np.isnan(a).any() and a.min() >= -1 and a.max() <= 1
    w = lambda x: np.sqrt(1 - x * x) * np.exp(k * x)  # noqa: E731
    Z = quad(w, -1, 1)[0]
    for p in (1, 2, 3):
        mp = quad(lambda x: x ** p * w(x), -1, 1)[0] / Z
        sig = np.sqrt(((a ** p) ** 2).mean() - (a ** p).mean() ** 2) / np.sqrt(a.size)
        assert abs((a ** p).mean() - mp) < 5 * sig


def test_heatbath_strong_coupling(oracle):
    # SU(3) strong coupling: <P> = beta/18 + beta^2/216 + O(beta^3)
    dims = (6, 6, 6, 6)
    U = lt.gen.cold_links(dims)
    beta, ps = 0.3, []
    for sw in range(120):
        assert oracle.heatbath_sweep(dims, U, beta, 11, sw) == 0
        if sw >= 20:
            ps.append(oracle.plaquette_sum(dims, U) / (18 * np.prod(dims)))
    err = np.std(ps) / np.sqrt(len(ps))
    assert abs(np.mean(ps) - (beta / 18 + beta ** 2 / 216)) < 4 * err + 2e-4  # + beta^3 term


def test_heatbath_agrees_with_hmc(oracle):
    # both sample exp(beta/3 sum Re tr P): average plaquettes at beta = 5.5 on 4^4
    dims = (4, 4, 4, 4)
    beta = 5.5
    U = lt.gen.cold_links(dims)
    hb = []
    for sw in range(160):
        oracle.heatbath_sweep(dims, U, beta, 3, sw)
        if sw >= 40:
            hb.append(oracle.plaquette_sum(dims, U) / (18 * 256))
    rng = np.random.default_rng(5)
    W = U.copy()
    hmc = []
    for tr in range(160):
        P = rng.standard_normal((4, 256, 8))
        W1, P1 = W.copy(), P.copy()
        H = oracle.md_trajectory(dims, W1, P1, beta / 2, 10, 1.0)
        if rng.random() < np.exp(-H[2]):
            W = W1
        if tr >= 40:
            hmc.append(oracle.plaquette_sum(dims, W) / (18 * 256))
    # integrated autocorrelation ~ a few sweeps / trajectories: widen the error by 3
    e = 3 * np.hypot(np.std(hb) / np.sqrt(len(hb)), np.std(hmc) / np.sqrt(len(hmc)))
    assert abs(np.mean(hb) - np.mean(hmc)) < 4 * e


def test_overrelaxation_keeps_action(oracle):
    # overrelaxation is microcanonical: every link update keeps Re tr(U V)
    dims = (4, 4, 4, 4)
    U = alg_links(oracle, dims, 0.5, 41)
    p0 = oracle.plaquette_sum(dims, U)
    for sw in range(3):
        oracle.heatbath_sweep(dims, U, 6.0, 5, sw, overrelax=True)
    assert oracle.plaquette_sum(dims, U) == pytest.approx(p0, rel=1e-13)
    M = lt.gen.links_to_matrices(U)
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-14
    assert np.abs(U - alg_links(oracle, dims, 0.5, 41)).max() > 0.1  # the links did move
