"""The workload configurations of the hot path, run end to end on the GPU against the oracle.

1. plain quenched HMC, 4^4, beta 6, Gram-Schmidt hot start, QPQ 10 steps tau 1: 10
   thermalization trajectories accepted unconditionally (from the hot start every trajectory
   has dH of 5-7 and Metropolis would never move the chain), then 20 trajectories with
   Metropolis: plaquette, dH and accept decision per trajectory, most of them accepted;
2. kernel level, 16^4 and 32^4, omega ~ N(0, 0.3^2): plaquette, staples of all four
   directions, force, one exp update with step 0.02;
3. Wilson flow (RK3 eps 0.01, 300 steps) with E(t) every 10 steps — the full 300 steps on an
   8^4 lattice against the oracle (the 32^4 FP64/FP32 comparison is in bench.py);
4. 3-level stout (rho 0.1) at 24^4: the GPU stout force against the oracle's stout force at
   24^4 and against central finite differences (h = 1e-5) of the smeared potential on 64
   (link, generator) pairs; stored-C and recomputed-C modes; one Omelyan 2MN trajectory
   (10 steps, tau 1, FP64) against the oracle;
5. production size, 48^4 FP32 bulk, 3 stout levels at rho 0.12, 2MN 16 steps, tau 1, from
   omega ~ N(0, 0.3^2): the first trajectory and, after 20 thermalization trajectories, two
   more with their FP64 reruns (dH deviation, plaquette), reversibility, and the stored-C
   memory mode.
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
    L = gf.Lattice(dims)

    def run_gpu(U0):
        L.upload_links(U0)

        def step(P, r, metropolis, rec):
            L.save_links()
            L.upload_momenta(P)
            H = L.md_trajectory(COEFF, 10, 1.0, "QPQ")
            acc = (not metropolis) or np.exp(-H[2]) >= r
            if not acc:
                L.restore_links()
            Href = rec["H"]
            assert acc == rec["accepted"], (rec["traj"], H, Href)
            assert abs(H[2] - Href[2]) < 1e-12 * abs(Href[0]), (rec["traj"], H, Href)
            assert abs(L.plaquette_sum() - rec["plaquette_sum"]) < 1e-12 * 6 * 256 * 3
        return step

    recs = lt.config1_chain(oracle, run_gpu)
    L.close()
    # the accept-and-evolve path is exercised: most Metropolis trajectories are accepted and the
    # GPU continues from its own evolved links each time
    n_acc = sum(r["accepted"] for r in recs if r["metropolis"])
    assert n_acc >= 10, n_acc


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


@pytest.fixture(scope="module")
def config4(gf, oracle):
    dims = (24, 24, 24, 24)
    U = alg_links(oracle, dims, 0.3, 401)
    L = gf.Lattice(dims)
    L.upload_links(U)
    yield dims, U, L
    L.close()


def test_config4_stout_force_24(config4, oracle):
    dims, U, L = config4
    L.stout_config(3, 0.1, "store")
    L.force(COEFF)
    F_store = L.download_force()
    L.stout_config(3, 0.1, "recompute")
    L.force(COEFF)
    F = L.download_force()
    assert np.abs(F - F_store).max() <= 1e-13 * np.abs(F).max()
    Fref = oracle.stout_force(dims, U, 3, 0.1, COEFF)
    assert np.abs(F_store - Fref).max() < 1e-12 * np.abs(Fref).max()
    # central differences, h = 1e-5, of the smeared potential on 64 (link, generator) pairs:
    # each evaluated from the plaquettes the changed link reaches (lattice_tools), so that the
    # difference quotient keeps ~1e-8 accuracy instead of the 2e-5 a global FP64 sum of the
    # O(7e6) potential would leave it at this volume
    rng = np.random.default_rng(402)
    V = 24 ** 4
    for _ in range(64):
        mu, s, a = int(rng.integers(4)), int(rng.integers(V)), int(rng.integers(8))
        fd = lt.stout_force_fd_local(oracle, dims, U, 3, 0.1, COEFF, mu, s, a, 1e-5)
        assert abs(F_store[mu, s, a] - fd) < 1e-7, (mu, s, a, F_store[mu, s, a], fd)


def test_config4_stout_trajectory_24(config4, oracle):
    """Omelyan 2MN, 10 steps, tau 1, FP64, stored and recomputed C, against the oracle (the
    oracle merges the adjacent kicks of consecutive steps as the fused GPU sequence does)."""
    dims, U, L = config4
    P = gen.gaussian_momenta(dims, 403)
    H = {}
    for mode in ("store", "recompute"):
        L.stout_config(3, 0.1, mode)
        L.upload_links(U)
        L.upload_momenta(P)
        H[mode] = L.md_trajectory(COEFF, 10, 1.0, "2MN", fused=True)
        if mode == "store":
            Ug, Pg = L.download_links(), L.download_momenta()
    Uo, Po = U.copy(), P.copy()
    Href = oracle.md_trajectory(dims, Uo, Po, COEFF, 10, 1.0, "2MN-merged", nlev=3, rho=0.1)
    scale = abs(Href[0])
    for mode in H:
        assert abs(H[mode][0] - Href[0]) < 1e-12 * scale, (mode, H[mode], Href)
        assert abs(H[mode][2] - Href[2]) < 1e-12 * scale, (mode, H[mode], Href)
    assert np.abs(Ug - Uo).max() < 1e-10
    assert np.abs(Pg - Po).max() < 1e-9 * np.abs(Po).max()


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


def test_config5_production_48(gf):
    """48^4 FP32 bulk, 3 stout levels at rho 0.12, Omelyan 2MN 16 steps, tau 1, beta 6, links
    exp(i omega.T) with omega ~ N(0, 0.3^2).  The first trajectory from that start is far from
    equilibrium (dH ~ -4e4): its FP64 rerun agrees to a relative bound.  After 20 unconditionally
    accepted thermalization trajectories two more trajectories (device momenta) are each rerun in
    FP64 from the same links and momenta (absolute dH bound), then a reversibility run."""
    dims = (48, 48, 48, 48)
    L32 = gf.Lattice(dims, precision="f32")
    L32.links_from_algebra(gen.algebra_coefficients(dims, 0.3, 5001))
    L32.stout_config(3, 0.12, "auto")
    assert L32.stout_mode() == "store"  # 288 GB leave room for every level's C
    L64 = gf.Lattice(dims, precision="f64")
    L64.stout_config(3, 0.12, "auto")

    def rerun(sweep):
        L32.gaussian_momenta(5002, sweep)
        U0, P0 = L32.download_links(), L32.download_momenta()
        H32 = L32.md_trajectory(COEFF, 16, 1.0, "2MN", fused=True)
        L64.upload_links(U0)
        L64.upload_momenta(P0)
        del U0, P0
        H64 = L64.md_trajectory(COEFF, 16, 1.0, "2MN", fused=True)
        # the FP32-bulk Hamiltonian is an FP64 evaluation of the same links and momenta
        assert abs(H32[0] - H64[0]) < 1e-12 * abs(H64[0]), (H32, H64)
        p32, p64 = L32.plaquette(), L64.plaquette()
        assert 0.05 < p32 < 0.95
        assert abs(p32 - p64) < 1e-7, (p32, p64)
        print(f"config5 sweep {sweep}: dH32 {H32[2]:.6g} dH64 {H64[2]:.6g} plaquette {p32:.8f}", flush=True)
        return H32, H64

    # first trajectory from the start: dH of order 1e4, FP32 within 1e-4 of it (1.7e-5 measured)
    H32, H64 = rerun(0)
    assert abs(H32[2] - H64[2]) < 1e-4 * abs(H64[2]) + 3e-2, (H32, H64)
    for k in range(20):
        L32.gaussian_momenta(5002, 1000 + k)
        L32.md_trajectory(COEFF, 16, 1.0, "2MN", fused=True, diagnostics=False)
    for sweep in (1, 2):
        H32, H64 = rerun(sweep)
        # the FP32 trajectory's deviation after thermalization: 7e-4 .. 6.4e-3 measured at this
        # volume (bench.py --production); stated bound 3e-2
        assert abs(H32[2] - H64[2]) < 3e-2, (H32, H64)
    # reversibility |U(tau -> -tau) - U0| of the FP32-bulk form: 7e-7 measured; bound 2e-6
    U0 = L32.download_links()
    L32.gaussian_momenta(5002, 99)
    L32.md_trajectory(COEFF, 16, 1.0, "2MN", fused=True, diagnostics=False)
    L32.md_trajectory(COEFF, 16, -1.0, "2MN", fused=True, diagnostics=False)
    rev = float(np.abs(L32.download_links() - U0).max())
    assert rev < 2e-6, rev
    L32.close()
    L64.close()
