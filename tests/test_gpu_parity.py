"""GPU parity: libgfhip (through its C ABI) against the CPU oracle on identical seeded inputs.

FP64 tolerances follow the stated bar of relative 1e-12 for plaquette, action, force and
Hamiltonian (for dH, relative to H: the Hamiltonian is summed over the lattice and its
difference inherits the absolute rounding of the sum).  FP32-bulk results are compared
against the FP64 oracle with the tolerances written in each test.
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

    def make(dims=D4, precision="f64"):
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


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def maxdiff(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max())


# ------------------------------------------------------------------ layout, plaquette


def test_roundtrip_exact(Lat):
    # the device keeps rows 0 and 1 of each SU(3) link and rebuilds row 2 as conj(row0 x row1)
    L = Lat()
    U = gen.hot_start_links(D4, 1)
    P = gen.gaussian_momenta(D4, 2)
    L.upload_links(U)
    L.upload_momenta(P)
    M, Md = gen.links_to_matrices(U), gen.links_to_matrices(L.download_links())
    assert np.array_equal(Md[..., :2, :], M[..., :2, :])
    assert np.abs(Md[..., 2, :] - M[..., 2, :]).max() < 1e-15
    assert np.array_equal(L.download_momenta(), P)


def test_upload_rejects_non_su3(Lat, gf):
    L = Lat()
    U = gen.hot_start_links(D4, 3)
    M = gen.links_to_matrices(U).copy()
    M[1, 17, 2, 0] += 1e-3
    with pytest.raises(gf.GFError, match="links must be SU\\(3\\)"):
        L.upload_links(gen.matrices_to_links(M))
    M[1, 17, 2, 0] = np.nan
    with pytest.raises(gf.GFError, match="links must be SU"):
        L.upload_links(gen.matrices_to_links(M))

    def with_row2(M):  # row 2 = conj(row0 x row1): passes the row-2 test whatever rows 0, 1 are
        M = M.copy()
        M[..., 2, :] = np.conj(np.cross(M[..., 0, :], M[..., 1, :]))
        return gen.matrices_to_links(M)

    M = gen.links_to_matrices(U).copy()
    M[2, 5, 0, :] *= 1.0 + 1e-4  # row 0 not of unit norm
    with pytest.raises(gf.GFError, match="direction 3, site 5 row 0 is not of unit norm"):
        L.upload_links(with_row2(M))
    M = gen.links_to_matrices(U).copy()
    M[0, 40, 1, :] += 1e-4 * M[0, 40, 0, :]  # rows 0 and 1 not orthogonal
    with pytest.raises(gf.GFError, match="direction 1, site 40 rows 0 and 1 are not orthogonal"):
        L.upload_links(with_row2(M))
    # a non-unitary link is rejected and the configuration on the device is left as it was
    L.upload_links(U)
    bad = gen.links_to_matrices(U).copy()
    bad[3, 7] *= 1.01
    with pytest.raises(gf.GFError, match="links must be SU"):
        L.upload_links(gen.matrices_to_links(bad))
    assert np.array_equal(gen.links_to_matrices(L.download_links())[..., :2, :],
                          gen.links_to_matrices(U)[..., :2, :])
    # within the limit (links written in single precision) is accepted
    L.upload_links(U.astype(np.complex64).astype(np.complex128))


@pytest.mark.parametrize("dims", [D4, (2, 2, 2, 2), (4, 2, 3, 2), (2, 4, 2, 6), (8, 4, 6, 2)])
def test_plaquette_parity(Lat, oracle, dims):
    L = Lat(dims)
    assert L.plaquette() == 1.0  # cold start
    U = gen.hot_start_links(dims, 5)
    L.upload_links(U)
    assert rel(L.plaquette_sum(), oracle.plaquette_sum(dims, U)) < 1e-12
    assert rel(L.action(2.75), 2 * 2.75 * oracle.plaquette_sum(dims, U)) < 1e-12


def test_instanton_golden_gpu(Lat):
    # test/init.jl:351-371 (4x4x4x1 lattice) embedded in SU(3): (2*0.9796864531099871 + 1)/3
    dims = (4, 4, 4, 1)
    L = Lat(dims)
    L.upload_links(gen.matrices_to_links(lt.instanton_links(dims, nc=3)))
    assert abs(L.plaquette() - (2 * 0.9796864531099871 + 1) / 3) < 1e-14


def test_plaquette_16_gauge_invariant(Lat, oracle):
    dims = (16, 16, 16, 16)
    L = Lat(dims)
    U = alg_links(oracle, dims, 0.3, 8)
    L.upload_links(U)
    p = L.plaquette_sum()
    assert rel(p, oracle.plaquette_sum(dims, U)) < 1e-12
    g = lt.random_su3(np.prod(dims), np.random.default_rng(1))
    L.upload_links(lt.gauge_transform(U, dims, g))
    assert rel(L.plaquette_sum(), p) < 1e-12


# ------------------------------------------------------------------ derivative, force, exp


@pytest.mark.parametrize("dims", [D4, (4, 2, 3, 2), (16, 8, 4, 8)])
def test_dsdu_parity(Lat, oracle, dims):
    L = Lat(dims)
    U = gen.hot_start_links(dims, 11)
    L.upload_links(U)
    assert maxdiff(L.dsdu(COEFF), oracle.dsdu(dims, U, COEFF)) < 1e-13 * COEFF * 6


# stencil forms: per-thread staples (2,4,2,6: V not a multiple of 64), shared plaquettes
# (12,4,4,4: NX does not divide 64), x-rows from LDS (the rest; 64,2,2,2: one row per workgroup)
@pytest.mark.parametrize("dims", [D4, (2, 4, 2, 6), (8, 8, 8, 8), (16, 8, 4, 8), (12, 4, 4, 4), (64, 2, 2, 2)])
def test_force_parity(Lat, oracle, dims):
    L = Lat(dims)
    U = alg_links(oracle, dims, 0.3, 12)
    L.upload_links(U)
    L.force(COEFF)
    F = oracle.force(dims, U, COEFF)
    assert maxdiff(L.download_force(), F) < 1e-12 * np.abs(F).max()


def test_kick_and_exp_parity(Lat, oracle):
    L = Lat()
    U = alg_links(oracle, D4, 0.3, 13)
    P = gen.gaussian_momenta(D4, 14)
    L.upload_links(U)
    L.upload_momenta(P)
    L.force_kick(COEFF, 0.02)
    P2 = P.copy()
    oracle.kick(D4, P2, oracle.force(D4, U, COEFF), 0.02)
    assert maxdiff(L.download_momenta(), P2) < 1e-14
    L.exp_update(0.02)
    U2 = U.copy()
    oracle.exp_update(D4, U2, P2, 0.02)
    assert maxdiff(L.download_links(), U2) < 1e-14


def test_fused_md_update(Lat, oracle):
    L = Lat()
    U = alg_links(oracle, D4, 0.3, 15)
    P = gen.gaussian_momenta(D4, 16)
    L.upload_links(U)
    L.upload_momenta(P)
    L.md_update(COEFF, 0.03, 0.02)
    U2, P2 = U.copy(), P.copy()
    oracle.kick(D4, P2, oracle.force(D4, U, COEFF), 0.03)
    oracle.exp_update(D4, U2, P2, 0.02)
    assert maxdiff(L.download_momenta(), P2) < 1e-14
    assert maxdiff(L.download_links(), U2) < 1e-14


def test_exp_large_and_tiny_momenta(Lat, oracle):
    L = Lat()
    # c1 ~ 2 scale^2 per link (~5 scale^2 the largest of a wave): 3.0 the closed form, 0.45 both
    # sides of c1 = 3/4 within a wave, 0.2 the 18-term series, 0.09 waves on both sides of the
    # 13-term limit (c1 = 0.046875), 0.08 the 13-term series, 0.03, 1e-9 and 0 the 11-term one
    for scale in (3.0, 0.45, 0.2, 0.09, 0.08, 0.03, 1e-9, 0.0):
        P = scale * gen.gaussian_momenta(D4, 17)
        L.set_cold()
        L.upload_momenta(P)
        L.exp_update(1.0)
        U = gen.cold_links(D4)
        oracle.exp_update(D4, U, P, 1.0)
        assert maxdiff(L.download_links(), U) < 5e-15 * max(1.0, scale)


# ------------------------------------------------------------------ MD


@pytest.mark.parametrize("integ", ["QPQ", "PQP", "2MN"])
@pytest.mark.parametrize("fused", [False, True])
def test_md_trajectory_parity(Lat, oracle, integ, fused):
    L = Lat()
    U = gen.hot_start_links(D4, 21)
    P = gen.gaussian_momenta(D4, 22)
    L.upload_links(U)
    L.upload_momenta(P)
    H = L.md_trajectory(COEFF, 10, 1.0, integ, fused=fused)
    Href = oracle.md_trajectory(D4, U, P, COEFF, 10, 1.0, integ)
    assert rel(H[0], Href[0]) < 1e-12 and rel(H[1], Href[1]) < 1e-12
    assert abs(H[2] - Href[2]) < 1e-12 * abs(Href[0])
    assert maxdiff(L.download_links(), U) < 1e-11
    assert maxdiff(L.download_momenta(), P) < 1e-10


def test_md_cold_start_gpu(Lat):
    L = Lat()
    H = L.md_trajectory(COEFF, 5, 1.0, "QPQ")
    assert H[2] == 0.0 and L.plaquette() == 1.0


def test_md_reversibility_gpu(Lat):
    L = Lat()
    U0 = gen.hot_start_links(D4, 23)
    P0 = gen.gaussian_momenta(D4, 24)
    L.upload_links(U0)
    L.upload_momenta(P0)
    L.md_trajectory(COEFF, 5, 0.2, "2MN", fused=True)
    L.md_trajectory(COEFF, 5, -0.2, "2MN", fused=True)
    assert maxdiff(L.download_links(), U0) < 2e-12
    assert maxdiff(L.download_momenta(), P0) < 2e-12


# ------------------------------------------------------------------ flow


def test_flow_parity(Lat, oracle):
    dims = (4, 4, 4, 6)
    L = Lat(dims)
    U = alg_links(oracle, dims, 0.4, 31)
    L.upload_links(U)
    E = L.flow(0.01, 20, 5)
    Eref = oracle.flow(dims, U, 0.01, 20, 5)
    assert np.all(np.abs(E - Eref) < 1e-12 * np.abs(Eref))
    assert maxdiff(L.download_links(), U) < 1e-13


# ------------------------------------------------------------------ stout


@pytest.mark.parametrize("mode", ["store", "recompute"])
def test_stout_smear_and_force_parity(Lat, oracle, mode):
    L = Lat()
    U = alg_links(oracle, D4, 0.3, 41)
    L.upload_links(U)
    L.stout_config(3, 0.1, mode)
    assert maxdiff(L.stout_smear(), oracle.stout_smear(D4, U, 3, 0.1)) < 1e-14
    L.force(COEFF)
    F = oracle.stout_force(D4, U, 3, 0.1, COEFF)
    assert maxdiff(L.download_force(), F) < 1e-12 * np.abs(F).max()
    H = L.hamiltonian(COEFF)
    assert rel(H, oracle.stout_potential(D4, U, 3, 0.1, COEFF)) < 1e-12  # momenta are zero


# stencil forms of the backward stencil: x-rows from LDS (4^4), the site's own U and Lam from LDS
# (12,4,4,4: NX does not divide 64), every load from memory (2,4,2,6: V not a multiple of 64)
# rho 0.12 keeps every exponent's spectral radius below 1 (the series coefficients), rho 0.5 takes
# part of the links past c1 = 3/4 (the closed form) within the same waves
@pytest.mark.parametrize("rho", [0.12, 0.5])
@pytest.mark.parametrize("dims", [D4, (12, 4, 4, 4), (2, 4, 2, 6)])
def test_stout_backward_level_parity(Lat, oracle, dims, rho):
    L = Lat(dims)
    U = gen.hot_start_links(dims, 42)
    D = gen.hot_start_links(dims, 43)  # any link-shaped derivative
    L.upload_links(U)
    L.stout_config(1, rho, "store")
    L.stout_smear(download=False)
    assert maxdiff(L.stout_backward(D), oracle.stout_backward(dims, U, rho, D)) < 1e-13


# the fused 2MN sequence: the thin-link backward stencil forms the force and the link update, in
# the x-row (4^4) and shared (12,4,4,4) stencil forms
@pytest.mark.parametrize("dims", [D4, (12, 4, 4, 4)])
def test_stout_md_parity(Lat, oracle, dims):
    L = Lat(dims)
    U = alg_links(oracle, dims, 0.3, 44)
    P = gen.gaussian_momenta(dims, 45)
    L.upload_links(U)
    L.upload_momenta(P)
    L.stout_config(3, 0.1, "store")
    H = L.md_trajectory(COEFF, 4, 0.2, "2MN", fused=True)
    Href = oracle.md_trajectory(dims, U, P, COEFF, 4, 0.2, "2MN", nlev=3, rho=0.1)
    assert abs(H[2] - Href[2]) < 1e-12 * abs(Href[0])
    assert maxdiff(L.download_links(), U) < 1e-12


# ------------------------------------------------------------------ FP32 bulk


def test_fp32_bulk_against_fp64_oracle(Lat, oracle):
    L = Lat(precision="f32")
    U = alg_links(oracle, D4, 0.3, 51)
    P = gen.gaussian_momenta(D4, 52)
    L.upload_links(U)
    assert rel(L.plaquette_sum(), oracle.plaquette_sum(D4, U)) < 2e-7
    L.force(COEFF)
    F = oracle.force(D4, U, COEFF)
    assert maxdiff(L.download_force(), F) < 2e-6 * np.abs(F).max()
    L.upload_momenta(P)
    H = L.md_trajectory(COEFF, 10, 1.0, "QPQ", fused=True)
    Href = oracle.md_trajectory(D4, U.copy(), P.copy(), COEFF, 10, 1.0, "QPQ")
    assert abs(H[2] - Href[2]) < 1e-3


# FP32 bulk (quad-plane storage) in the three stencil forms: 3-level stout force and stout force
# + kick + link update (the fused 2MN sequence) against the FP64 oracle on the same float-rounded
# links.  Measured: force 4.7e-6 ... 6.1e-6 of its maximum, dH within 7e-7 ... 3.8e-5; bars 2e-5
# and 1e-4.
@pytest.mark.parametrize("dims", [D4, (12, 4, 4, 4), (2, 4, 2, 6)])
def test_fp32_bulk_stout_forms(Lat, oracle, dims):
    L = Lat(dims, precision="f32")
    U = alg_links(oracle, dims, 0.3, 53)
    P = gen.gaussian_momenta(dims, 54)
    L.upload_links(U)
    U = L.download_links()  # the float-rounded links, for both sides
    L.stout_config(3, 0.1, "store")
    L.force(COEFF)
    F = oracle.stout_force(dims, U, 3, 0.1, COEFF)
    assert maxdiff(L.download_force(), F) < 2e-5 * np.abs(F).max()
    L.upload_momenta(P)
    P = L.download_momenta()
    H = L.md_trajectory(COEFF, 4, 0.2, "2MN", fused=True)
    Href = oracle.md_trajectory(dims, U.copy(), P.copy(), COEFF, 4, 0.2, "2MN", nlev=3, rho=0.1)
    assert abs(H[2] - Href[2]) < 1e-4


# ------------------------------------------------------------------ workgroup orderings
# Extents whose xy-plane holds a multiple of 64 sites run the stencil kernels in a non-linear
# group order: the t-sweep tiles when the (xy-group, z) plane splits into 8 tiles ((8,8,8,8),
# (16,8,4,8), (24,8,8,2); (32,32,16,2) takes 2 x 16 tiles split into 2 sub-tiles of 2 x 8
# columns), else 4D Morton order for power-of-two extents ((32,4,2,4), (8,8,2,2)); these
# cases pin both paths.


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (16, 8, 4, 8), (24, 8, 8, 2), (32, 32, 16, 2), (32, 4, 2, 4),
                                  (8, 8, 2, 2)])
def test_morton_path_md_and_flow(Lat, oracle, dims):
    L = Lat(dims)
    U = alg_links(oracle, dims, 0.3, 61)
    P = gen.gaussian_momenta(dims, 62)
    L.upload_links(U)
    L.upload_momenta(P)
    H = L.md_trajectory(COEFF, 3, 0.3, "2MN", fused=True)
    Href = oracle.md_trajectory(dims, U, P, COEFF, 3, 0.3, "2MN")
    assert abs(H[2] - Href[2]) < 1e-12 * abs(Href[0])
    assert maxdiff(L.download_links(), U) < 1e-12
    E = L.flow(0.02, 4, 2)
    Eref = oracle.flow(dims, U, 0.02, 4, 2)
    assert np.all(np.abs(E - Eref) < 1e-12 * np.abs(Eref))
    assert rel(L.plaquette_sum(), oracle.plaquette_sum(dims, U)) < 1e-12


def test_morton_path_stout(Lat, oracle):
    dims = (8, 8, 8, 8)
    L = Lat(dims)
    U = alg_links(oracle, dims, 0.3, 63)
    L.upload_links(U)
    L.stout_config(3, 0.1, "recompute")
    L.force(COEFF)
    F = oracle.stout_force(dims, U, 3, 0.1, COEFF)
    assert maxdiff(L.download_force(), F) < 1e-12 * np.abs(F).max()


def test_stout_auto_mode_stores_when_it_fits(Lat, oracle):
    L = Lat()
    L.stout_config(2, 0.1, "auto")
    assert L.stout_mode() == "store"
    U = alg_links(oracle, D4, 0.3, 64)
    L.upload_links(U)
    L.force(COEFF)
    F = oracle.stout_force(D4, U, 2, 0.1, COEFF)
    assert maxdiff(L.download_force(), F) < 1e-12 * np.abs(F).max()


# ------------------------------------------------------------------ volumes that are not a multiple of 64
# (the stencil kernels then run their unshared form: no workgroup barrier, no LDS products)


@pytest.mark.parametrize("mode", ["store", "recompute"])
def test_unaligned_volume_stout_flow_md(Lat, oracle, mode):
    dims = (3, 4, 5, 2)  # V = 120
    print(f"[parity] unaligned volume {dims} stout {mode}", flush=True)
    L = Lat(dims)
    U = alg_links(oracle, dims, 0.3, 71)
    P = gen.gaussian_momenta(dims, 72)
    L.upload_links(U)
    L.stout_config(2, 0.1, mode)
    assert maxdiff(L.stout_smear(), oracle.stout_smear(dims, U, 2, 0.1)) < 1e-14
    L.force(COEFF)
    F = oracle.stout_force(dims, U, 2, 0.1, COEFF)
    assert maxdiff(L.download_force(), F) < 1e-12 * np.abs(F).max()
    L.stout_config(0, 0.0, "store")
    L.upload_momenta(P)
    H = L.md_trajectory(COEFF, 3, 0.3, "2MN", fused=True)
    U2, P2 = U.copy(), P.copy()
    Href = oracle.md_trajectory(dims, U2, P2, COEFF, 3, 0.3, "2MN")
    assert abs(H[2] - Href[2]) < 1e-12 * abs(Href[0])
    assert maxdiff(L.download_links(), U2) < 1e-12
    E = L.flow(0.02, 4, 2)
    Eref = oracle.flow(dims, U2, 0.02, 4, 2)
    assert np.all(np.abs(E - Eref) < 1e-12 * np.abs(Eref))
    assert maxdiff(L.download_links(), U2) < 1e-13


# ------------------------------------------------------------------ precision-study rounding


def _gram_schmidt_rows(M):
    """rows 0 and 1 of each 3x3 (..., 3, 3) orthonormalized in double, row 2 = conj(r0 x r1)"""
    r0 = M[..., 0, :] / np.linalg.norm(M[..., 0, :], axis=-1, keepdims=True)
    r1 = M[..., 1, :] - np.sum(np.conj(r0) * M[..., 1, :], axis=-1, keepdims=True) * r0
    r1 = r1 / np.linalg.norm(r1, axis=-1, keepdims=True)
    return np.stack([r0, r1, np.conj(np.cross(r0, r1))], axis=-2)


def test_round_to_float_is_the_fp32_storage_rounding(Lat, oracle, gf):
    """gf_round_to_float on an FP64 context: momenta equal numpy's float32 rounding; links equal
    rows 0-1 of the double Gram-Schmidt projection rounded to float, with or without the
    projection; an FP32 context refuses it"""
    dims = (4, 4, 4, 6)
    U = alg_links(oracle, dims, 0.3, 61)
    P = gen.gaussian_momenta(dims, 62)
    L = Lat(dims)
    L.upload_links(U)
    L.upload_momenta(P)
    Ud = gen.links_to_matrices(L.download_links())
    L.round_to_float(momenta=True)
    assert np.array_equal(L.download_momenta(), P.astype(np.float32).astype(np.float64))
    assert np.array_equal(gen.links_to_matrices(L.download_links()), Ud)  # links untouched
    for project in (False, True):
        L.upload_links(U)
        L.round_to_float(links=True, project=project)
        got = gen.links_to_matrices(L.download_links())
        M = _gram_schmidt_rows(Ud) if project else Ud.copy()
        M[..., :2, :] = M[..., :2, :].astype(np.complex64).astype(np.complex128)
        M[..., 2, :] = np.conj(np.cross(M[..., 0, :], M[..., 1, :]))
        assert np.abs(got - M).max() < 1e-15, project
    L32 = Lat(dims, "f32")
    with pytest.raises(gf.GFError):
        L32.round_to_float(momenta=True)
