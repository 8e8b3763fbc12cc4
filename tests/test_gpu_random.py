"""Device random fields (Philox4x32-10 per-link streams) against the oracle's restatement.

The integer stream is the same function on both sides, so the momenta and hot-start links
agree to the rounding of log/sqrt/sin/cos (FP64: 1e-13 absolute); FP32 storage rounds the
results (1e-6).  Statistics at 16^4 check the distribution itself.
"""
import numpy as np
import pytest

import lattice_tools as lt

pytestmark = pytest.mark.gpu

gen = lt.gen


@pytest.fixture(scope="module")
def Lat(gf):
    made = []

    def make(dims, precision="f64"):
        L = gf.Lattice(dims, precision=precision)
        made.append(L)
        return L

    yield make
    for L in made:
        L.close()


@pytest.mark.parametrize("dims,seed,sweep", [((4, 4, 4, 4), 1, 0), ((3, 5, 2, 4), 2 ** 40 + 7, 3), ((8, 8, 8, 8), 99, 2 ** 33)])
def test_gaussian_momenta_match_oracle(Lat, oracle, dims, seed, sweep):
    print(f"[random] momenta {dims} seed={seed} sweep={sweep}", flush=True)
    L = Lat(dims)
    L.gaussian_momenta(seed, sweep, sigma=1.5)
    ref = oracle.gaussian_momenta(dims, seed, sweep, sigma=1.5)
    assert np.abs(L.download_momenta() - ref).max() < 1e-13


@pytest.mark.parametrize("dims,seed", [((4, 4, 4, 4), 5), ((3, 5, 2, 4), 2 ** 50 + 1)])
def test_hot_start_matches_oracle(Lat, oracle, dims, seed):
    print(f"[random] hot start {dims} seed={seed}", flush=True)
    L = Lat(dims)
    L.hot_start(seed)
    ref = oracle.hot_links(dims, seed)
    assert np.abs(L.download_links() - ref).max() < 1e-13
    assert L.plaquette_sum() == pytest.approx(oracle.plaquette_sum(dims, ref), rel=1e-12, abs=1e-11)


def test_fp32_random_fields(Lat, oracle):
    dims = (4, 4, 4, 4)
    L = Lat(dims, "f32")
    L.gaussian_momenta(11, 1)
    assert np.abs(L.download_momenta() - oracle.gaussian_momenta(dims, 11, 1)).max() < 5e-6
    L.hot_start(12)
    assert np.abs(L.download_links() - oracle.hot_links(dims, 12)).max() < 1e-6


def test_momenta_statistics_and_streams(Lat):
    dims = (16, 16, 16, 16)
    L = Lat(dims)
    L.gaussian_momenta(2024, 0)
    a = L.download_momenta()
    n = a.size
    assert abs(a.mean()) < 5 / np.sqrt(n)
    assert abs(a.var() - 1) < 5 * np.sqrt(2 / n)
    assert abs((a ** 4).mean() - 3) < 5 * np.sqrt(96 / n)
    # the 8 coefficients of a link are uncorrelated
    c = np.corrcoef(a.reshape(-1, 8).T)
    assert np.abs(c - np.eye(8)).max() < 5 / np.sqrt(a.shape[0] * a.shape[1])
    # same stream twice -> identical; another sweep or seed -> a different stream
    L.gaussian_momenta(2024, 0)
    assert np.array_equal(L.download_momenta(), a)
    L.gaussian_momenta(2024, 1)
    b = L.download_momenta()
    assert abs(np.corrcoef(a.ravel(), b.ravel())[0, 1]) < 5 / np.sqrt(n)
    L.gaussian_momenta(2025, 0)
    assert not np.array_equal(L.download_momenta(), a)
    # kinetic energy sum p^2 / 2 has mean 4 V 8 / 2
    assert L.momenta_dot() / 2 == pytest.approx(16 * np.prod(dims), rel=5 * np.sqrt(2 / n))


def test_hot_start_is_su3_and_disordered(Lat):
    dims = (8, 8, 8, 8)
    L = Lat(dims)
    L.hot_start(3)
    M = gen.links_to_matrices(L.download_links())
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14
    assert abs(L.plaquette()) < 0.02
