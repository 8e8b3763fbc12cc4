"""GPU field strengths: topological charge (plaquette, clover, rectangle, improved), clover
energy density and the flow with measurements, through the C ABI, against the oracle.

FP64: densities to absolute 1e-13 per site, sums to relative 1e-12 (absolute 1e-12 where
the charge itself is near zero).  FP32 bulk: links are stored in float (rows 0 and 1; row 2
rebuilt in float), the loop products and sums run in double, so the comparison is against the
oracle on the float-rounded links to 1e-6 relative, and against the FP64 links to 1e-5.
"""
import numpy as np
import pytest

import lattice_tools as lt

pytestmark = pytest.mark.gpu

gen = lt.gen
METHODS = ["plaquette", "clover", "rectangle", "improved"]


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


def alg_links(oracle, dims, sigma, seed):
    U = gen.cold_links(dims)
    oracle.exp_update(dims, U, gen.algebra_coefficients(dims, sigma, seed), 1.0)
    return U


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (3, 4, 5, 2), (8, 8, 8, 8)])
def test_topological_charge_parity(Lat, oracle, dims):
    print(f"[fstrength] parity {dims}", flush=True)
    U = alg_links(oracle, dims, 0.6, 31)
    L = Lat(dims)
    L.upload_links(U)
    for m in METHODS:
        Qref, qref = oracle.topological_charge_density(dims, U, m)
        q = L.topological_charge_density(m)
        Q = L.topological_charge(m)
        assert np.abs(q - qref).max() < 1e-13, m
        assert abs(Q - Qref) <= 1e-12 * max(abs(Qref), 1.0), m
    Eref = oracle.energy_density_clover(dims, U)
    assert L.energy_density_clover() == pytest.approx(Eref, rel=1e-12)


def test_constant_abelian_field_closed_form(Lat):
    dims = (4, 6, 6, 4)
    U, t, tp = lt.abelian_flux_links(dims, 2, -1)
    V = int(np.prod(dims))
    L = Lat(dims)
    L.upload_links(U)
    qc = np.sin(t) * np.sin(tp) / (2 * np.pi ** 2)
    qr = np.sin(2 * t) * np.sin(2 * tp) / np.pi ** 2
    expect = {"plaquette": qc, "clover": qc, "rectangle": qr, "improved": 5 / 3 * qc - 1 / 12 * qr}
    for m in METHODS:
        assert L.topological_charge(m) == pytest.approx(V * expect[m], rel=1e-12)
        assert np.allclose(L.topological_charge_density(m), expect[m], rtol=1e-12, atol=1e-15)
    assert L.energy_density_clover() == pytest.approx(2 * (np.sin(t) ** 2 + np.sin(tp) ** 2), rel=1e-12)


def test_instanton_charge_embeddings(Lat, oracle):
    dims = (4, 4, 4, 4)
    L = Lat(dims)
    vals = []
    for blk in [(0, 1), (0, 2), (1, 2)]:
        U = gen.matrices_to_links(lt.instanton_links(dims, 3, blk))
        L.upload_links(U)
        vals.append([L.topological_charge(m) for m in ("plaquette", "clover", "improved")])
        ref = [oracle.topological_charge_density(dims, U, m)[0] for m in ("plaquette", "clover", "improved")]
        assert np.allclose(vals[-1], ref, rtol=1e-12, atol=1e-13)
    assert np.allclose(vals, vals[0], rtol=1e-12, atol=1e-13)


def test_fp32_bulk_topological_charge(Lat, oracle):
    dims = (4, 4, 4, 4)
    U = alg_links(oracle, dims, 0.6, 33)
    U32 = U.astype(np.complex64).astype(np.complex128)
    L = Lat(dims, "f32")
    L.upload_links(U)
    # the float links keep rows 0 and 1 and rebuild row 2 in float: ~1e-7 relative per link
    for m in METHODS:
        Qr32, qr32 = oracle.topological_charge_density(dims, U32, m)
        assert np.abs(L.topological_charge_density(m) - qr32).max() < 1e-6 * max(np.abs(qr32).max(), 1e-3), m
        Qr64 = oracle.topological_charge_density(dims, U, m)[0]
        assert abs(L.topological_charge(m) - Qr64) < 1e-5 * max(abs(Qr64), 1.0), m
    assert L.energy_density_clover() == pytest.approx(oracle.energy_density_clover(dims, U32), rel=1e-6)


def test_flow_measure_matches_oracle(Lat, oracle):
    # samples/measurements/topologicalcharge.jl:236-245: flow, then plaquette, Q_plaq,
    # Q_clover, Q_improved (and the clover energy, energydensity.jl:112-118) every few steps
    dims = (4, 4, 4, 4)
    print("[fstrength] flow_measure", flush=True)
    U = alg_links(oracle, dims, 0.5, 35)
    L = Lat(dims)
    L.upload_links(U)
    eps, nsteps, every = 0.02, 12, 4
    rows = L.flow_measure(eps, nsteps, every)
    assert rows.shape == (3, 7)
    W = U.copy()
    V = int(np.prod(dims))
    for k in range(3):
        oracle.flow(dims, W, eps, every)
        r = rows[k]
        assert r[0] == pytest.approx(eps * every * (k + 1), rel=1e-15)
        assert r[1] == pytest.approx(oracle.plaquette_sum(dims, W) / (18 * V), rel=1e-12)
        assert r[2] == pytest.approx(oracle.energy_density(dims, W), rel=1e-11)
        assert r[3] == pytest.approx(oracle.energy_density_clover(dims, W), rel=1e-11)
        for j, m in ((4, "plaquette"), (5, "clover"), (6, "improved")):
            Qref = oracle.topological_charge_density(dims, W, m)[0]
            assert abs(r[j] - Qref) <= 1e-11 * max(abs(Qref), 1.0), m
    # the flow smooths: the clover energy decreases monotonically
    assert np.all(np.diff(rows[:, 3]) < 0)


def test_api_topological_charge(gf):
    from gaugefields_jl_amd import api

    U = api.gauge_configuration((4, 4, 4, 4), start="hot", seed=4)
    q = api.topological_charge_density(U, method="clover")
    assert q.shape == (4, 4, 4, 4)
    assert api.topological_charge(U, method="clover") == pytest.approx(q.sum(), rel=1e-12, abs=1e-12)
    with pytest.raises(api.ArgumentError, match="supported topological_charge methods"):
        api.topological_charge(U, method="rect")
    with pytest.raises(api.ArgumentError, match="supported topological_charge_density methods"):
        api.topological_charge_density(U, method="rect")
    cold = api.gauge_configuration((4, 4, 4, 4), start="cold")
    assert api.topological_charge(cold, method="improved") == 0.0
    assert api.calculate_energy_density(cold) == 0.0
