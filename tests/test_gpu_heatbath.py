"""GPU heatbath / overrelaxation sweeps against the oracle on the same Philox streams.

Same streams and the same floating-point recipe on both sides: the links after a sweep agree
to rounding amplified by the Kennedy-Pendleton map (a1 = 1 - delta near +-1 enters through
sqrt(1 - a1^2), which multiplies last-bit differences of log/cos between the device and host
math libraries by up to ~1e5): 1e-9 absolute after two sweeps.  An accept/reject decision
flipping on a last-bit difference would show up as an O(1) link difference.
Physics: strong-coupling plaquette, and overrelaxation keeps the action.
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


@pytest.mark.parametrize("dims,beta", [((4, 4, 4, 4), 5.7), ((4, 2, 6, 4), 1.0), ((8, 8, 8, 8), 6.0)])
def test_heatbath_matches_oracle(Lat, oracle, dims, beta):
    print(f"[heatbath] {dims} beta={beta}", flush=True)
    U = oracle.hot_links(dims, 17)
    L = Lat(dims)
    L.upload_links(U)
    W = U.copy()
    for sw in range(2):
        f = L.heatbath(beta, 99, sw)
        fo = oracle.heatbath_sweep(dims, W, beta, 99, sw)
        assert f == fo == 0
    assert np.abs(L.download_links() - W).max() < 1e-9
    assert L.plaquette_sum() == pytest.approx(oracle.plaquette_sum(dims, W), rel=1e-11)


# 8^4: the tiled thread order (V a multiple of 4096); 4^4: linear order
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 8, 8, 8)])
def test_overrelaxation_matches_oracle_and_keeps_action(Lat, oracle, dims):
    print(f"[overrelaxation] {dims}", flush=True)
    U = oracle.hot_links(dims, 18)
    W = U.copy()
    for sw in range(3):
        oracle.heatbath_sweep(dims, W, 6.0, 1, sw)
    L = Lat(dims)
    L.upload_links(W)
    p0 = L.plaquette_sum()
    L.overrelaxation(5, 0)
    oracle.heatbath_sweep(dims, W, 6.0, 5, 0, overrelax=True)
    assert np.abs(L.download_links() - W).max() < 1e-12
    assert L.plaquette_sum() == pytest.approx(p0, rel=1e-13)


def test_heatbath_strong_coupling_and_fp32(Lat):
    dims = (8, 8, 8, 8)
    for prec in ("f64", "f32"):
        L = Lat(dims, prec)
        L.set_cold()
        beta, ps = 0.3, []
        for sw in range(60):
            assert L.heatbath(beta, 123, sw) == 0
            if sw >= 10:
                ps.append(L.plaquette())
        err = np.std(ps) / np.sqrt(len(ps))
        assert abs(np.mean(ps) - (beta / 18 + beta ** 2 / 216)) < 4 * err + 2e-4, prec


def test_heatbath_rejects_odd_extents(Lat, gf):
    L = Lat((3, 4, 4, 4))
    with pytest.raises(gf.GFError, match="even lattice extents"):
        L.heatbath(6.0, 1, 0)


def test_api_heatbath_object(gf):
    from gaugefields_jl_amd import api

    U = api.gauge_configuration((4, 4, 4, 4), start="cold")
    action = api.GaugeAction(U)
    plaq = api.make_loops_fromname("plaquette")
    action.push(5.7 / 2, plaq + api.dagger(plaq))
    h = api.Heatbath_update(U, action, seed=3)
    assert h.beta == pytest.approx(5.7)
    for _ in range(3):
        api.heatbath(U, h)
        api.overrelaxation(U, h)
    assert h.sweep == 3 and h.overrelaxation_sweep == 3 and h.failures == 0
    assert 0.3 < api.measure_plaquette(U) < 0.8
    with pytest.raises(api.ArgumentError):
        api.Heatbath(U, 6.0, ITERATION_MAX=0)
    odd = api.gauge_configuration((3, 4, 4, 4), start="cold")
    with pytest.raises(api.ArgumentError, match="even lattice extents"):
        api.Heatbath(odd, 6.0)
