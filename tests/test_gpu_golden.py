"""libgfhip against the committed golden vectors (tests/golden/vectors_4x4x4x4.npz).

Every row of the path on the same stored 4^4 inputs: plaquette, action, dS/dU, force, kick and
exp update, a QPQ trajectory, one stout level forward and backward, five flow steps, the four
topological charges, the clover energy, a heatbath + overrelaxation sweep and the Philox
momenta.  FP64 bar: relative 1e-12 on sums and forces, 1e-13 absolute on links (1e-9 for the
heatbath, whose Kennedy-Pendleton acceptance amplifies rounding of the staple).  FP32 bulk is
checked on the plaquette and the force against the same FP64 data.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_goldens  # noqa: E402

D = make_goldens.DIMS
TOPO = make_goldens.TOPO


@pytest.fixture(scope="module")
def G():
    return make_goldens.load()


@pytest.fixture
def L(gf):
    lat = gf.Lattice(D)
    yield lat
    lat.close()


def md(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max())


def test_golden_plaquette_action_force(L, G):
    print("[golden] plaquette/action/dsdu/force", flush=True)
    c = float(G["coeff"])
    L.upload_links(G["U0"])
    assert L.plaquette_sum() == pytest.approx(float(G["plaq"]), rel=1e-12)
    # evaluate_GaugeAction (GaugeActions.jl:132-142) is 2 coeff S_P; the MD potential is -(2 coeff/3) S_P
    assert L.action(c) == pytest.approx(2 * c * float(G["plaq"]), rel=1e-12)
    L.clear_momenta()
    assert L.hamiltonian(c) == pytest.approx(float(G["potential"]), rel=1e-12)
    assert md(L.dsdu(c), G["dsdu"]) < 1e-12 * np.abs(G["dsdu"]).max()
    L.force(c)
    assert md(L.download_force(), G["force"]) < 1e-12 * np.abs(G["force"]).max()


def test_golden_kick_exp_and_trajectory(L, G):
    print("[golden] kick/exp/QPQ", flush=True)
    c = float(G["coeff"])
    L.upload_links(G["U0"])
    L.upload_momenta(G["P0"])
    L.force_kick(c, 0.02)
    assert md(L.download_momenta(), G["P1"]) < 1e-14
    L.exp_update(0.02)
    assert md(L.download_links(), G["U1"]) < 1e-14
    L.upload_links(G["U0"])
    L.upload_momenta(G["P0"])
    H = L.md_trajectory(c, 4, 0.2, "QPQ")
    H0 = float(G["H_qpq"][0])
    assert H[0] == pytest.approx(H0, rel=1e-12)
    assert abs(H[2] - G["H_qpq"][2]) < 1e-12 * abs(H0)
    assert md(L.download_links(), G["Uqpq"]) < 1e-12


def test_golden_stout_level(L, G):
    print("[golden] stout", flush=True)
    L.upload_links(G["U0"])
    L.stout_config(1, 0.12, "store")
    assert md(L.stout_smear(), G["Ustout"]) < 1e-13
    assert md(L.stout_backward(G["Dout"]), G["Din"]) < 1e-13


def test_golden_flow_and_field_strengths(L, G):
    print("[golden] flow/topology", flush=True)
    L.upload_links(G["U0"])
    for i, m in enumerate(TOPO):
        assert abs(L.topological_charge(m) - G["Q"][i]) <= 1e-12 * max(abs(G["Q"][i]), 1.0), m
    assert L.energy_density_clover() == pytest.approx(float(G["Eclov"]), rel=1e-12)
    E = L.flow(0.02, 5, 1)
    assert np.all(np.abs(E - G["E_flow"]) < 1e-12 * np.abs(G["E_flow"]))
    assert md(L.download_links(), G["W"]) < 1e-13


def test_golden_heatbath_and_rng(L, G):
    print("[golden] heatbath/rng", flush=True)
    L.upload_links(G["U0"])
    L.heatbath(6.0, 104, 0)
    L.overrelaxation(104, 1)
    assert md(L.download_links(), G["Uhb"]) < 1e-9
    L.gaussian_momenta(105, 3)
    assert md(L.download_momenta(), G["Prng"]) < 1e-13


def test_golden_fp32_bulk(gf, G):
    print("[golden] fp32", flush=True)
    c = float(G["coeff"])
    with gf.Lattice(D, precision="f32") as L32:
        L32.upload_links(G["U0"])
        assert L32.plaquette_sum() == pytest.approx(float(G["plaq"]), rel=1e-6)
        L32.force(c)
        assert md(L32.download_force(), G["force"]) < 1e-5 * np.abs(G["force"]).max()
