"""Golden vectors (tests/golden/vectors_4x4x4x4.npz, written by tests/golden/make_goldens.py).

The oracle must reproduce its stored output, and the stored data must satisfy checks that do
not go through the oracle: numpy brute-force plaquette and topological charge, scipy expm for
the link update, the TA relation between dS/dU and the force, the Hamiltonian's parts, the
monotone flow energy, SU(3) heatbath output and Gaussian moments of the Philox momenta.
"""
import os
import sys

import numpy as np
import pytest
from scipy.linalg import expm

import lattice_tools as lt

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_goldens  # noqa: E402

D = make_goldens.DIMS


@pytest.fixture(scope="module")
def G():
    return make_goldens.load()


def test_oracle_reproduces_goldens(G, oracle):
    new = make_goldens.compute()
    assert set(new) == set(G)
    for k in G:
        scale = max(float(np.abs(G[k]).max()), 1.0)
        assert np.abs(new[k] - G[k]).max() <= 1e-13 * scale, k


def test_golden_plaquette_and_action_numpy(G):
    assert G["plaq"] == pytest.approx(lt.plaquette_field_sum(G["U0"], D), rel=1e-13)
    # potential -(beta/N_c) sum Re tr P with coeff = beta/2 (molecular_dynamics.jl:247-249)
    assert G["potential"] == pytest.approx(-2 * float(G["coeff"]) / 3 * G["plaq"], rel=1e-14)
    H0, H1, dH = G["H_qpq"]
    assert H0 == pytest.approx(G["potential"] + 0.5 * (G["P0"] ** 2).sum(), rel=1e-13)
    assert dH == pytest.approx(H1 - H0, abs=1e-9)
    assert abs(dH) < 5.0  # a 4-step trajectory of tau 0.2 at beta 6 stays near the surface


def test_golden_force_is_ta_of_dsdu(G):
    U = lt.gen.links_to_matrices(G["U0"])
    S = lt.gen.links_to_matrices(G["dsdu"])
    F = lt.coeffs_to_matrix(G["force"])
    assert np.abs(F - (-1.0 / 3.0) * lt.ta_matrix(U @ S)).max() < 1e-12


def test_golden_kick_and_exp_scipy(G):
    assert np.abs(G["P1"] - (G["P0"] + 0.02 * G["force"])).max() < 1e-15
    X = lt.coeffs_to_matrix(0.02 * G["P1"])
    U0 = lt.gen.links_to_matrices(G["U0"])
    U1 = lt.gen.links_to_matrices(G["U1"])
    for mu in range(4):
        for s in range(0, U0.shape[1], 7):
            assert np.abs(U1[mu, s] - expm(X[mu, s]) @ U0[mu, s]).max() < 1e-14


def _su3_defect(U):
    M = lt.gen.links_to_matrices(U)
    unit = np.abs(M @ lt.dg(M) - np.eye(3)).max()
    det = np.abs(np.linalg.det(M) - 1).max()
    return max(unit, det)


def test_golden_links_are_su3(G):
    for k in ("U0", "U1", "Uqpq", "Ustout", "W", "Uhb"):
        assert _su3_defect(G[k]) < 1e-13, k
    # the heatbath changed every link it visited; overrelaxation kept them SU(3)
    assert np.all(np.abs(G["Uhb"] - G["U0"]).max(axis=2) > 1e-6)


def test_golden_topology_numpy(G):
    V = int(np.prod(D))
    for i, m in enumerate(("plaquette", "clover")):
        assert G["Q"][i] == pytest.approx(lt.topo_density_numpy(G["U0"], D, m).sum(), rel=1e-12, abs=1e-13)
    assert 0 < G["Eclov"] < 18 and V > 0


def test_golden_flow_energy_decreases(G):
    E = G["E_flow"]
    assert E.shape == (5,) and np.all(np.diff(E) < 0)
    assert G["plaq"] < lt.plaquette_field_sum(G["W"], D)


def test_golden_rng_moments(G):
    p = G["Prng"].ravel()
    assert abs(p.mean()) < 0.05 and abs(p.var() - 1) < 0.05
