"""The host-side API mirror: argument validation on the CPU, behaviour through libgfhip on the GPU.

Mirrors the reference's own suites test/high_level_api.jl and test/md_driver.jl where they
apply to the SU(3) plaquette path.
"""
import math

import numpy as np
import pytest

import lattice_tools as lt

gen = lt.gen
D4 = (4, 4, 4, 4)


@pytest.fixture(scope="module")
def api(gf):
    from gaugefields_jl_amd import api as a

    return a


class _FakeU:
    """stand-in configuration for validation paths that never reach the device"""

    class _L:
        precision = "f64"

    lattice = _L()
    NC = 3

    def __len__(self):
        return 4


# ------------------------------------------------------------------ CPU: validation


@pytest.mark.parametrize(
    "kwargs",
    [dict(colors=2), dict(start="warm"), dict(halo=-1), dict(boundary=(1, 1, 1, -1)),
     dict(eltype=np.float64)],
)
def test_gauge_configuration_rejects(api, kwargs):
    # API.jl:178-251 argument checks
    with pytest.raises(api.ArgumentError):
        api.gauge_configuration(D4, **kwargs)


def test_gauge_configuration_rejects_shape(api):
    with pytest.raises(api.ArgumentError):
        api.gauge_configuration((4, 4, 4))
    with pytest.raises(api.ArgumentError):
        api.gauge_configuration((4, 4, 0, 4))


@pytest.mark.parametrize("kwargs", [dict(steps=0), dict(steps=-3), dict(steps=4, trajectory_length=0.0),
                                    dict(steps=4, trajectory_length=math.inf),
                                    dict(steps=4, trajectory_length=math.nan)])
def test_md_driver_validation(api, kwargs):
    # molecular_dynamics.jl:447-466
    U = _FakeU()
    action = api.GaugeAction(U)
    with pytest.raises(api.ArgumentError):
        api.md_driver(U, action, **kwargs)


def test_md_driver_float32_conversion(api):
    # molecular_dynamics.jl:459-465: trajectory_length that underflows in Float32
    U = _FakeU()
    U.lattice = type("L", (), {"precision": "f32"})()
    with pytest.raises(api.ArgumentError):
        api.md_driver(U, api.GaugeAction(U), steps=2, trajectory_length=1e-50)
    d = api.md_driver(U, api.GaugeAction(U), steps=3, trajectory_length=0.1)
    assert d.trajectory_length == float(np.float32(0.1))


def test_action_loop_set_restricted(api):
    U = _FakeU()
    action = api.GaugeAction(U)
    plaq = api.make_loops_fromname("plaquette")
    with pytest.raises(api.ArgumentError):
        action.push(1.0, plaq)  # not closed under conjugation
    action.push(2.75, plaq + api.dagger(plaq))
    assert action.coefficient == 2.75
    with pytest.raises(api.ArgumentError):
        api.make_loops_fromname("rectangular")


def test_unknown_action_provider(api):
    with pytest.raises(api.ArgumentError):
        api.md_action_workspace(object(), _FakeU())


# ------------------------------------------------------------------ GPU behaviour


def _action(api, U, coeff=3.0):
    a = api.GaugeAction(U)
    plaq = api.make_loops_fromname("plaquette")
    a.push(coeff, plaq + api.dagger(plaq))
    return a


@pytest.mark.gpu
def test_configurations_and_plaquette(api, oracle):
    U = api.gauge_configuration(D4)
    assert api.measure_plaquette(U) == 1.0
    assert api.calculate_Plaquette(U) == 6 * 256 * 3
    H1 = api.gauge_configuration(D4, start="hot", seed=99)
    H2 = api.gauge_configuration(D4, start="hot", seed=99)
    H3 = api.gauge_configuration(D4, start="hot", seed=98)
    assert api.measure_plaquette(H1) == api.measure_plaquette(H2) != api.measure_plaquette(H3)
    # the hot start is the device Philox stream of the seed, restated by the oracle
    ref = oracle.plaquette_sum(D4, oracle.hot_links(D4, 99))
    assert abs(api.calculate_Plaquette(H1) - ref) < 1e-12 * abs(ref) + 1e-11
    F32 = api.gauge_configuration(D4, start="hot", seed=99, eltype=np.complex64)
    assert abs(api.measure_plaquette(F32) - api.measure_plaquette(H1)) < 1e-6


@pytest.mark.gpu
def test_potential_force_and_derivative(api, oracle):
    U = api.gauge_configuration(D4, start="hot", seed=5)
    a = _action(api, U, 2.75)
    host = U.to_host()
    assert abs(api.md_potential(a, U) - oracle.potential(D4, host, 2.75)) < 1e-12 * abs(oracle.potential(D4, host, 2.75))
    F = api.md_force(None, a, U)
    ref = oracle.force(D4, host, 2.75)
    assert np.abs(F - ref).max() < 1e-12 * np.abs(ref).max()
    D = api.calc_dSdUmu(a, 2, U)
    Dref = gen.links_to_matrices(oracle.dsdu(D4, host, 2.75))[1]
    assert np.abs(D - Dref).max() < 1e-12
    with pytest.raises(api.ArgumentError):
        api.calc_dSdUmu(a, 5, U)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,precision", [(D4, "f64"), ((4, 2, 6, 4), "f64"), ((8, 8, 8, 8), "f32")])
def test_untraced_action(api, dims, precision):
    # evaluate_GaugeAction_untraced (GaugeActions.jl:144-213): the per-site loop sum against a
    # numpy np.roll restatement, and tr of it summed over sites = evaluate_GaugeAction
    print(f"[untraced action] {dims} {precision}", flush=True)
    U = api.gauge_configuration(dims, start="hot", seed=61)
    if precision == "f32":  # the links as float holds them (row 2 rebuilt in double)
        W = api.gauge_configuration(dims, eltype=np.complex64)
        W.from_host(U.to_host())
        U = W
    host = U.to_host()
    a = _action(api, U, 2.5)
    M = api.evaluate_GaugeAction_untraced(a, U)
    ref = lt.untraced_plaquette_action(host, dims, 2.5)
    assert np.abs(M - ref).max() < 1e-12 * np.abs(ref).max()
    tr = np.trace(M, axis1=1, axis2=2).sum()
    S = api.evaluate_GaugeAction(a, U)
    assert abs(tr.imag) < 1e-10 and tr.real == pytest.approx(S, rel=1e-12)
    # Hermitian by construction (the loop set is closed under conjugation)
    assert np.abs(M - np.conj(np.swapaxes(M, 1, 2))).max() < 1e-13 * np.abs(M).max()


@pytest.mark.gpu
def test_md_cold_start_and_custom_integrator(api):
    # md_driver.jl:258-282 cold start; :417-464 custom QPQ == built-in QPQ
    U = api.gauge_configuration(D4)
    P = api.gauge_momenta(U)
    d = api.md_driver(U, _action(api, U), steps=5)
    r = api.md_trajectory(U, P, d)
    assert r["delta_hamiltonian"] == 0.0 and api.measure_plaquette(U) == 1.0

    def my_qpq(U, P, eps, driver):
        api.update_gaugefields(U, P, 0.5 * eps, driver)
        api.update_momenta(P, U, eps, driver)
        api.update_gaugefields(U, P, 0.5 * eps, driver)

    Ua = api.gauge_configuration(D4, start="hot", seed=7)
    Ub = api.gauge_configuration(D4, start="hot", seed=7)
    Pa = api.gaussian_momenta(Ua, seed=8)
    Pb = api.gaussian_momenta(Ub, seed=8)
    ra = api.md_trajectory(Ua, Pa, api.md_driver(Ua, _action(api, Ua), steps=4, trajectory_length=0.2))
    rb = api.md_trajectory(Ub, Pb, api.md_driver(Ub, _action(api, Ub), steps=4, trajectory_length=0.2,
                                                 integrator=my_qpq))
    assert abs(ra["delta_hamiltonian"] - rb["delta_hamiltonian"]) < 2e-12 * abs(ra["initial_hamiltonian"])
    assert np.abs(Ua.to_host() - Ub.to_host()).max() < 2e-12
    with pytest.raises(api.ArgumentError):
        api.md_trajectory(Ua, Pb, api.md_driver(Ua, _action(api, Ua), steps=1))


@pytest.mark.gpu
@pytest.mark.parametrize("integ", ["QPQ", "PQP", "OMELYAN"])
def test_md_reversibility_api(api, integ):
    # md_driver.jl:328-368: forward then trajectory_length -> -tau, < 2e-12
    it = {"QPQ": api.QPQ(), "PQP": api.PQP(), "OMELYAN": api.OmelyanPQP()}[integ]
    U = api.gauge_configuration(D4, start="hot", seed=17)
    P = api.gaussian_momenta(U, seed=18)
    U0, P0 = U.to_host(), P.to_host()
    a = _action(api, U)
    api.md_trajectory(U, P, api.md_driver(U, a, steps=5, trajectory_length=0.2, integrator=it))
    api.md_trajectory(U, P, api.md_driver(U, a, steps=5, trajectory_length=-0.2, integrator=it))
    assert np.abs(U.to_host() - U0).max() < 2e-12 and np.abs(P.to_host() - P0).max() < 2e-12


@pytest.mark.gpu
def test_flow_and_stout_api(api, oracle):
    U = api.gauge_configuration(D4, start="hot", seed=31)
    host = U.to_host()
    g = api.gradient_flow(U, steps=10, step_size=0.01)
    E = api.flow(U, g, measure_every=5)
    Eref = oracle.flow(D4, host.copy(), 0.01, 10, 5)
    assert np.abs(E - Eref).max() < 1e-12 * np.abs(Eref).max()
    with pytest.raises(api.ArgumentError):
        api.gradient_flow(U, steps=0)

    V = api.gauge_configuration(D4, start="hot", seed=32)
    vh = V.to_host()
    sm = api.stout_smearing(V, rho=0.1, levels=3)
    assert np.abs(api.smear(V, sm) - oracle.stout_smear(D4, vh, 3, 0.1)).max() < 1e-13
    Dout = gen.hot_start_links(D4, 33)
    ref = oracle.stout_backward(D4, oracle.stout_smear(D4, vh, 2, 0.1), 0.1, Dout)
    sm2 = api.stout_smearing(V, rho=0.1, levels=1)
    # one-level back_prop of the configuration smeared twice by the oracle
    W = api.gauge_configuration(D4)
    W.from_host(oracle.stout_smear(D4, vh, 2, 0.1))
    assert np.abs(api.back_prop(Dout, sm2, W) - ref).max() < 1e-13
    act = api.SmearedGaugeAction(_action(api, V), sm)
    assert abs(api.md_potential(act, V) - oracle.stout_potential(D4, vh, 3, 0.1, 3.0)) < 1e-11
    P = api.gaussian_momenta(V, seed=34)
    r = api.md_trajectory(V, P, api.md_driver(V, act, steps=4, trajectory_length=0.1,
                                             integrator=api.OmelyanPQP()))
    assert abs(r["delta_hamiltonian"]) < 1.0
