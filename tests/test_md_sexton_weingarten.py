"""Multiple-time-scale MD: MDActionSet, MDForceGroup and the Sexton-Weingarten integrator.

Mirrors the reference's md_driver.jl testsets for these types (validation :150-196, force-call
counts and reversibility :197-255) and checks the schedule against the oracle: the same sequence
of kicks and link updates (molecular_dynamics.jl:618-700) stepped on the oracle's force, kick and
exponential, with a stout-smeared slow member and a thin-link fast member.
"""
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


def _action(api, U, coeff):
    a = api.GaugeAction(U)
    plaq = api.make_loops_fromname("plaquette")
    a.push(coeff, plaq + api.dagger(plaq))
    return a


# ------------------------------------------------------------------ CPU: validation


def test_action_set_and_force_group_validation(api):
    U = _FakeU()
    with pytest.raises(api.ArgumentError, match="at least one action provider"):
        api.MDActionSet()
    with pytest.raises(api.ArgumentError, match="at least one action name"):
        api.MDForceGroup()
    with pytest.raises(api.ArgumentError, match="duplicate action names"):
        api.MDForceGroup("a", "a")
    s = api.MDActionSet(first=_action(api, U, 1.0), second=_action(api, U, 2.0))
    assert s.keys() == ("first", "second")
    assert api.MDForceGroup(("first", "second")).names == ("first", "second")


@pytest.mark.parametrize(
    "slow,fast,msg",
    [("first", "first", "overlap"), ("missing", "first", "unknown actions"),
     ("second", ("first",), None)],
)
def test_sexton_weingarten_driver_validation(api, slow, fast, msg):
    # md_driver.jl:150-196: overlapping groups, unknown names; a valid split is accepted
    U = _FakeU()
    s = api.MDActionSet(first=_action(api, U, 1.0), second=_action(api, U, 2.0))
    integ = api.SextonWeingarten(slow=slow, fast=fast, n_fast=1)
    if msg is None:
        d = api.md_driver(U, s, steps=1, integrator=integ)
        assert d.integrator.slow.names == ("second",) and d.integrator.fast.names == ("first",)
    else:
        with pytest.raises(api.ArgumentError, match=msg):
            api.md_driver(U, s, steps=1, integrator=integ)


def test_sexton_weingarten_rejects(api):
    U = _FakeU()
    three = api.MDActionSet(a=_action(api, U, 1.0), b=_action(api, U, 1.0), c=_action(api, U, 1.0))
    with pytest.raises(api.ArgumentError, match="does not schedule"):
        api.md_driver(U, three, steps=1, integrator=api.SextonWeingarten(slow="a", fast="b", n_fast=2))
    with pytest.raises(api.ArgumentError, match="requires an MDActionSet"):
        api.md_driver(U, _action(api, U, 1.0), steps=1,
                      integrator=api.SextonWeingarten(slow="a", fast="b", n_fast=2))
    with pytest.raises(api.ArgumentError, match="n_fast must be positive"):
        api.SextonWeingarten(slow="a", fast="b", n_fast=0)
    with pytest.raises(api.ArgumentError, match="ordering must be PQP"):
        api.SextonWeingarten(slow="a", fast="b", n_fast=1, ordering=api.OmelyanPQP())
    with pytest.raises(api.ArgumentError, match="selects members of an MDActionSet"):
        api.md_force(None, _action(api, U, 1.0), U, group=api.MDForceGroup("a"))


# ------------------------------------------------------------------ GPU


def _oracle_sw(oracle, dims, U, P, members, slow, fast, n_fast, ordering, steps, tau):
    """the Sexton-Weingarten schedule (molecular_dynamics.jl:618-700) on the oracle's primitives;
    members: name -> force(U) callable"""
    def kick(names, eps):
        for n in names:
            oracle.kick(dims, P, members[n](U), eps)

    def fast_qpq(duration):
        h = duration / n_fast
        oracle.exp_update(dims, U, P, h / 2)
        for k in range(1, n_fast + 1):
            kick(fast, h)
            oracle.exp_update(dims, U, P, h / 2 if k == n_fast else h)

    eps = tau / steps
    for _ in range(steps):
        if ordering == "QPQ":
            fast_qpq(eps / 2)
            kick(slow, eps)
            fast_qpq(eps / 2)
        else:
            kick(slow, eps / 2)
            fast_qpq(eps)
            kick(slow, eps / 2)


@pytest.mark.gpu
@pytest.mark.parametrize("ordering", ["QPQ", "PQP"])
def test_sexton_weingarten_counts_and_reversibility(api, ordering):
    # md_driver.jl:197-255: slow coupling 0.75, fast 0.25, n_fast 3, one step of length 0.1:
    # QPQ ordering 1 slow and 6 fast force evaluations, PQP 2 and 3; forward then back < 3e-12
    print(f"[sexton-weingarten] counts {ordering}", flush=True)
    U = api.gauge_configuration(D4, start="hot", seed=0x1234)
    P = api.gaussian_momenta(U, seed=0x5678)
    U0, P0 = U.to_host(), P.to_host()
    slow, fast = _action(api, U, 0.75), _action(api, U, 0.25)
    s = api.MDActionSet(slow=slow, fast=fast)
    o = api.QPQ() if ordering == "QPQ" else api.PQP()
    integ = api.SextonWeingarten(slow="slow", fast="fast", n_fast=3, ordering=o)
    calls = {0.75: 0, 0.25: 0}
    lat = U.lattice
    orig = lat.force_kick

    def counting(coeff, eps):
        calls[coeff] += 1
        return orig(coeff, eps)

    lat.force_kick = counting
    try:
        api.md_trajectory(U, P, api.md_driver(U, s, steps=1, trajectory_length=0.1, integrator=integ),
                          diagnostics=False)
    finally:
        lat.force_kick = orig
    assert (calls[0.75], calls[0.25]) == ((1, 6) if ordering == "QPQ" else (2, 3))
    api.md_trajectory(U, P, api.md_driver(U, s, steps=1, trajectory_length=-0.1, integrator=integ),
                      diagnostics=False)
    assert np.abs(U.to_host() - U0).max() < 3e-12
    assert np.abs(P.to_host() - P0).max() < 3e-12


@pytest.mark.gpu
@pytest.mark.parametrize("ordering", ["QPQ", "PQP"])
def test_sexton_weingarten_matches_oracle(api, oracle, ordering):
    # slow: a 2-level stout-smeared plaquette action (rho 0.1, coefficient 2.5); fast: the thin
    # plaquette action (coefficient 0.5); n_fast 2, 3 steps of tau 0.3
    print(f"[sexton-weingarten] oracle {ordering}", flush=True)
    U = api.gauge_configuration(D4, start="hot", seed=41)
    P = api.gaussian_momenta(U, seed=42)
    Uo, Po = U.to_host().copy(), np.ascontiguousarray(P.to_host().copy())
    sm = api.stout_smearing(U, rho=0.1, levels=2, mode="store")
    s = api.MDActionSet(smeared=api.SmearedGaugeAction(_action(api, U, 2.5), sm), thin=_action(api, U, 0.5))
    o = api.QPQ() if ordering == "QPQ" else api.PQP()
    integ = api.SextonWeingarten(slow="smeared", fast=("thin",), n_fast=2, ordering=o)
    r = api.md_trajectory(U, P, api.md_driver(U, s, steps=3, trajectory_length=0.3, integrator=integ))

    members = {"smeared": lambda W: oracle.stout_force(D4, W, 2, 0.1, 2.5),
               "thin": lambda W: oracle.force(D4, W, 0.5)}

    def H(W, Q):
        return (oracle.stout_potential(D4, W, 2, 0.1, 2.5) + oracle.potential(D4, W, 0.5)
                + 0.5 * oracle.momenta_dot(D4, Q))

    h0 = H(Uo, Po)
    _oracle_sw(oracle, D4, Uo, Po, members, ["smeared"], ["thin"], 2, ordering, 3, 0.3)
    h1 = H(Uo, Po)
    assert np.abs(U.to_host() - Uo).max() < 1e-11
    assert np.abs(P.to_host() - Po).max() < 1e-11
    assert r["initial_hamiltonian"] == pytest.approx(h0, rel=1e-12)
    assert abs(r["delta_hamiltonian"] - (h1 - h0)) < 1e-12 * abs(h0)


@pytest.mark.gpu
def test_action_set_force_is_the_sum(api, oracle):
    # md_force! of an MDActionSet (molecular_dynamics.jl:128-236): the members' forces added,
    # all of them or those of a group
    U = api.gauge_configuration(D4, start="hot", seed=51)
    host = U.to_host()
    sm = api.stout_smearing(U, rho=0.12, levels=3, mode="store")
    s = api.MDActionSet(a=_action(api, U, 1.5), b=api.SmearedGaugeAction(_action(api, U, 2.0), sm),
                        c=_action(api, U, -0.5))
    Fa, Fb, Fc = oracle.force(D4, host, 1.5), oracle.stout_force(D4, host, 3, 0.12, 2.0), oracle.force(D4, host, -0.5)
    F = api.md_force(None, s, U)
    assert np.abs(F - (Fa + Fb + Fc)).max() < 1e-12 * np.abs(Fa + Fb).max()
    G = api.md_force(None, s, U, group=api.MDForceGroup("c", "b"))
    assert np.abs(G - (Fc + Fb)).max() < 1e-12 * np.abs(Fb).max()
    pot = api.md_potential(s, U)
    ref = oracle.potential(D4, host, 1.5) + oracle.stout_potential(D4, host, 3, 0.12, 2.0) + oracle.potential(D4, host, -0.5)
    assert pot == pytest.approx(ref, rel=1e-12)
