"""Write the golden vectors of the SU(3) gauge-link path: tests/golden/vectors_4x4x4x4.npz.

Inputs come from the package's seeded numpy generators (Philox bit streams) and the
outputs from the CPU oracle (oracle/gf_oracle.c), once, so that the CPU tests can check the
oracle against its own earlier output and every other check it carries (numpy brute force,
scipy expm, finite differences), and the GPU tests can compare libgfhip against fixed data
without recomputing the reference side.

    python tests/golden/make_goldens.py          # rewrite the fixture
    python tests/golden/make_goldens.py --check  # recompute and compare, write nothing

Contents (4x4x4x4 lattice, site s = x + 4(y + 4(z + 4t)); links complex (4, V, 9) row-major
3x3, momenta float (4, V, 8) coefficients of sum_a p_a i lambda_a / 2):
  U0        links exp(i sum_a w_a lambda_a/2), w_a ~ N(0, 0.3^2), seed 101
  P0        momenta N(0, 1), seed 102
  plaq      sum_{x, mu<nu} Re tr P_munu(x) of U0        (AbstractGaugefields.jl:2684-2699)
  potential MD potential -(2 coeff/3) plaq at coeff 3 (beta = 6)
                                                         (molecular_dynamics.jl:247-249)
  dsdu      dS/dU_mu at coeff 3                          (GaugeActions.jl:95-123)
  force     -(1/3) TA(U_mu dS/dU_mu)                     (molecular_dynamics.jl:251-267)
  P1, U1    P0 + 0.02 force, then exp(0.02 P1) U0        (molecular_dynamics.jl:513-551)
  H_qpq     (H0, H1, dH) of one QPQ trajectory, 4 steps, tau 0.2, from (U0, P0)
  Uqpq      links at the end of that trajectory          (molecular_dynamics.jl:604-730)
  Ustout    one stout level rho 0.12 of U0               (stout_fast.jl:250-274)
  Dout, Din a link-shaped cotangent (hot links, seed 103) and its pullback through that level
                                                         (stout_fast.jl:317-407, 712-785)
  W, E_flow U0 after 5 RK3 flow steps eps 0.02, E(t) after steps 1..5 (gradientflow.jl:171-238)
  Q         topological charge of U0, loop sets plaquette, clover, improved, rectangle
  Eclov     clover energy density of U0                  (samples/measurements/energydensity.jl)
  Uhb       U0 after one heatbath and one overrelaxation sweep, beta 6, seed 104
  Prng      device-stream Gaussian momenta, seed 105, sweep 3
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

DIMS = (4, 4, 4, 4)
COEFF = 3.0
PATH = os.path.join(HERE, "vectors_4x4x4x4.npz")
TOPO = ("plaquette", "clover", "improved", "rectangle")


def compute():
    import __graft_entry__
    import oracle.gforacle as o

    gen = __graft_entry__.load_package().generators
    o.lib()
    d = DIMS
    out = {"dims": np.array(d), "coeff": np.array(COEFF)}
    U0 = gen.cold_links(d)
    o.exp_update(d, U0, gen.algebra_coefficients(d, 0.3, 101), 1.0)
    P0 = gen.gaussian_momenta(d, 102)
    out["U0"], out["P0"] = U0.copy(), P0.copy()
    out["plaq"] = np.array(o.plaquette_sum(d, U0))
    out["potential"] = np.array(o.potential(d, U0, COEFF))
    out["dsdu"] = o.dsdu(d, U0, COEFF)
    F = o.force(d, U0, COEFF)
    out["force"] = F
    P1, U1 = P0.copy(), U0.copy()
    o.kick(d, P1, F, 0.02)
    o.exp_update(d, U1, P1, 0.02)
    out["P1"], out["U1"] = P1, U1
    Ut, Pt = U0.copy(), P0.copy()
    out["H_qpq"] = np.array(o.md_trajectory(d, Ut, Pt, COEFF, 4, 0.2, "QPQ"))
    out["Uqpq"] = Ut
    out["Ustout"] = o.stout_forward(d, U0, 0.12)
    Dout = gen.hot_start_links(d, 103)
    out["Dout"] = Dout
    out["Din"] = o.stout_backward(d, U0, 0.12, Dout)
    W = U0.copy()
    out["E_flow"] = o.flow(d, W, 0.02, 5, 1)
    out["W"] = W
    out["Q"] = np.array([o.topological_charge_density(d, U0, m)[0] for m in TOPO])
    out["Eclov"] = np.array(o.energy_density_clover(d, U0))
    Uh = U0.copy()
    o.heatbath_sweep(d, Uh, 6.0, 104, 0)
    o.heatbath_sweep(d, Uh, 6.0, 104, 1, overrelax=True)
    out["Uhb"] = Uh
    out["Prng"] = o.gaussian_momenta(d, 105, 3)
    return out


def load():
    with np.load(PATH, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true")
    a = ap.parse_args()
    new = compute()
    if a.check:
        old = load()
        worst = max(float(np.abs(new[k] - old[k]).max()) for k in new)
        print(f"max |new - stored| = {worst:.3e}")
        sys.exit(0 if worst < 1e-12 else 1)
    np.savez(PATH, **new)
    print(f"wrote {PATH} ({os.path.getsize(PATH)} bytes)")


if __name__ == "__main__":
    main()
