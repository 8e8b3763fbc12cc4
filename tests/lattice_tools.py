"""Independent numpy restatements used by the tests as a second check on the oracle.

Whole-lattice np.roll formulations (not the oracle's site loops), the reference's
instanton constructor, gauge transformations and finite differences.
"""
import os
import sys

import numpy as np
from scipy.linalg import expm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

gen = __graft_entry__.load_package().generators
LAMBDA = gen.LAMBDA
T_GEN = LAMBDA / 2


def field(U, dims):
    """(4, V, 9) reference layout -> (4, NT, NZ, NY, NX, 3, 3)."""
    NX, NY, NZ, NT = dims
    return gen.links_to_matrices(U).reshape(4, NT, NZ, NY, NX, 3, 3)


def unfield(M):
    return gen.matrices_to_links(M.reshape(M.shape[0], -1, 3, 3))


_AX = {0: 3, 1: 2, 2: 1, 3: 0}


def sh(A, mu, d=1):
    """A(x + d*mu) for A of shape (NT, NZ, NY, NX, ...)."""
    return np.roll(A, -d, axis=_AX[mu])


def dg(A):
    return np.conj(np.swapaxes(A, -1, -2))


def plaquette_field_sum(U, dims):
    """sum_{x, mu<nu} Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag (any NC)."""
    M = field(U, dims) if U.ndim == 3 else U
    tot = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            P = M[mu] @ sh(M[nu], mu) @ dg(sh(M[mu], nu)) @ dg(M[nu])
            tot += np.trace(P, axis1=-2, axis2=-1).real.sum()
    return tot


def untraced_plaquette_action(U, dims, coeff):
    """coeff sum_{mu<nu} (P_munu(x) + P_munu(x)^dag) per site as (V, 3, 3): the untraced loop sum
    of the loop set plaquette u plaquette^dag (every loop evaluated from its starting site x)"""
    M = field(U, dims)
    out = 0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            P = M[mu] @ sh(M[nu], mu) @ dg(sh(M[mu], nu)) @ dg(M[nu])
            out = out + P + dg(P)
    return coeff * out.reshape(-1, 3, 3)


def ta_matrix(M):
    A = 0.5 * (M - dg(M))
    tr = np.trace(A, axis1=-2, axis2=-1)
    return A - (tr / 3.0)[..., None, None] * np.eye(3)


def coeffs_to_matrix(c):
    """sum_a c_a i lambda_a / 2 for c of shape (..., 8)."""
    return np.einsum("...a,aij->...ij", c, 1j * T_GEN)


def su2_instanton_link(mu, ix, iy, iz, it, L):
    """_su2_instanton_link (AbstractGaugefields.jl:1126-1170), default centre/radius, sign +1."""
    center = [L[k] / 2 + 0.5 for k in range(4)]
    radius = L[0] // 2
    s1 = np.array([[0, 1], [1, 0]], complex)
    s2 = np.array([[0, -1j], [1j, 0]], complex)
    s3 = np.array([[1, 0], [0, -1]], complex)
    En = np.eye(2, dtype=complex)
    ss = (1j * s1, 1j * s2, 1j * s3, En)
    sd = (-1j * s1, -1j * s2, -1j * s3, En)
    nv = np.array([ix - 1 - center[0], iy - 1 - center[1], iz - 1 - center[2], it - 1 - center[3]])
    n2 = float(nv @ nv)
    tau = np.zeros((2, 2), complex)
    for nu in range(4):
        tau += (sd[mu] @ ss[nu] - sd[nu] @ ss[mu]) * nv[nu]
    return expm(1j * tau * 0.5 * (1.0 / n2) * (1j * radius ** 2 / (n2 + radius ** 2)))


def instanton_links(dims, nc=3, block=(0, 1)):
    """Oneinstanton_4D_wing (gaugefields_4D_wing.jl:388-431), embedded in SU(nc) at `block`."""
    NX, NY, NZ, NT = dims
    V = NX * NY * NZ * NT
    M = np.zeros((4, V, nc, nc), complex)
    M[..., range(nc), range(nc)] = 1.0
    i, j = block
    for it in range(1, NT + 1):
        for iz in range(1, NZ + 1):
            for iy in range(1, NY + 1):
                for ix in range(1, NX + 1):
                    s = (ix - 1) + NX * ((iy - 1) + NY * ((iz - 1) + NZ * (it - 1)))
                    for mu in range(4):
                        u2 = su2_instanton_link(mu, ix, iy, iz, it, dims)
                        M[mu, s, i, i] = u2[0, 0]
                        M[mu, s, i, j] = u2[0, 1]
                        M[mu, s, j, i] = u2[1, 0]
                        M[mu, s, j, j] = u2[1, 1]
    return M


def random_su3(n, rng):
    g = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    q, r = np.linalg.qr(g)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    return q / (det ** (1.0 / 3.0))[:, None, None]


def gauge_transform(U, dims, g):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^dag; g has shape (V, 3, 3)."""
    NX, NY, NZ, NT = dims
    M = field(U, dims)
    G = g.reshape(NT, NZ, NY, NX, 3, 3)
    out = np.empty_like(M)
    for mu in range(4):
        out[mu] = G @ M[mu] @ dg(sh(G, mu))
    return unfield(out)


def perturb_link(U, mu, s, a, h):
    """copy of U with U_mu(s) -> exp(h i T_a) U_mu(s)."""
    V = U.copy()
    M = gen.links_to_matrices(V)
    E = expm(h * 1j * T_GEN[a])
    M[mu, s] = E @ M[mu, s]
    return gen.matrices_to_links(M)


def stout_force_fd_local(oracle, dims, U, nlev, rho, coeff, mu, s, a, h):
    """Central finite difference -(V(+h) - V(-h)) / 2h of the stout-smeared Wilson potential
    V = -(2 coeff / 3) sum Re tr P(smeared) under U_mu(s) -> exp(+-h i T_a) U_mu(s), evaluated
    locally so that its rounding stays far below the O(h^2) truncation even on large lattices.

    Changing one thin link changes the level-l smeared links only within max-norm distance l of
    its site, hence the plaquettes only at base sites within nlev + 1.  The smeared links there
    are exact on a periodic window of the lattice covering offsets -(nlev+4) .. nlev+5 around the
    site (the wrap-around at the window's faces corrupts one layer per level from outside in), so
    each evaluation smears that window with the oracle and sums the per-plaquette differences
    tr P(+h) - tr P(-h) over the base sites within nlev + 1 (all other differences are exactly
    zero).  Axes not longer than the window are taken whole (exactly periodic)."""
    NX, NY, NZ, NT = dims
    x = [s % NX, (s // NX) % NY, (s // (NX * NY)) % NZ, s // (NX * NY * NZ)]
    lo, hi = nlev + 4, nlev + 5
    W = field(U, dims)
    wdims, cen, box = [], [], []
    for d in range(4):
        n = dims[d]
        if n <= lo + hi + 1:  # the whole axis, rolled so that the site sits at index n // 2
            c = n // 2
            idx = (x[d] + np.arange(n) - c) % n
            box.append((0, n))
        else:
            c = lo
            idx = (x[d] + np.arange(-lo, hi + 1)) % n
            box.append((c - nlev - 1, c + nlev + 2))
        W = np.take(W, idx, axis=_AX[d] + 1)
        wdims.append(len(idx))
        cen.append(c)
    wdims = tuple(wdims)
    site = (cen[3], cen[2], cen[1], cen[0])
    tr = []
    for sgn in (1, -1):
        Wp = W.copy()
        Wp[(mu,) + site] = expm(sgn * h * 1j * T_GEN[a]) @ W[(mu,) + site]
        S = field(oracle.stout_smear(wdims, np.ascontiguousarray(unfield(Wp)), nlev, rho), wdims)
        t = 0.0
        for m in range(4):
            for n_ in range(m + 1, 4):
                P = S[m] @ sh(S[n_], m) @ dg(sh(S[m], n_)) @ dg(S[n_])
                sl = tuple(slice(*box[d]) for d in (3, 2, 1, 0))
                t = t + np.trace(P[sl], axis1=-2, axis2=-1).real
        tr.append(t)
    dV = -(2.0 * coeff / 3.0) * float(np.sum(tr[0] - tr[1]))
    return -dV / (2 * h)


def config1_chain(oracle, run_gpu=None, n_therm=10, n_traj=20, coeff=3.0):
    """The plain-HMC workload on 4^4 (beta 6, QPQ 10 steps, tau 1) from the Gram-Schmidt hot
    start of seed 101: n_therm trajectories accepted unconditionally, then n_traj with the
    Metropolis test (uniforms from Philox stream 103); momenta from Philox stream 102.  The
    oracle runs the chain; run_gpu(U0), if given, receives the start links and returns a callback
    step(P, r, metropolis, record) that the GPU side runs after each oracle trajectory.
    Returns the oracle's per-trajectory records (H, accepted, plaquette sum after it)."""
    dims = (4, 4, 4, 4)
    Uo = gen.hot_start_links(dims, 101)
    rng_mom = np.random.Generator(np.random.Philox(102))
    rng_acc = np.random.Generator(np.random.Philox(103))
    step = run_gpu(Uo.copy()) if run_gpu else None
    recs = []
    for traj in range(n_therm + n_traj):
        P = rng_mom.standard_normal((4, 256, 8))
        metropolis = traj >= n_therm
        r = rng_acc.random() if metropolis else 0.0
        Ut, Pt = Uo.copy(), P.copy()
        Href = oracle.md_trajectory(dims, Ut, Pt, coeff, 10, 1.0, "QPQ")
        acc = bool((not metropolis) or np.exp(-Href[2]) >= r)
        if acc:
            Uo = Ut
        rec = {"traj": traj, "H": Href, "accepted": acc, "metropolis": metropolis,
               "plaquette_sum": oracle.plaquette_sum(dims, Uo)}
        if step:
            step(P, r, metropolis, rec)
        recs.append(rec)
    return recs


# ---------------------------------------------------------------- field strengths


def path_field(M, path):
    """Whole-lattice product of links along path [(mu, n), ...] from every site x:
    (mu, +n) takes U_mu(y) and moves y -> y + mu, n times; (mu, -n) moves y -> y - mu and
    takes U_mu(y)^dag, n times.  M: (4, NT, NZ, NY, NX, 3, 3)."""
    disp = [0, 0, 0, 0]
    out = None
    for mu, n in path:
        for _ in range(abs(n)):
            if n > 0:
                A = M[mu]
                for d in range(4):
                    A = sh(A, d, disp[d]) if disp[d] else A
                disp[mu] += 1
            else:
                disp[mu] -= 1
                A = M[mu]
                for d in range(4):
                    A = sh(A, d, disp[d]) if disp[d] else A
                A = dg(A)
            out = A if out is None else out @ A
    return out


def _loops(method, m, n):
    if method == "plaquette":
        return [[(m, 1), (n, 1), (m, -1), (n, -1)]]
    if method == "clover":
        return [[(m, 1), (n, 1), (m, -1), (n, -1)], [(n, 1), (m, -1), (n, -1), (m, 1)],
                [(n, -1), (m, 1), (n, 1), (m, -1)], [(m, -1), (n, -1), (m, 1), (n, 1)]]
    return [[(m, 2), (n, 1), (m, -2), (n, -1)], [(n, 1), (m, -2), (n, -1), (m, 2)],
            [(n, -1), (m, 2), (n, 1), (m, -2)], [(m, -2), (n, -1), (m, 2), (n, 1)],
            [(m, 1), (n, 2), (m, -1), (n, -2)], [(n, 2), (m, -1), (n, -2), (m, 1)],
            [(n, -2), (m, 1), (n, 2), (m, -1)], [(m, -1), (n, -2), (m, 1), (n, 2)]]


def field_strengths(U, dims, method):
    M = field(U, dims)
    F = {}
    for m in range(4):
        for n in range(4):
            if m != n:
                F[m, n] = ta_matrix(sum(path_field(M, p) for p in _loops(method, m, n)))
    return F


def _levi_civita(p):
    if len(set(p)) < 4:
        return 0
    inv = sum(1 for i in range(4) for j in range(i + 1, 4) if p[i] > p[j])
    return -1 if inv % 2 else 1


def topo_density_numpy(U, dims, method):
    """density in site order, literal epsilon sum over all index quadruples"""
    if method == "improved":
        return (5 / 3) * topo_density_numpy(U, dims, "clover") - (1 / 12) * topo_density_numpy(U, dims, "rectangle")
    nl = {"plaquette": 1, "clover": 4, "rectangle": 8}[method]
    norm = 2.0 if method == "rectangle" else 1.0
    F = field_strengths(U, dims, method)
    q = 0.0
    import itertools
    for p in itertools.permutations(range(4)):
        e = _levi_civita(p)
        q = q + e * np.einsum("...ij,...ji->...", F[p[0], p[1]], F[p[2], p[3]])
    return (-norm * q.real / (32 * np.pi ** 2 * nl ** 2)).reshape(-1)


def abelian_flux_links(dims, n01, n23):
    """Constant abelian field strength: every (0,1) plaquette equals diag(e^{i t}, e^{-i t}, 1)
    with t = 2 pi n01 / (NX NY), every (2,3) plaquette the same with t' = 2 pi n23 / (NZ NT),
    all other plaquettes 1 (twisted boundary links carry the flux quantum)."""
    NX, NY, NZ, NT = dims
    ph = np.zeros((4, NT, NZ, NY, NX))
    x = np.arange(NX)[None, None, None, :]
    y = np.arange(NY)[None, None, :, None]
    z = np.arange(NZ)[None, :, None, None]
    t = np.arange(NT)[:, None, None, None]
    th = 2 * np.pi * n01 / (NX * NY)
    tp = 2 * np.pi * n23 / (NZ * NT)
    ph[1] = th * x + 0 * (y + z + t)
    ph[0] = np.where(x == NX - 1, -2 * np.pi * n01 * y / NY, 0.0) + 0 * (z + t)
    ph[3] = tp * z + 0 * (x + y + t)
    ph[2] = np.where(z == NZ - 1, -2 * np.pi * n23 * t / NT, 0.0) + 0 * (x + y)
    M = np.zeros((4, NT, NZ, NY, NX, 3, 3), complex)
    M[..., 0, 0] = np.exp(1j * ph)
    M[..., 1, 1] = np.exp(-1j * ph)
    M[..., 2, 2] = 1.0
    return unfield(M), th, tp
