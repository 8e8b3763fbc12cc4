"""Seeded synthetic inputs for the SU(3) hot path (numpy, host side).

The reference draws its hot starts and momenta from LatticeMatrices' site-local
Philox/PCG/Xoshiro streams (gaugefields_4D_MPILattice.jl:430-472,
TA_gaugefields_4D_MPILattice.jl:157-193), which are not reproducible outside Julia.
The workloads here therefore define their own generators, driven by numpy's Philox
bit generator with a committed seed, and feed the same arrays to the HIP library and
to the CPU oracle:

* ``hot_start_links``: each link a complex-Gaussian 3x3 matrix made special unitary by
  the row Gram-Schmidt + cross-product completion of normalize3!/m3complv3!
  (AbstractGaugefields.jl:3378-3457).
* ``algebra_coefficients``: omega_a ~ N(0, sigma^2) per (direction, site, generator); the
  links U = exp(i sum_a omega_a T_a) are then formed by the exp-update kernel acting on
  unit links (U <- exp(1 * P) U with P = omega).
* ``gaussian_momenta``: p_a ~ N(0, 1) per generator (gauss_distribution!,
  TA_gaugefields_4D_serial.jl:49-73).

Array layouts are the reference's Julia layouts: links (4, V, 9) complex128 with
element (i, j) at i + 3*j; momenta (4, V, 8) float64; site s = x + NX*(y + NY*(z + NZ*t)).
"""
import numpy as np

__all__ = [
    "volume",
    "cold_links",
    "hot_start_links",
    "algebra_coefficients",
    "gaussian_momenta",
    "links_to_matrices",
    "matrices_to_links",
    "LAMBDA",
]


def volume(dims):
    dims = tuple(int(d) for d in dims)
    if len(dims) != 4 or any(d <= 0 for d in dims):
        raise ValueError(f"lattice must have four positive extents; got {dims}")
    return dims[0] * dims[1] * dims[2] * dims[3]


def _rng(seed):
    return np.random.Generator(np.random.Philox(int(seed)))


def links_to_matrices(U):
    """(4, V, 9) reference layout -> (4, V, 3, 3) with [..., i, j]."""
    U = np.asarray(U)
    return U.reshape(U.shape[0], U.shape[1], 3, 3).transpose(0, 1, 3, 2)


def matrices_to_links(M):
    """(4, V, 3, 3) with [..., i, j] -> (4, V, 9) reference layout."""
    M = np.asarray(M)
    return np.ascontiguousarray(M.transpose(0, 1, 3, 2)).reshape(M.shape[0], M.shape[1], 9)


def cold_links(dims):
    V = volume(dims)
    M = np.zeros((4, V, 3, 3), np.complex128)
    M[..., 0, 0] = M[..., 1, 1] = M[..., 2, 2] = 1.0
    return matrices_to_links(M)


def _special_unitary_rows(M):
    """normalize3! (AbstractGaugefields.jl:3378-3428) + m3complv3! (:3431-3457), vectorized:
    orthonormalize rows 1 and 2, row 3 = conj(row1 x row2)."""
    r1 = M[..., 0, :]
    r2 = M[..., 1, :]
    w2 = np.sum(r1 * np.conj(r1), axis=-1).real
    if np.any(w2 == 0):
        raise ValueError("zero first row in SU(3) normalization")
    w1 = -np.sum(r2 * np.conj(r1), axis=-1) / w2
    x = r2 + w1[..., None] * r1
    w3 = np.sum(x * np.conj(x), axis=-1).real
    if np.any(w3 == 0):
        raise ValueError("dependent rows in SU(3) normalization")
    r1 = r1 / np.sqrt(w2)[..., None]
    r2 = x / np.sqrt(w3)[..., None]
    r3 = np.conj(np.cross(r1, r2))
    return np.stack([r1, r2, r3], axis=-2)


def hot_start_links(dims, seed):
    """Complex-Gaussian links, Gram-Schmidt orthonormalized and det-fixed to SU(3)."""
    V = volume(dims)
    g = _rng(seed)
    M = g.standard_normal((4, V, 3, 3)) + 1j * g.standard_normal((4, V, 3, 3))
    return matrices_to_links(_special_unitary_rows(M))


def algebra_coefficients(dims, sigma, seed):
    """omega_a ~ N(0, sigma^2), shape (4, V, 8); U = exp(i sum_a omega_a lambda_a / 2)."""
    V = volume(dims)
    return float(sigma) * _rng(seed).standard_normal((4, V, 8))


def gaussian_momenta(dims, seed, sigma=1.0):
    V = volume(dims)
    return float(sigma) * _rng(seed).standard_normal((4, V, 8))


def _gell_mann():
    lam = np.zeros((8, 3, 3), np.complex128)
    lam[0][0, 1] = lam[0][1, 0] = 1
    lam[1][0, 1] = -1j
    lam[1][1, 0] = 1j
    lam[2][0, 0] = 1
    lam[2][1, 1] = -1
    lam[3][0, 2] = lam[3][2, 0] = 1
    lam[4][0, 2] = -1j
    lam[4][2, 0] = 1j
    lam[5][1, 2] = lam[5][2, 1] = 1
    lam[6][1, 2] = -1j
    lam[6][2, 1] = 1j
    lam[7] = np.diag([1.0, 1.0, -2.0]) / np.sqrt(3.0)
    return lam


#: Gell-Mann matrices lambda_1..lambda_8 in the order of TA_gaugefields_4D_serial.jl:243-260
LAMBDA = _gell_mann()
