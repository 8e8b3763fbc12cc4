"""Truncation of the Cayley-Hamilton exponential series the device kernels use (su3.hpp
ch_coeffs_series / exp_coeffs): exp(iH) = f0 + f1 H + f2 H^2 for traceless Hermitian H, with
H^n = a_n + b_n H + c_n H^2, a_{n+1} = c0 c_n, b_{n+1} = a_n + c1 c_n, c_{n+1} = b_n
(c0 = det H, c1 = tr H^2 / 2) and f_j = sum_{n<N} i^n / n! (a_n, b_n, c_n)_j.

The term counts and the c1 limits of each tier are read from su3.hpp, and the truncated series is
checked against scipy.linalg.expm (and its c0/c1 derivatives against expm_frechet) on spectra at
each tier's limit: the largest eigenvalue modulus allowed is r with r^2 = 4 c1 / 3.
"""
import math
import os
import re

import numpy as np
import pytest
from scipy.linalg import expm, expm_frechet

HERE = os.path.dirname(os.path.abspath(__file__))
SU3 = os.path.join(HERE, "..", "gaugefields.jl_amd", "csrc", "su3.hpp")


def _tiers():
    src = open(SU3).read()
    c1 = {m.group(1): float(m.group(2)) for m in re.finditer(r"#define (GF_CH_\w+_C1) ([0-9.]+)", src)}
    body = src[src.index("__device__ __forceinline__ void exp_coeffs("):]
    body = body[:body.index("\n}\n")]
    short = [int(n) for n in re.findall(r"ch_coeffs_series<false, (\d+)>", body)]
    full = re.search(r"int N = WANT_B \? (\d+) : (\d+)>", src)
    return [(c1["GF_CH_SHORT1_C1"], short[0]), (c1["GF_CH_SHORT2_C1"], short[1]),
            (c1["GF_CH_SERIES_C1"], int(full.group(2)))], (c1["GF_CH_SERIES_C1"], int(full.group(1)))


def _series(H, N, want_b=False):
    c0 = np.linalg.det(H).real
    c1 = 0.5 * np.trace(H @ H).real
    a, b, c = 1.0, 0.0, 0.0
    da = np.zeros(2); db = np.zeros(2); dc = np.zeros(2)  # d/dc0, d/dc1
    f = np.zeros(3, complex); g = np.zeros((2, 3), complex)
    for n in range(N):
        w = 1j ** n / math.factorial(n)
        f += w * np.array([a, b, c])
        g += w * np.stack([da, db, dc], axis=1)
        na, nb = c0 * c, c1 * c + a
        nda = np.array([c + c0 * dc[0], c0 * dc[1]])
        ndb = np.array([da[0] + c1 * dc[0], da[1] + c + c1 * dc[1]])
        dc, da, db = db, nda, ndb
        a, b, c = na, nb, b
    E = f[0] * np.eye(3) + f[1] * H + f[2] * H @ H
    return (E, g, c0, c1) if want_b else E


def _herm_with_spectrum(lams, rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
    return q @ np.diag(lams) @ q.conj().T


def _spectra(c1):
    # traceless spectra with tr H^2 / 2 = c1: (2, -1, -1) (largest modulus, sqrt(4 c1 / 3)),
    # (1, -1, 0), (-2, 1, 1) and (3, -1, -2), each scaled to c1
    out = []
    for v in ([2.0, -1.0, -1.0], [1.0, -1.0, 0.0], [-2.0, 1.0, 1.0], [3.0, -1.0, -2.0]):
        v = np.array(v)
        out.append(v * math.sqrt(c1 / (0.5 * (v @ v))))
    return out


def test_tier_constants():
    tiers, deriv = _tiers()
    assert [t[1] for t in tiers] == [11, 13, 18] and deriv == (0.75, 20)
    assert tiers[0][0] < tiers[1][0] < tiers[2][0]


@pytest.mark.parametrize("tier", [0, 1, 2])
def test_exp_series_tier_at_its_limit(tier):
    (c1max, N) = _tiers()[0][tier]
    rng = np.random.default_rng(7 + tier)
    short = []
    for lams in _spectra(c1max):
        H = _herm_with_spectrum(lams, rng)
        assert 0.5 * np.trace(H @ H).real <= c1max * (1 + 1e-12)
        assert np.abs(lams).max() <= math.sqrt(4.0 * c1max / 3.0) * (1 + 1e-12)
        err = np.abs(_series(H, N) - expm(1j * H)).max()
        assert err < 1e-15, (tier, lams, err)
        short.append(np.abs(_series(H, N - 2) - expm(1j * H)).max())
    # two terms fewer would not do at the limit (one term of margin over 1e-15)
    assert max(short) > 1e-15


def test_exp_series_small_spectra_all_tiers():
    # inside the smallest tier every count agrees with expm to rounding
    rng = np.random.default_rng(3)
    for c1 in (1e-16, 1e-8, 2e-3):
        for lams in _spectra(c1):
            H = _herm_with_spectrum(lams, rng)
            for _, N in _tiers()[0]:
                assert np.abs(_series(H, N) - expm(1j * H)).max() < 1e-15


@pytest.mark.parametrize("c1", [0.0075, 0.2, 0.75])
def test_derivative_series_vs_expm_frechet(c1):
    # d exp(iH)/dH along a traceless Hermitian direction from the series' c0/c1 derivatives:
    # d exp(iH) = sum_j [df_j H^j + f_j d(H^j)], df_j = b1_j dc1 + b2_j dc0
    _, (c1max, N) = _tiers()
    rng = np.random.default_rng(11)
    for lams in _spectra(c1):
        H = _herm_with_spectrum(lams, rng)
        D = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        D = 0.5 * (D + D.conj().T)
        D -= np.trace(D) / 3 * np.eye(3)
        E, g, c0, c1 = _series(H, N, want_b=True)
        # f_j themselves at N terms
        a = _series(H, N)
        dc1 = np.trace(H @ D).real
        dc0 = (np.linalg.det(H) * np.trace(np.linalg.solve(H, D)) if abs(np.linalg.det(H)) > 1e-14
               else _ddet(H, D)).real
        f = _fvec(H, N)
        dE = sum((g[1, j] * dc1 + g[0, j] * dc0) * np.linalg.matrix_power(H, j) for j in range(3))
        dE = dE + f[1] * D + f[2] * (D @ H + H @ D)
        ref = expm_frechet(1j * H, 1j * D, compute_expm=False)
        assert np.abs(dE - ref).max() < 2e-15 * max(1.0, np.abs(ref).max())


def _fvec(H, N):
    c0 = np.linalg.det(H).real
    c1 = 0.5 * np.trace(H @ H).real
    a, b, c = 1.0, 0.0, 0.0
    f = np.zeros(3, complex)
    for n in range(N):
        f += 1j ** n / math.factorial(n) * np.array([a, b, c])
        a, b, c = c0 * c, c1 * c + a, b
    return f


def _ddet(H, D):
    # derivative of det along D: sum over columns replaced by D's
    return sum(np.linalg.det(np.where(np.arange(3)[None, :] == k, D, H)) for k in range(3))
