"""CPU oracle of the FP64 aux-metric functions — TEST INFRASTRUCTURE ONLY.

Restates, with scipy's LAPACK (the library the reference calls through the
lapack crate), the sequences of
  _dpotrf                     src/matrix/matrix_blas_lapack.rs:518-534
  cholesky_decompose_inverse  src/matrix/matrix_blas_lapack.rs:931-1002
  lapack_power                src/matrix/matrix_blas_lapack.rs:1004-1064
The product never imports this module.

Pinning: dpotrf by the doc example of matrix_blas_lapack.rs:502-517
(U = [2,6,1,-8,5,3]); the eigen-decomposition step of lapack_power by the
_dsyev doc eigenvalues (matrix_blas_lapack.rs:286-318); lapack_power(-0.5) by
X S X = I and cholesky_decompose_inverse by M M^T = S^-1 on the 8x8 matrix of
the reference's own test (src/lib.rs:170-190), and both against the closed
forms on diagonal and 2x2 matrices.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lapack


def dpotrf(A: np.ndarray, uplo: str):
    """LAPACK dpotrf on the uplo triangle; the other strict triangle is
    returned as it came in.  Returns (factor matrix, info)."""
    c, info = lapack.dpotrf(np.asfortranarray(A), lower=uplo.upper() == "L", clean=0)
    return c, int(info)


def cholesky_decompose_inverse(A: np.ndarray, uplo: str):
    """dpotrf, dpotri, mirror the triangle, dpotrf again, zero the other
    triangle (the reference's sequence).  None when the first dpotrf fails."""
    lower = uplo.upper() == "L"
    c, info = lapack.dpotrf(np.asfortranarray(A), lower=lower, clean=0)
    if info != 0:
        return None
    inv, info = lapack.dpotri(c, lower=lower)
    if info != 0:
        return None
    inv = np.tril(inv) + np.tril(inv, -1).T if lower else np.triu(inv) + np.triu(inv, 1).T
    m, _ = lapack.dpotrf(np.asfortranarray(inv), lower=lower, clean=0)
    return np.tril(m) if lower else np.triu(m)


def lapack_power(A: np.ndarray, p: float, threshold: float):
    """Eigen-decomposition of -A (dsyev, ascending), so eigenvalues of A come
    largest first; those >= threshold are kept, their eigenvectors scaled by
    sqrt(lambda)^p, and the result is E E^T (the others are dropped)."""
    w, v = np.linalg.eigh(-np.asarray(A))
    w = -w
    keep = w >= threshold
    E = v[:, keep] * (np.sqrt(w[keep]) ** p)[None, :]
    return E @ E.T


def spd(n: int, seed: int, lo: float = 1e-2, hi: float = 10.0) -> np.ndarray:
    """Seeded SPD test matrix with eigenvalues log-spaced in [lo, hi]."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.logspace(np.log10(lo), np.log10(hi), n)
    S = (Q * lam[None, :]) @ Q.T
    return np.asfortranarray(0.5 * (S + S.T))
