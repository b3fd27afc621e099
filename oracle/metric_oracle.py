"""CPU oracle of the FP64 aux-metric functions — TEST INFRASTRUCTURE ONLY.

Restates, with scipy's LAPACK (the library the reference calls through the
lapack crate), the sequences of
  _dpotrf                     src/matrix/matrix_blas_lapack.rs:518-534
  cholesky_decompose_inverse  src/matrix/matrix_blas_lapack.rs:931-1002
  _dsyev                      src/matrix/matrix_blas_lapack.rs:319-348 (DSYEV, uplo 'L')
  lapack_power                src/matrix/matrix_blas_lapack.rs:1004-1064
  _dspgvx / lapack_dspgvx     src/matrix/matrix_blas_lapack.rs:2123-2186, 1096-1147
  lapack_dspevx               src/matrix/matrix_blas_lapack.rs:1075-1095
and DSTEQR with vectors from the identity (through DSTEV), the tridiagonal
stage of DSYEV.  The product never imports this module.

Pinning: dpotrf by the doc example of matrix_blas_lapack.rs:502-517
(U = [2,6,1,-8,5,3]); dsyev by the _dsyev doc example (eigenvalues and the
sign-sensitive eigenvectors, matrix_blas_lapack.rs:286-318); dstev by the
closed form of the [-1, 2, -1] tridiagonal (eigenvalues 2 - 2 cos(k pi/(n+1)),
eigenvectors sin(j k pi/(n+1)) up to sign); lapack_power(-0.5) by X S X = I
and cholesky_decompose_inverse by M M^T = S^-1 on the 8x8 matrix of the
reference's own test (src/lib.rs:170-190), and both against the closed forms
on diagonal and 2x2 matrices; the threshold drop of lapack_power against a
matrix with known eigenpairs; dspgvx by its closed forms (B = I gives dsyev's
eigenvalues; diagonal A, B give a_ii/b_ii with vectors e_i/sqrt(b_ii)) and by
A Z = B Z W, Z^T B Z = I.
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


def dsyev(A: np.ndarray, jobz: str = "V"):
    """LAPACK dsyev on the lower triangle (what _dsyev calls).  Returns
    (eigenvectors or None, ascending eigenvalues, info)."""
    vec = jobz.upper() == "V"
    w, v, info = lapack.dsyev(np.asfortranarray(A), compute_v=1 if vec else 0, lower=1)
    return (v if vec else None), w, int(info)


def dstev(d: np.ndarray, e: np.ndarray):
    """LAPACK dstev with vectors = DSTEQR from the identity on the
    tridiagonal (d, e).  Returns (eigenvalues, eigenvectors, info)."""
    d = np.array(d, float)
    if d.shape[0] == 1:                  # scipy's wrapper wants a non-empty e
        return d.copy(), np.ones((1, 1)), 0
    w, z, info = lapack.dstev(d, np.array(e, float), compute_v=1)
    return w, z, int(info)


def _unpack_upper(packed: np.ndarray, n: int) -> np.ndarray:
    A = np.zeros((n, n))
    iu = np.triu_indices(n)
    # packed order: column j, rows 0..=j  (j(j+1)/2 + i)
    order = np.argsort(iu[1] * (iu[1] + 1) // 2 + iu[0])
    A[iu[0][order], iu[1][order]] = packed
    return np.triu(A) + np.triu(A, 1).T


def dspgvx(Ap: np.ndarray, Bp: np.ndarray, n: int, num_orb: int):
    """_dspgvx (src/matrix/matrix_blas_lapack.rs:2123-2186): DSPGVX itype 1,
    jobz 'V', range 'I' (1..num_orb), uplo 'U' on packed A, B; restated with
    LAPACK's DSYGVX on the unpacked matrices (the same algorithm on full
    storage).  Returns (Z [n, num_orb], w [n] with zeros after num_orb)."""
    import scipy.linalg as sl
    A, B = _unpack_upper(Ap, n), _unpack_upper(Bp, n)
    w = np.zeros(n)
    if num_orb == 0:
        return np.zeros((n, 0)), w
    ww, Z = sl.eigh(A, B, lower=False, driver="gvx", subset_by_index=[0, num_orb - 1])
    w[:num_orb] = ww
    return Z, w


def dspevx(Ap: np.ndarray, n: int):
    """lapack_dspevx (src/matrix/matrix_blas_lapack.rs:1075-1095): all
    eigenpairs of a packed-upper matrix (DSYEVX on the unpacked one)."""
    import scipy.linalg as sl
    A = _unpack_upper(Ap, n)
    w, Z = sl.eigh(A, lower=False, driver="evx")
    return Z, w


def lapack_power(A: np.ndarray, p: float, threshold: float):
    """Eigen-decomposition of -A (dsyev, ascending), so eigenvalues of A come
    largest first; those >= threshold are kept, their eigenvectors scaled by
    sqrt(lambda)^p, and the result is E E^T (the others are dropped)."""
    v, w, info = dsyev(-np.asarray(A), "V")
    assert info == 0
    w = -w
    f = np.zeros_like(w)
    keep = w >= threshold
    f[keep] = np.sqrt(w[keep]) ** p
    E = v * f[None, :]
    return E @ E.T


def spd(n: int, seed: int, lo: float = 1e-2, hi: float = 10.0) -> np.ndarray:
    """Seeded SPD test matrix with eigenvalues log-spaced in [lo, hi]."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.logspace(np.log10(lo), np.log10(hi), n)
    S = (Q * lam[None, :]) @ Q.T
    return np.asfortranarray(0.5 * (S + S.T))
