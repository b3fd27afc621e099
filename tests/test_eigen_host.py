"""Symmetric eigensolver: the oracle pinned by the reference's _dsyev doc
example and by closed forms, and the host stage of rt_dsyev_dev (the DSTEQR
restatement, rt_dsteqr_host) against LAPACK's own DSTEQR (through DSTEV).
CPU only: the host stage makes no HIP call."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import metric_oracle as M

# _dsyev doc example (src/matrix/matrix_blas_lapack.rs:286-318): the packed
# upper [1..6] unpacked to a symmetric 3 x 3; eigenvectors column-major
DOC_A = np.array([[1.0, 2.0, 4.0], [2.0, 3.0, 5.0], [4.0, 5.0, 6.0]])
DOC_W = np.array([-1.5066326307865059, -0.05739624271478554, 11.564028873501286])
DOC_V = np.array([-0.6827362941552275, -0.38559063640162244, 0.6206375864887483,
                  0.6202872696512856, -0.7547821848190948, 0.21341873532627673,
                  -0.3861539275363389, -0.5306823104260265, -0.7544941548144386]
                 ).reshape(3, 3).T


def test_oracle_dsyev_doc_example():
    v, w, info = M.dsyev(DOC_A, "V")
    assert info == 0
    assert np.sum((w - DOC_W) ** 2) < 1e-6 and np.abs(w - DOC_W).max() < 1e-13
    # sign-sensitive, as the doc test is
    assert np.sum((v - DOC_V) ** 2) < 1e-6 and np.abs(v - DOC_V).max() < 1e-13
    _, w2, _ = M.dsyev(DOC_A, "N")
    assert np.abs(w2 - DOC_W).max() < 1e-13


@pytest.mark.parametrize("n", [1, 2, 5, 40])
def test_oracle_dstev_closed_form(n):
    """[-1, 2, -1]: lambda_k = 2 - 2 cos(k pi/(n+1)), v_k(j) ~ sin(j k pi/(n+1))."""
    w, z, info = M.dstev(np.full(n, 2.0), np.full(n - 1, -1.0))
    assert info == 0
    k = np.arange(1, n + 1)
    assert np.abs(w - (2 - 2 * np.cos(k * np.pi / (n + 1)))).max() < 1e-14
    j = np.arange(1, n + 1)[:, None]
    ref = np.sin(j * k[None, :] * np.pi / (n + 1))
    ref /= np.linalg.norm(ref, axis=0)
    assert np.abs(np.abs(z) - np.abs(ref)).max() < 1e-12


def _tridiag_cases():
    rng = np.random.default_rng(0x5EED7001)
    cases = []
    for n in [1, 2, 3, 4, 7, 10, 33, 100, 257]:
        cases.append((f"gauss{n}", rng.standard_normal(n), rng.standard_normal(max(n - 1, 0))))
    d = rng.standard_normal(30)
    e = rng.standard_normal(29)
    e[[5, 6, 17]] = 0.0                     # splits into blocks, incl. a 1 x 1
    cases.append(("split", d, e))
    e2 = rng.standard_normal(29)
    e2[[3, 5, 9, 11]] = 0.0                 # 2 x 2 blocks (DLAEV2 path)
    cases.append(("blocks2", rng.standard_normal(30), e2))
    n = 60                                  # graded: QL on one end, QR on the other
    cases.append(("graded_down", 10.0 ** -np.linspace(0, 12, n), 10.0 ** -np.linspace(0.5, 12.5, n - 1)))
    cases.append(("graded_up", 10.0 ** np.linspace(-12, 0, n), 10.0 ** np.linspace(-12.5, -0.5, n - 1)))
    cases.append(("zero_diag", np.zeros(25), rng.standard_normal(24)))
    cases.append(("tiny", 1e-200 * rng.standard_normal(20), 1e-200 * rng.standard_normal(19)))
    cases.append(("huge", 1e200 * rng.standard_normal(20), 1e200 * rng.standard_normal(19)))
    return cases


@pytest.fixture(scope="module")
def D():
    from rest_tensors_amd import device as D
    return D


@pytest.mark.parametrize("name,d,e", _tridiag_cases(), ids=lambda x: x if isinstance(x, str) else "")
def test_dsteqr_host_matches_lapack(D, name, d, e):
    n = d.shape[0]
    w_ref, z_ref, info = M.dstev(d, e)
    assert info == 0
    dd, ee = d.copy(), e.copy()
    z = np.zeros((n, n), order="F")
    assert D.dsteqr_host("I", dd, ee, z) == 0
    scale = max(np.abs(d).max(initial=0.0), np.abs(e).max(initial=0.0))
    assert np.abs(dd - w_ref).max(initial=0.0) <= 1e-13 * scale
    # same algorithm, same rotations: eigenvectors agree with their signs
    assert np.abs(z - z_ref).max(initial=0.0) < 1e-10
    # eigenvalues only
    d2, e2 = d.copy(), e.copy()
    assert D.dsteqr_host("N", d2, e2) == 0
    assert np.abs(d2 - w_ref).max(initial=0.0) <= 1e-13 * scale


def test_dsteqr_host_residual_clustered(D):
    """Wilkinson W21+: eigenvalue pairs agree to ~1e-14, so eigenvectors are
    compared through the residual and orthogonality, not entry by entry."""
    n = 21
    d = np.abs(np.arange(n) - 10.0)
    e = np.ones(n - 1)
    w_ref, _, _ = M.dstev(d, e)
    dd, ee = d.copy(), e.copy()
    z = np.zeros((n, n), order="F")
    assert D.dsteqr_host("I", dd, ee, z) == 0
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.abs(dd - w_ref).max() < 1e-13 * 10
    assert np.abs(T @ z - z * dd[None, :]).max() < 1e-13 * 10
    assert np.abs(z.T @ z - np.eye(n)).max() < 1e-13


def test_oracle_lapack_power_threshold_drop():
    """Known eigenpairs: the eigenvalue below the threshold is dropped."""
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    lam = np.array([9.0, 4.0, 1.0, 0.25, 1e-9, 1e-12])
    S = (Q * lam) @ Q.T
    S = 0.5 * (S + S.T)
    X = M.lapack_power(S, -0.5, 1e-6)
    keep = lam >= 1e-6
    ref = (Q[:, keep] * lam[keep] ** -0.5) @ Q[:, keep].T
    assert np.abs(X - ref).max() < 1e-9


def _pack_upper(A):
    n = A.shape[0]
    return np.concatenate([A[:j + 1, j] for j in range(n)]) if n else np.zeros(0)


def test_oracle_dspgvx_closed_forms():
    rng = np.random.default_rng(11)
    n = 9
    X = rng.standard_normal((n, n))
    A = 0.5 * (X + X.T)
    # B = I: the standard problem
    Z, w = M.dspgvx(_pack_upper(A), _pack_upper(np.eye(n)), n, 4)
    _, wr, _ = M.dsyev(A, "N")
    assert np.abs(w[:4] - wr[:4]).max() < 1e-13 and np.all(w[4:] == 0.0)
    # diagonal A, B: lambda = a/b, vectors e_i / sqrt(b_i)
    a = rng.standard_normal(n)
    bb = rng.uniform(0.5, 2.0, n)
    Z, w = M.dspgvx(_pack_upper(np.diag(a)), _pack_upper(np.diag(bb)), n, n)
    order = np.argsort(a / bb)
    assert np.abs(w - (a / bb)[order]).max() < 1e-14
    ref = np.zeros((n, n))
    ref[order, np.arange(n)] = 1.0 / np.sqrt(bb[order])
    assert np.abs(np.abs(Z) - ref).max() < 1e-14
    # general: residual and B-orthonormality
    Y = rng.standard_normal((n, n))
    B = Y @ Y.T + n * np.eye(n)
    Z, w = M.dspgvx(_pack_upper(A), _pack_upper(B), n, 5)
    assert np.abs(A @ Z - B @ Z * w[None, :5]).max() < 1e-12
    assert np.abs(Z.T @ B @ Z - np.eye(5)).max() < 1e-12
