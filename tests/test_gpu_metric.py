"""FP64 aux-metric functions on the device (rt_dpotrf_dev,
rt_cholesky_decompose_inverse_dev, rt_matrix_power_dev) against the oracle's
LAPACK restatements (oracle/metric_oracle.py), the reference's doc example
and the matrix of the reference's own test (src/lib.rs:170-190)."""
import json
import os

import numpy as np
import pytest

from oracle import metric_oracle as M
from oracle import ri_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from rest_tensors_amd import device
    assert device.device_count() >= 1, "no GPU visible"
    return device


def lib_rs_8x8():
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "matfull_inverse_and_power_8x8.json")))
    return np.array(d["data"]).reshape(d["size"], order="F")


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_dpotrf_doc_example(D):
    A = O.upper_unpack(np.array([4.0, 12.0, 37.0, -16.0, -43.0, 98.0]))
    dA = D.DeviceArray.from_host(A)
    assert D.dpotrf("U", 3, dA, 3) == 0
    assert np.array_equal(O.upper_pack(dA.to_host((3, 3))), [2.0, 6.0, 1.0, -8.0, 5.0, 3.0])


@pytest.mark.parametrize("n", [8, 100, 257, 1000])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dpotrf(D, n, uplo):
    S = M.spd(n, seed=n)
    rng = np.random.default_rng(1)
    junk = rng.standard_normal((n, n))
    lower = uplo == "L"
    A = np.where(np.tril(np.ones((n, n), bool)) if lower else np.triu(np.ones((n, n), bool)),
                 S, junk)                         # the other triangle holds junk
    lda = n + 5
    Ap = np.zeros((lda, n), order="F")
    Ap[:n] = A
    dA = D.DeviceArray.from_host(Ap)
    assert D.dpotrf(uplo, n, dA, lda) == 0
    out = dA.to_host((lda, n))[:n]
    ref, info = M.dpotrf(A, uplo)
    assert info == 0
    tri = np.tril(np.ones((n, n), bool)) if lower else np.triu(np.ones((n, n), bool))
    assert relf(out[tri], ref[tri]) < 1e-12
    assert np.array_equal(out[~tri], junk[~tri])


def test_dpotrf_info_not_positive_definite(D):
    S = M.spd(150, seed=3)
    S[100, 100] = -5.0                         # breaks definiteness at some order
    _, info_ref = M.dpotrf(S, "L")
    assert info_ref > 0
    assert D.dpotrf("L", 150, D.DeviceArray.from_host(S), 150) == info_ref


@pytest.mark.parametrize("n", [8, 200, 513])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_cholesky_decompose_inverse(D, n, uplo):
    S = lib_rs_8x8() if n == 8 else M.spd(n, seed=10 + n)
    dA, dO = D.DeviceArray.from_host(S), D.DeviceArray(n * n)
    D.cholesky_decompose_inverse(uplo, n, dA, n, dO, n)
    out = dO.to_host((n, n))
    ref = M.cholesky_decompose_inverse(S, uplo)
    assert relf(out, ref) < 1e-10
    assert np.array_equal(out, np.tril(out) if uplo == "L" else np.triu(out))


@pytest.mark.parametrize("n", [700, 1025])
def test_cholesky_decompose_inverse_ill_conditioned(D, n):
    """Eigenvalues over 5 decades (cond 1e5): the recursive triangular inverse
    stays within the conditioning of LAPACK's result; n = 1025 leaves a
    one-row block for the top-level join."""
    S = M.spd(n, seed=40 + n, lo=1e-5, hi=1.0)
    dA, dO = D.DeviceArray.from_host(S), D.DeviceArray(n * n)
    D.cholesky_decompose_inverse("L", n, dA, n, dO, n)
    out = dO.to_host((n, n))
    ref = M.cholesky_decompose_inverse(S, "L")
    assert relf(out, ref) < 1e-9
    # M M^T = S^-1: M^T S M = I to the conditioning
    assert np.abs(out.T @ S @ out - np.eye(n)).max() < 1e-9


def test_cholesky_decompose_inverse_not_pd(D):
    from rest_tensors_amd._lib import RtError
    S = np.diag([1.0, -1.0, 2.0])
    with pytest.raises(RtError):
        D.cholesky_decompose_inverse("L", 3, D.DeviceArray.from_host(S), 3, D.DeviceArray(9), 3)


@pytest.mark.parametrize("p", [-0.5, 0.5, -1.0, 1.0])
@pytest.mark.parametrize("n", [8, 300])
def test_matrix_power(D, p, n):
    S = lib_rs_8x8() if n == 8 else M.spd(n, seed=20 + n, lo=0.05, hi=20.0)
    dA, dO = D.DeviceArray.from_host(S), D.DeviceArray(n * n)
    D.matrix_power(n, dA, n, p, 1e-5, dO, n)
    out = dO.to_host((n, n))
    ref = M.lapack_power(S, p, 1e-5)
    assert relf(out, ref) < 1e-10
    if p == -0.5:
        assert np.abs(out @ S @ out - np.eye(n)).max() < 1e-9


@pytest.mark.parametrize("p", [-0.5, 0.5])
def test_matrix_power_ill_conditioned(D, p):
    """cond = 1e6, nothing dropped."""
    n = 400
    S = M.spd(n, seed=99, lo=1e-6, hi=1.0)
    dA, dO = D.DeviceArray.from_host(S), D.DeviceArray(n * n)
    D.matrix_power(n, dA, n, p, 1e-10, dO, n)
    out = dO.to_host((n, n))
    assert relf(out, M.lapack_power(S, p, 1e-10)) < 1e-8
    assert np.array_equal(out, out.T)


def test_matrix_power_threshold_drop(D):
    """The reference drops eigenvalues below the threshold: 1e-12 is dropped
    and its eigenvector contributes nothing."""
    S = np.diag([4.0, 1e-12, 2.0])
    dO = D.DeviceArray(9)
    D.matrix_power(3, D.DeviceArray.from_host(S), 3, -0.5, 1e-10, dO, 3)
    out = dO.to_host((3, 3))
    assert np.abs(out - np.diag([0.5, 0.0, 2 ** -0.5])).max() < 1e-14
    assert np.abs(out - M.lapack_power(S, -0.5, 1e-10)).max() < 1e-14
