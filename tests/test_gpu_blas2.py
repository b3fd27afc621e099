"""Level-2 / rank-k primitives (rt_dgemv_dev, rt_dsyrk_dev,
rt_einsum_ip_ip_p_dev) against the oracle restatements of _dgemv, _dsyrk and
einsum "ip,ip->p", and the RI-J composition REST builds from them on the
packed ri3ao (equal to the fused RI-J kernels)."""
import numpy as np
import pytest

from oracle import ri_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from rest_tensors_amd import device
    assert device.device_count() >= 1, "no GPU visible"
    return device


def relf(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("pad", [3, 4])
@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (37, 23), (5000, 3), (3, 5000), (70000, 40),
                                   (4099, 17), (1030, 9)])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgemv(D, shape, trans, pad):
    """lda = m + pad: an even lda takes the 16-byte kernels (odd m leaves one
    row alone), an odd one the 8-byte kernels."""
    m, n = shape
    rng = np.random.default_rng(m * 7 + n)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    lda = m + pad
    Ap = np.zeros((lda, n), order="F")
    Ap[:m] = A
    for incx, incy, beta in ((1, 1, 0.0), (2, 3, -0.5)):
        lx, ly = (m, n) if trans == "T" else (n, m)
        x = rng.standard_normal(1 + (lx - 1) * incx)
        y = rng.standard_normal(1 + (ly - 1) * incy)
        dA, dx, dy = (D.DeviceArray.from_host(v) for v in (Ap, x, y))
        D.dgemv(trans, m, n, 1.25, dA, lda, dx, beta, dy, incx=incx, incy=incy)
        out = dy.to_host()
        ref = O.dgemv(A, x, y, trans, 1.25, beta, incx, incy)
        assert relf(out, ref) < 1e-13
        # entries between the strided ones are untouched
        keep = np.ones(len(y), bool)
        keep[::incy] = False
        assert np.array_equal(out[keep], y[keep])


@pytest.mark.parametrize("trans,n,k", [("N", 70, 33), ("T", 70, 33), ("T", 256, 512),
                                       ("N", 1, 5)])
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_dsyrk(D, trans, n, k, uplo):
    rng = np.random.default_rng(n + k)
    A = np.asfortranarray(rng.standard_normal((n, k) if trans == "N" else (k, n)))
    Cm = np.asfortranarray(rng.standard_normal((n, n)))
    dA, dC = D.DeviceArray.from_host(A), D.DeviceArray.from_host(Cm)
    D.dsyrk(uplo, trans, n, k, 0.5, dA, A.shape[0], 2.0, dC, n)
    out = dC.to_host((n, n))
    ref = O.dsyrk(A, Cm, uplo, trans, 0.5, 2.0)
    assert relf(out, ref) < 1e-13
    other = np.tril(np.ones((n, n), bool), -1) if uplo == "U" else np.triu(np.ones((n, n), bool), 1)
    assert np.array_equal(out[other], Cm[other])


def test_einsum_ip_ip_p(D):
    rng = np.random.default_rng(3)
    A, B = rng.standard_normal((900, 50)), rng.standard_normal((700, 64))
    dA, dB = D.DeviceArray.from_host(A), D.DeviceArray.from_host(B)
    out = D.DeviceArray(50)
    D.einsum_ip_ip_p(dA, 900, 50, dB, 700, 64, out)
    assert relf(out.to_host(), O.einsum_ip_ip_p(A, B)) < 1e-14


def test_rij_composition_on_packed_layout(D):
    """gamma = Ap^T (w o D)_packed, J_packed = Ap gamma with the device
    primitives equals the fused RI-J kernels (and the oracle)."""
    nbas, naux, nocc = 96, 288, 12
    dA = D.DeviceArray(nbas * nbas * naux)
    D.gen_ri3ao(dA, nbas, naux, seed=77)
    Cm = O.gen_orthonormal(nbas, seed=78)
    dC = D.DeviceArray.from_host(Cm)
    dD, dJ = D.DeviceArray(nbas * nbas), D.DeviceArray(nbas * nbas)
    D.density(dC, nbas, nocc, dD)
    e_j = D.rij(dA, nbas, naux, dD, dJ)
    npair = nbas * (nbas + 1) // 2
    dAp = D.DeviceArray(npair * naux)
    D.rifull_to_matfull_symm(dA, nbas, naux, dAp)
    Dm = dD.to_host((nbas, nbas))
    dw = D.DeviceArray.from_host(O.upper_pack((2.0 - np.eye(nbas)) * Dm))
    dg, dJp = D.DeviceArray(naux), D.DeviceArray(npair)
    D.dgemv("T", npair, naux, 1.0, dAp, npair, dw, 0.0, dg)
    D.dgemv("N", npair, naux, 1.0, dAp, npair, dg, 0.0, dJp)
    dJ2 = D.DeviceArray(nbas * nbas)
    D.upper_unpack(dJp, npair, dJ2)
    J, J2 = dJ.to_host((nbas, nbas)), dJ2.to_host((nbas, nbas))
    assert relf(J2, J) < 1e-13
    assert abs(0.5 * float(np.sum(Dm * J2)) - e_j) < 1e-12 * abs(e_j)
