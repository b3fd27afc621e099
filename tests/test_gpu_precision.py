"""Reduced-precision RI-K builds against the FP64 build and the oracle.

Tolerances stated per form: exponent-aligned BF16 slices (Ozaki-style,
RT_PREC_OZAKI_BF16) must stay within 1e-8 relative Frobenius error of the
FP64 K and within 1e-8 relative error in E_x; the unaligned BF16x3 form is
reported against a looser 1e-4 bar (its BF16 hi/lo split keeps ~16 bits).
"""
import numpy as np
import pytest

from oracle import ri_oracle as O

pytestmark = pytest.mark.gpu


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def D():
    from rest_tensors_amd import device
    assert device.device_count() >= 1, "no GPU visible"
    return device


def make_B(D, nbas, naux, nocc, seed):
    dA = D.DeviceArray(nbas * nbas * naux)
    D.gen_ri3ao(dA, nbas, naux, seed=seed)
    dC = D.DeviceArray.from_host(O.gen_orthonormal(nbas, seed=seed + 1))
    dB = D.DeviceArray(nocc * nbas * naux)
    D.ao2mo_occ(dC, nbas, nbas, nocc, dA, naux, dB)
    del dA
    return dB


@pytest.mark.parametrize("shape", [(128, 384, 16), (96, 50, 5), (160, 64, 40),
                                   (512, 2048, 64)])
def test_ozaki_and_bf16x3_k(D, shape):
    nbas, naux, nocc = shape
    dB = make_B(D, nbas, naux, nocc, seed=1000 + nbas)
    K64, Koz, Kb3 = (D.DeviceArray(nbas * nbas) for _ in range(3))
    ex64 = D.rik(dB, nocc, nbas, naux, K64, precision="fp64")
    exoz = D.rik(dB, nocc, nbas, naux, Koz, precision="ozaki_bf16")
    exb3 = D.rik(dB, nocc, nbas, naux, Kb3, precision="bf16x3")
    k64 = K64.to_host((nbas, nbas))
    koz = Koz.to_host((nbas, nbas))
    kb3 = Kb3.to_host((nbas, nbas))
    e_oz, e_b3 = relf(koz, k64), relf(kb3, k64)
    print(f"{shape}: K rel err ozaki {e_oz:.2e}  bf16x3 {e_b3:.2e}; "
          f"E_x rel err ozaki {abs(exoz - ex64) / abs(ex64):.2e} "
          f"bf16x3 {abs(exb3 - ex64) / abs(ex64):.2e}")
    assert e_oz < 1e-8 and abs(exoz - ex64) < 1e-8 * abs(ex64)
    assert e_b3 < 1e-4
    assert np.array_equal(koz, koz.T)
    if nbas <= 160:
        B = dB.to_host((nocc, nbas, naux))
        K, ex, _ = O.rik(B)
        assert relf(koz, K) < 1e-8


@pytest.mark.parametrize("shape", [(128, 384, 16), (96, 50, 5), (160, 64, 40),
                                   (200, 300, 70), (512, 2048, 64)])
def test_ozaki_int8_k(D, shape):
    """INT8 Ozaki form: within 1e-8 of the FP64 K and E_x, exactly symmetric,
    and equal (to FP64 rounding of the fold, 1e-14) to the oracle's
    restatement of the same slicing and exact level sums."""
    from rest_tensors_amd._lib import lib
    nbas, naux, nocc = shape
    print(f"ozaki_int8 shape {shape}", flush=True)
    dB = make_B(D, nbas, naux, nocc, seed=3000 + nbas)
    K64, K8 = D.DeviceArray(nbas * nbas), D.DeviceArray(nbas * nbas)
    ex64 = D.rik(dB, nocc, nbas, naux, K64, precision="fp64")
    ex8 = D.rik(dB, nocc, nbas, naux, K8, precision="ozaki_int8")
    k64, k8 = K64.to_host((nbas, nbas)), K8.to_host((nbas, nbas))
    err = relf(k8, k64)
    print(f"{shape}: K rel err ozaki_int8 {err:.2e}, E_x rel err "
          f"{abs(ex8 - ex64) / abs(ex64):.2e}")
    assert err < 1e-8 and abs(ex8 - ex64) < 1e-8 * abs(ex64)
    assert np.array_equal(k8, k8.T)
    if nocc * nbas * naux <= 4_000_000:
        gp = lib().rt_ozaki_i8_group_size(nocc, nbas, naux)
        B = dB.to_host((nocc, nbas, naux))
        Ke = O.rik_ozaki_i8(B, gp)
        assert relf(k8, Ke) < 1e-14
        assert relf(k8, O.rik(B)[0]) < 1e-8


@pytest.mark.parametrize("shape", [(128, 384, 16), (96, 50, 5), (200, 300, 70),
                                   (512, 2048, 64)])
def test_split_fp32_k(D, shape):
    """Split FP32 (hi/lo planes, 3 FP32-MFMA products, FP64 fold per 32-deep
    k-chunk): the FP32 rounding of each chunk's hi*hi sum bounds the
    off-diagonal K entries near the FP32 unit roundoff, so the stated
    tolerance is 5e-8 for K (relative Frobenius) and 1e-8 for E_x (a sum of
    diagonal entries); exactly symmetric."""
    nbas, naux, nocc = shape
    print(f"split_fp32 shape {shape}", flush=True)
    dB = make_B(D, nbas, naux, nocc, seed=5000 + nbas)
    K64, K32 = D.DeviceArray(nbas * nbas), D.DeviceArray(nbas * nbas)
    ex64 = D.rik(dB, nocc, nbas, naux, K64, precision="fp64")
    ex32 = D.rik(dB, nocc, nbas, naux, K32, precision="split_fp32")
    k64, k32 = K64.to_host((nbas, nbas)), K32.to_host((nbas, nbas))
    err = relf(k32, k64)
    print(f"{shape}: K rel err split_fp32 {err:.2e}, E_x rel err "
          f"{abs(ex32 - ex64) / abs(ex64):.2e}")
    assert err < 5e-8 and abs(ex32 - ex64) < 1e-8 * abs(ex64)
    assert np.array_equal(k32, k32.T)
