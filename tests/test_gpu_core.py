"""GPU parity of the FP64 hot path against the oracle, through the C ABI.

Sizes: tiny cases against the golden vectors of the compiled reference, the
CPU-reference workload (nbas=128, naux=384, nocc=16) in full, and the GPU
parity workload (nbas=512, naux=2048, nocc=64) with the tolerances the
workload states: bit-exact index layout, <= 1e-12 relative Frobenius error.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import ri_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative Frobenius error bar of the FP64 path


def relf(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="module")
def D():
    from rest_tensors_amd import device
    assert device.device_count() >= 1, "no GPU visible"
    print("device:", device.device_name(), flush=True)
    return device


def ip(v):
    return C.byref(C.c_int(int(v)))


def dp(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- GEMM engine -----------------------------------------------------------------
@pytest.mark.parametrize("opa", "NT")
@pytest.mark.parametrize("opb", "NT")
@pytest.mark.parametrize("mnk", [(1, 1, 1), (17, 5, 33), (64, 64, 16), (130, 257, 67),
                                 (300, 129, 1000), (130, 257, 96), (5, 3, 32),
                                 (256, 384, 512)])
def test_dgemm_ops_shapes(D, opa, opb, mnk):
    m, n, k = mnk
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = np.asfortranarray(rng.standard_normal((m, k) if opa == "N" else (k, m)))
    B = np.asfortranarray(rng.standard_normal((k, n) if opb == "N" else (n, k)))
    Cm = np.asfortranarray(rng.standard_normal((m, n)))
    ref = 0.7 * ((A if opa == "N" else A.T) @ (B if opb == "N" else B.T)) - 1.3 * Cm
    dA, dB, dC = D.DeviceArray.from_host(A), D.DeviceArray.from_host(B), D.DeviceArray.from_host(Cm)
    D.dgemm(opa, opb, m, n, k, 0.7, dA, A.shape[0], dB, B.shape[0], -1.3, dC, m)
    assert relf(dC.to_host((m, n)), ref) < 1e-13


def test_dgemm_batched_big_tiles(D):
    # 128 batches of 256x256x96: enough tiles to select the 128x128 variant
    from rest_tensors_amd._lib import lib, check
    b, m, n, k = 128, 256, 256, 96
    rng = np.random.default_rng(5)
    A = rng.standard_normal((b, k, m))   # per batch column-major m x k
    B = rng.standard_normal((b, n, k))   # per batch column-major k x n
    dA, dB = D.DeviceArray.from_host(A.reshape(-1)), D.DeviceArray.from_host(B.reshape(-1))
    dC = D.DeviceArray(b * m * n)
    check(lib().rt_dgemm_strided_batched_dev(b"N", b"N", m, n, k, 1.0, dA.ptr, m, m * k,
                                             dB.ptr, k, k * n, 0.0, dC.ptr, m, m * n, b,
                                             None))
    out = dC.to_host().reshape(b, n, m)
    for q in (0, 77, 127):
        Aq = A[q].reshape(-1).reshape((m, k), order="F")
        Bq = B[q].reshape(-1).reshape((k, n), order="F")
        assert relf(out[q].reshape(-1).reshape((m, n), order="F"), Aq @ Bq) < 1e-13


@pytest.mark.parametrize("mnkb", [(256, 256, 96, 128), (1024, 128, 64, 24), (77, 130, 32, 9)])
def test_dgemm_batched_kmajor(D, mnkb):
    # op(A) = A^T, op(B) = B with k-major storage: the direct-to-LDS engine
    from rest_tensors_amd._lib import lib, check
    m, n, k, b = mnkb
    rng = np.random.default_rng(m + n + k + b)
    A = rng.standard_normal((b, m, k))   # per batch k x m column-major (k fastest)
    B = rng.standard_normal((b, n, k))   # per batch k x n column-major
    dA, dB = D.DeviceArray.from_host(A.reshape(-1)), D.DeviceArray.from_host(B.reshape(-1))
    Cm = rng.standard_normal((b, n, m))
    dC = D.DeviceArray.from_host(Cm.reshape(-1))
    check(lib().rt_dgemm_strided_batched_dev(b"T", b"N", m, n, k, 1.5, dA.ptr, k, m * k,
                                             dB.ptr, k, k * n, 0.25, dC.ptr, m, m * n, b,
                                             None))
    out = dC.to_host().reshape(b, n, m)
    for q in (0, b // 2, b - 1):
        Aq = A[q].reshape(-1).reshape((k, m), order="F")
        Bq = B[q].reshape(-1).reshape((k, n), order="F")
        Cq = Cm[q].reshape(-1).reshape((m, n), order="F")
        ref = 1.5 * Aq.T @ Bq + 0.25 * Cq
        assert relf(out[q].reshape(-1).reshape((m, n), order="F"), ref) < 1e-13


def test_dgemm_rejects_bad_shapes(D):
    from rest_tensors_amd._lib import RtError
    a = D.DeviceArray(16)
    with pytest.raises(RtError):
        D.dgemm("N", "N", 4, 4, 4, 1.0, a, 2, a, 4, 0.0, a, 4)   # lda < m
    with pytest.raises(RtError):
        D.dgemm("X", "N", 4, 4, 4, 1.0, a, 4, a, 4, 0.0, a, 4)


# ---- drop-in symbols against the compiled reference's golden vectors ------------------
def dropin_ao2mo(Cm, A):
    from rest_tensors_amd._lib import lib
    nbas, ns = Cm.shape
    naux = A.shape[2]
    R = np.full((naux, ns, ns), np.nan, order="F")
    Cm = np.asfortranarray(Cm)
    A = np.asfortranarray(A)
    lib().ri_ao2mo_f_(dp(Cm), dp(A), dp(R), ip(ns), ip(nbas), ip(naux))
    return R


@pytest.mark.parametrize("case", ["ao2mo_a", "ao2mo_b", "ao2mo_c", "ao2mo_ns"])
def test_dropin_ao2mo_golden(D, golden, case):
    R = dropin_ao2mo(golden[f"{case}_C"], golden[f"{case}_A"])
    assert relf(R, golden[f"{case}_R"]) < TOL


def test_dropin_ao2mo_bench_closed_form(D, golden):
    R = dropin_ao2mo(np.ones((10, 10)), np.full((10, 10, 20), 2.0))
    assert np.array_equal(R, golden["ao2mo_bench_R"])  # every entry exactly 200.0


def test_dropin_ao2mo_rectangular_intended(D):
    # ns < nbas: the reference is undefined there; the intended result holds
    A = O.gen_ri3ao(40, 11, seed=3)
    Cm = np.asfortranarray(O.gen_orthonormal(40, seed=4)[:, :13])
    assert relf(dropin_ao2mo(Cm, A), O.ri_ao2mo_ref(Cm, A)) < TOL


def test_dropin_general_dgemm_golden(D, golden):
    from rest_tensors_amd._lib import lib
    for ci in range(int(golden["gemm_ncases"])):
        p = f"gemm{ci}_"
        A, B = golden[p + "A"], golden[p + "B"]
        Cm = np.asfortranarray(golden[p + "C"].copy())
        ra, rb, rc = golden[p + "ra"], golden[p + "rb"], golden[p + "rc"]
        opa, opb = (bytes([int(c)]) for c in golden[p + "ops"])
        alpha, beta = golden[p + "ab"]
        lib().general_dgemm_f_(
            dp(A), ip(A.shape[0]), ip(A.shape[1]), ip(ra[0][0]), ip(ra[0][1]), ip(ra[1][0]),
            ip(ra[1][1]), opa, dp(B), ip(B.shape[0]), ip(B.shape[1]), ip(rb[0][0]),
            ip(rb[0][1]), ip(rb[1][0]), ip(rb[1][1]), opb, dp(Cm), ip(Cm.shape[0]),
            ip(Cm.shape[1]), ip(rc[0][0]), ip(rc[0][1]), ip(rc[1][0]), ip(rc[1][1]),
            C.byref(C.c_double(alpha)), C.byref(C.c_double(beta)))
        if ci < 3:   # the doc examples are small integers: exact
            assert np.array_equal(Cm, golden[p + "out"]), ci
        else:
            assert relf(Cm, golden[p + "out"]) < 1e-14, ci


def test_dropin_general_dgemm_bad_shape_leaves_output(D, capfd):
    from rest_tensors_amd._lib import lib
    A = np.asfortranarray(np.ones((3, 3)))
    Cm = np.asfortranarray(np.full((3, 3), 2.0))
    lib().general_dgemm_f_(dp(A), ip(3), ip(3), ip(0), ip(2), ip(0), ip(3), b"N",
                           dp(A), ip(3), ip(3), ip(0), ip(2), ip(0), ip(2), b"N",
                           dp(Cm), ip(3), ip(3), ip(0), ip(2), ip(0), ip(2),
                           C.byref(C.c_double(1.0)), C.byref(C.c_double(1.0)))
    assert np.all(Cm == 2.0)
    assert "general_dgemm_f_" in capfd.readouterr().err


def test_dropin_copies_golden(D, golden):
    from rest_tensors_amd._lib import lib
    F, T = golden["cmm_F"], np.asfortranarray(golden["cmm_T"].copy())
    xl, yl, fxs, fys, txs, tys = golden["cmm_args"]
    lib().copy_mm_(ip(xl), ip(yl), dp(F), ip(9), ip(8), ip(fxs), ip(fys), dp(T), ip(7), ip(10),
                   ip(txs), ip(tys))
    assert np.array_equal(T, golden["cmm_out"])
    F, R0 = golden["cr_F"], golden["cr_R"]
    for mod in range(3):
        R = np.asfortranarray(R0.copy())
        xl, yl, fxs, fys, txs, tys, tx3 = golden[f"cmr{mod}_args"]
        lib().copy_mr_(ip(xl), ip(yl), dp(F), ip(9), ip(8), ip(fxs), ip(fys), dp(R), ip(6),
                       ip(7), ip(8), ip(txs), ip(tys), ip(tx3), ip(mod))
        assert np.array_equal(R, golden[f"cmr{mod}_out"]), mod
        Fc = np.asfortranarray(F.copy())
        xl, yl, fxs, fys, fx3, txs, tys = golden[f"crm{mod}_args"]
        lib().copy_rm_(ip(xl), ip(yl), dp(R0), ip(6), ip(7), ip(8), ip(fxs), ip(fys),
                       ip(fx3), ip(mod), dp(Fc), ip(9), ip(8), ip(txs), ip(tys))
        assert np.array_equal(Fc, golden[f"crm{mod}_out"]), mod
    R = np.asfortranarray(R0.copy())
    a = golden["crr_args"]
    F2 = golden["crr_F"]
    lib().copy_rr_(ip(a[0]), ip(a[1]), ip(a[2]), dp(F2), ip(5), ip(9), ip(4), ip(a[3]),
                   ip(a[4]), ip(a[5]), dp(R), ip(6), ip(7), ip(8), ip(a[6]), ip(a[7]), ip(a[8]))
    assert np.array_equal(R, golden["crr_out"])


def test_dropin_special_dgemm(D):
    from rest_tensors_amd._lib import lib
    rng = np.random.default_rng(9)
    T = np.asfortranarray(rng.standard_normal((8, 5, 12)))
    Bm = np.asfortranarray(rng.standard_normal((10, 9)))
    ref = O.special_dgemm_01(T, 2, 5, 3, 6, Bm, 1, 6, 2, 6, 0.5, -2.0)
    Tc = np.asfortranarray(T.copy())
    lib().special_dgemm_f_01_(dp(Tc), ip(8), ip(5), ip(12), ip(2), ip(5), ip(0), ip(3), ip(6),
                              dp(Bm), ip(10), ip(9), ip(1), ip(6), ip(2), ip(6),
                              C.byref(C.c_double(0.5)), C.byref(C.c_double(-2.0)))
    assert relf(Tc, ref) < 1e-14


@pytest.mark.parametrize("case", [(8, 5, 12, 2, 5, 3, 6, 10, 9, 1, 6, 2, 6, 0.5, -2.0),
                                  (40, 33, 70, 0, 40, 0, 70, 70, 70, 0, 70, 0, 70, 1.0, 0.0),
                                  (17, 3, 9, 4, 0, 2, 5, 5, 5, 0, 5, 0, 5, 1.0, 1.0)])
def test_special_dgemm_dev(D, case):
    """Device buffers, in place; same argument meaning as special_dgemm_f_01_."""
    xa, ya, za, sx, lx, sz, lz, rb, cb, srb, lrb, scb, lcb, al, be = case
    rng = np.random.default_rng(10)
    T = np.asfortranarray(rng.standard_normal((xa, ya, za)))
    Bm = np.asfortranarray(rng.standard_normal((rb, cb)))
    ref = O.special_dgemm_01(T, sx, lx, sz, lz, Bm, srb, lrb, scb, lcb, al, be)
    dT, dB = D.DeviceArray.from_host(T), D.DeviceArray.from_host(Bm)
    D.special_dgemm(dT, xa, ya, za, sx, lx, sz, lz, dB, rb, cb, srb, lrb, scb, lcb, al, be)
    assert relf(dT.to_host((xa, ya, za)), ref) < 1e-14


def test_ri3fn_producer(D):
    """ri3fn = O V^-1/2 (metric power, then the per-slice contraction), and
    O M with V^-1 = M M^T (Cholesky route): both give ri3fn ri3fn^T = O V^-1 O^T."""
    from oracle import metric_oracle as MO
    nbas, naux = 24, 72
    dO = D.DeviceArray(nbas * nbas * naux)
    D.gen_ri3ao(dO, nbas, naux, seed=0x5EED7900)
    Om = dO.to_host((nbas * nbas, naux))
    V = MO.spd(naux, seed=0x5EED7901, lo=1e-4, hi=5.0)
    dV, dM = D.DeviceArray.from_host(V), D.DeviceArray(naux * naux)
    D.matrix_power(naux, dV, naux, -0.5, 1e-10, dM, naux)
    D.special_dgemm(dO, nbas, nbas, naux, 0, nbas, 0, naux, dM, naux, naux, 0, naux, 0, naux,
                    1.0, 0.0)
    R = dO.to_host((nbas * nbas, naux))
    ref = Om @ MO.lapack_power(V, -0.5, 1e-10)
    assert relf(R, ref) < 1e-11
    G = Om @ np.linalg.solve(V, Om.T)
    assert relf(R @ R.T, G) < 1e-10
    dO2, dL = D.DeviceArray.from_host(np.asfortranarray(Om)), D.DeviceArray(naux * naux)
    D.cholesky_decompose_inverse("L", naux, dV, naux, dL, naux)
    D.special_dgemm(dO2, nbas, nbas, naux, 0, nbas, 0, naux, dL, naux, naux, 0, naux, 0, naux,
                    1.0, 0.0)
    R2 = dO2.to_host((nbas * nbas, naux))
    assert relf(R2 @ R2.T, G) < 1e-10


# ---- device entry points -----------------------------------------------------------
def test_gen_ri3ao_matches_host_generator(D):
    nbas, naux = 37, 5
    d = D.DeviceArray(nbas * nbas * naux)
    D.gen_ri3ao(d, nbas, naux, seed=77)
    A = d.to_host((nbas, nbas, naux))
    ref = O.gen_ri3ao(nbas, naux, seed=77)
    assert np.array_equal(A, A.transpose(1, 0, 2))
    assert np.abs(A - ref).max() < 1e-13


def test_ao2mo_identity_bit_exact(D):
    nbas, naux = 96, 24
    dA = D.DeviceArray(nbas * nbas * naux)
    D.gen_ri3ao(dA, nbas, naux, seed=1)
    A = dA.to_host((nbas, nbas, naux))
    dC = D.DeviceArray.from_host(np.eye(nbas))
    dR = D.DeviceArray(naux * nbas * nbas)
    D.ao2mo(dC, dA, dR, nbas, nbas, naux)
    R = dR.to_host((naux, nbas, nbas))
    # ri3mo[P + naux*i + naux*ns*j] == ri3ao[i + nbas*j + nbas^2*P], exactly
    assert np.array_equal(R, np.transpose(A, (2, 0, 1)))
    dB = D.DeviceArray(8 * nbas * naux)
    D.ao2mo_occ(dC, nbas, nbas, 8, dA, naux, dB)
    assert np.array_equal(dB.to_host((8, nbas, naux)), A[:8])


def pipeline(D, nbas, naux, nocc, seed, square_subset=None):
    """Runs the FP64 pipeline on the GPU; returns host copies of inputs/outputs."""
    dA = D.DeviceArray(nbas * nbas * naux)
    D.gen_ri3ao(dA, nbas, naux, seed=seed)
    Cm = O.gen_orthonormal(nbas, seed=seed + 1)
    dC = D.DeviceArray.from_host(Cm)
    dD = D.DeviceArray(nbas * nbas)
    D.density(dC, nbas, nocc, dD)
    dJ = D.DeviceArray(nbas * nbas)
    dg = D.DeviceArray(naux)
    e_j = D.rij(dA, nbas, naux, dD, dJ, dg)
    dB = D.DeviceArray(nocc * nbas * naux)
    D.ao2mo_occ(dC, nbas, nbas, nocc, dA, naux, dB)
    dK = D.DeviceArray(nbas * nbas)
    e_x = D.rik(dB, nocc, nbas, naux, dK)
    e_x2 = D.exchange_energy(dB, nocc, nbas, naux)
    out = dict(A=dA.to_host((nbas, nbas, naux)), C=Cm, D=dD.to_host((nbas, nbas)),
               J=dJ.to_host((nbas, nbas)), gamma=dg.to_host(), e_j=e_j,
               B=dB.to_host((nocc, nbas, naux)), K=dK.to_host((nbas, nbas)), e_x=e_x,
               e_x2=e_x2)
    if square_subset is not None:
        dR = D.DeviceArray(naux * nbas * nbas)
        D.ao2mo(dC, dA, dR, nbas, nbas, naux)
        out["R"] = dR.to_host((naux, nbas, nbas))
    return out


def check_pipeline(g, nocc, square_subset=None):
    Dref = O.density(g["C"], nocc)
    assert relf(g["D"], Dref) < TOL
    gamma, J, e_j = O.rij(g["A"], g["D"])
    assert relf(g["gamma"], gamma) < TOL
    assert relf(g["J"], J) < TOL
    assert abs(g["e_j"] - e_j) <= TOL * abs(e_j)
    B = O.ao2mo_occ(g["C"], g["A"], nocc)
    assert relf(g["B"], B) < TOL
    K, e_x, e_x2 = O.rik(g["B"])
    assert relf(g["K"], K) < TOL
    assert np.array_equal(g["K"], g["K"].T)
    assert abs(g["e_x"] - e_x) <= TOL * abs(e_x)
    assert abs(g["e_x2"] - e_x2) <= TOL * abs(e_x2)
    if square_subset is not None:
        naux = g["A"].shape[2]
        rng = np.random.default_rng(0)
        Ps = np.sort(rng.choice(naux, size=min(square_subset, naux), replace=False))
        ref = O.ri_ao2mo_ref(g["C"], g["A"][:, :, Ps])
        assert relf(g["R"][Ps], ref) < TOL


def test_pipeline_cpu_reference_workload(D):
    # nbas=128, naux=384, nocc=16: everything against the oracle in full
    g = pipeline(D, 128, 384, 16, seed=0x5EED0001, square_subset=384)
    check_pipeline(g, 16, square_subset=384)


def test_pipeline_gpu_parity_workload(D):
    # nbas=512, naux=2048, nocc=64 (ri3ao 4.3 GB); square ao2mo checked on 48 slices
    g = pipeline(D, 512, 2048, 64, seed=0x5EED0002, square_subset=48)
    check_pipeline(g, 64, square_subset=48)


def test_odd_shapes_pipeline(D):
    g = pipeline(D, 77, 13, 5, seed=99, square_subset=13)
    check_pipeline(g, 5, square_subset=13)


@pytest.mark.parametrize("nbas,naux", [(2, 1), (2, 9), (6, 3), (64, 7), (130, 17), (258, 8)])
def test_rij_even_nbas_shapes(D, nbas, naux):
    """Even nbas takes the 16-byte gamma path (8 slices per block): naux not a
    multiple of 8, columns shorter than one 64-pair chunk and longer than
    two, against the oracle."""
    g = pipeline(D, nbas, naux, 1, seed=700 + nbas + naux)
    gamma, J, e_j = O.rij(g["A"], g["D"])
    assert relf(g["gamma"], gamma) < TOL
    assert relf(g["J"], J) < TOL
    assert abs(g["e_j"] - e_j) <= TOL * abs(e_j)


# ---- packed upper -------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 100, 1024])
def test_upper_unpack_pack_bit_exact(D, n):
    L = n * (n + 1) // 2
    packed = np.random.default_rng(n).standard_normal(L)
    dp_ = D.DeviceArray.from_host(packed)
    dfull = D.DeviceArray(n * n)
    D.upper_unpack(dp_, L, dfull)
    full = dfull.to_host((n, n))
    assert np.array_equal(full, O.upper_unpack(packed))
    dback = D.DeviceArray(L)
    D.upper_pack(dfull, n, dback)
    assert np.array_equal(dback.to_host(), packed)


def test_upper_unpack_rejects_non_triangular(D):
    from rest_tensors_amd._lib import RtError
    d = D.DeviceArray(16)
    with pytest.raises(RtError):
        D.upper_unpack(d, 7, d)


def test_rifull_to_matfull_symm_bit_exact(D):
    nao, naux = 45, 7
    d = D.DeviceArray(nao * nao * naux)
    D.gen_ri3ao(d, nao, naux, seed=5)
    ri = d.to_host((nao, nao, naux))
    out = D.DeviceArray(nao * (nao + 1) // 2 * naux)
    D.rifull_to_matfull_symm(d, nao, naux, out)
    assert np.array_equal(out.to_host((nao * (nao + 1) // 2, naux)), O.rifull_to_matfull_symm(ri))


@pytest.mark.parametrize("order", ["jik", "jki", "kji", "ikj"])
@pytest.mark.parametrize("size", [(3, 2, 2), (1, 1, 1), (33, 65, 7), (130, 70, 33), (1, 300, 5),
                                  (257, 1, 3), (64, 64, 64)])
def test_rifull_transpose_bit_exact(D, order, size):
    """RIFull::transpose_* (ri.rs:225-290) against the oracle, bit for bit."""
    print(f"[transpose] {order} {size}", flush=True)
    rng = np.random.default_rng(0x5EED7500)
    ri = np.asfortranarray(rng.standard_normal(size))
    ref = O.rifull_transpose(ri, order)
    dsrc = D.DeviceArray.from_host(ri)
    dout = D.DeviceArray(ri.size)
    D.rifull_transpose(dsrc, *size, order, dout)
    assert np.array_equal(dout.to_host(ref.shape), ref)


def test_rifull_transpose_ri3ao_round_trip(D):
    """A C2-sized ri3ao slab: jki three times is the identity."""
    nb, naux = 512, 64
    d0 = D.DeviceArray(nb * nb * naux)
    D.gen_ri3ao(d0, nb, naux, seed=0x5EED7501)
    d1, d2 = D.DeviceArray(nb * nb * naux), D.DeviceArray(nb * nb * naux)
    D.rifull_transpose(d0, nb, nb, naux, "jki", d1)      # [nb, naux, nb]
    D.rifull_transpose(d1, nb, naux, nb, "jki", d2)      # [naux, nb, nb]
    D.rifull_transpose(d2, naux, nb, nb, "jki", d1)      # [nb, nb, naux]
    assert np.array_equal(d1.to_host((nb * nb * naux,)), d0.to_host((nb * nb * naux,)))


def test_rifull_transpose_rejects(D):
    from rest_tensors_amd._lib import RtError
    a = D.DeviceArray(8)
    with pytest.raises(RtError):
        D.rifull_transpose(a, 2, 2, 2, "kij", D.DeviceArray(8))
    with pytest.raises(RtError):
        D.rifull_transpose(a, 2, 2, 2, "jik", a)          # in place is refused
