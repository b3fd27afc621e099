"""Pins the CPU oracle (oracle/ri_oracle.py) before it is trusted as the checker.

Each oracle function is checked against at least one of: golden vectors of the
reference's own compiled Fortran (tests/golden/ref_vectors.npz), doc values of
the reference, closed forms, invariants, or brute force on tiny inputs.
"""
import math

import numpy as np
import pytest

from oracle import ri_oracle as O


def relf(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


# ---- ao2mo ---------------------------------------------------------------------
@pytest.mark.parametrize("case", ["ao2mo_a", "ao2mo_b", "ao2mo_c", "ao2mo_ns"])
def test_ao2mo_matches_compiled_reference(golden, case):
    Cm, A, R = golden[f"{case}_C"], golden[f"{case}_A"], golden[f"{case}_R"]
    out = O.ri_ao2mo_ref(Cm, A)
    assert out.shape == R.shape
    assert relf(out, R) < 1e-13
    assert relf(O.ri_ao2mo_v01(Cm, A), R) < 1e-13


def test_ao2mo_bench_closed_form(golden):
    # benches/bench_tensors.rs:5-10: ri3fn == 2.0 [10,10,20], C == 1.0
    assert np.all(golden["ao2mo_bench_R"] == 200.0)
    out = O.ri_ao2mo_ref(np.ones((10, 10), order="F"), np.full((10, 10, 20), 2.0, order="F"))
    assert out.shape == (20, 10, 10) and np.all(out == 200.0)


def test_ao2mo_identity_is_permutation():
    A = O.gen_ri3ao(9, 4, seed=3)
    out = O.ri_ao2mo_ref(np.eye(9, order="F"), A)
    for P in range(4):
        assert np.array_equal(out[P], A[:, :, P])
    # flat layout: element (P,i,j) at P + naux*i + naux*ns*j
    flat = out.reshape(-1, order="F")
    assert flat[2 + 4 * 5 + 4 * 9 * 7] == A[5, 7, 2]


def test_ao2mo_brute_force():
    rng = np.random.default_rng(1)
    nbas, ns, naux = 5, 3, 3
    A = rng.standard_normal((nbas, nbas, naux))
    Cm = rng.standard_normal((nbas, ns))
    out = O.ri_ao2mo_ref(Cm, A)
    for P in range(naux):
        for i in range(ns):
            for j in range(ns):
                ref = math.fsum(Cm[mu, i] * A[mu, nu, P] * Cm[nu, j]
                                for mu in range(nbas) for nu in range(nbas))
                assert abs(out[P, i, j] - ref) < 1e-13 * max(1.0, abs(ref))


def test_ao2mo_orthogonal_invariance():
    A = O.gen_ri3ao(24, 6, seed=5)
    Cm = O.gen_orthonormal(24, seed=6)
    out = O.ri_ao2mo_ref(Cm, A)
    for P in range(6):
        assert abs(np.linalg.norm(out[P]) - np.linalg.norm(A[:, :, P])) < 1e-12
        assert abs(np.trace(out[P]) - np.trace(A[:, :, P])) < 1e-12
        assert np.allclose(out[P], out[P].T, atol=1e-13)


def test_ao2mo_occ_is_occupied_rows_of_reference():
    A = O.gen_ri3ao(20, 7, seed=8)
    Cm = O.gen_orthonormal(20, seed=9)
    full = O.ri_ao2mo_ref(Cm, A)          # [naux, ns, ns]
    B = O.ao2mo_occ(Cm, A, nocc=4)        # [nocc, nmo, naux]
    assert B.shape == (4, 20, 7)
    assert relf(B, np.transpose(full, (1, 2, 0))[:4]) < 1e-13


# ---- inputs --------------------------------------------------------------------
def test_generators():
    z = O.gen_normal(200000, seed=11)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.02
    A = O.gen_ri3ao(16, 3, seed=12)
    assert np.array_equal(A, A.transpose(1, 0, 2))
    # envelope: |A[i,j]| / exp(-|i-j|/16) is a normal draw
    Q = O.gen_orthonormal(32, seed=13)
    assert np.abs(Q.T @ Q - np.eye(32)).max() < 1e-14
    G = O.gen_normal(32 * 32, seed=13).reshape((32, 32), order="F")
    R = Q.T @ G
    assert np.all(np.diag(R) > 0) and np.abs(np.tril(R, -1)).max() < 1e-12


def test_splitmix64_known_values():
    # SplitMix64 (Vigna) output sequence for state 0: first two outputs
    assert int(O.splitmix64(np.uint64(0))) == 0xE220A8397B1DCDAF
    assert int(O.splitmix64(np.uint64(0x9E3779B97F4A7C15))) == 0x6E789E6AA1B965F4


# ---- density / J / K -------------------------------------------------------------
def test_density_projector():
    Cm = O.gen_orthonormal(30, seed=21)
    D = O.density(Cm, 7)
    assert np.allclose(D @ D, D, atol=1e-13)
    assert abs(np.trace(D) - 7) < 1e-12
    i, j = 3, 11
    assert abs(D[i, j] - sum(Cm[i, m] * Cm[j, m] for m in range(7))) < 1e-15


def test_rij_brute_force_and_energy():
    A = O.gen_ri3ao(6, 4, seed=31)
    Cm = O.gen_orthonormal(6, seed=32)
    D = O.density(Cm, 2)
    gamma, J, e_j = O.rij(A, D)
    for P in range(4):
        ref = math.fsum(A[k, l, P] * D[k, l] for k in range(6) for l in range(6))
        assert abs(gamma[P] - ref) < 1e-14
    for i in range(6):
        for j in range(6):
            ref = math.fsum(A[i, j, P] * gamma[P] for P in range(4))
            assert abs(J[i, j] - ref) < 1e-14
    assert abs(e_j - 0.5 * np.sum(gamma ** 2)) < 1e-13
    assert np.allclose(J, J.T, atol=1e-15)


def test_rij_identity_reduction():
    A = O.gen_ri3ao(8, 5, seed=33)
    D = O.density(np.eye(8, order="F"), 3)     # diag(1,1,1,0,...)
    gamma, _, _ = O.rij(A, D)
    assert np.allclose(gamma, A[0, 0] + A[1, 1] + A[2, 2], atol=1e-15)


def test_rik_brute_force_and_two_routes():
    A = O.gen_ri3ao(7, 5, seed=41)
    Cm = O.gen_orthonormal(7, seed=42)
    B = O.ao2mo_occ(Cm, A, 3)
    K, e_x, e_x2 = O.rik(B)
    for i in range(7):
        for j in range(7):
            ref = math.fsum(B[m, i, P] * B[m, j, P] for m in range(3) for P in range(5))
            assert abs(K[i, j] - ref) < 1e-14
    assert abs(e_x - e_x2) < 1e-13
    assert np.allclose(K, K.T, atol=1e-15)
    assert np.linalg.eigvalsh(K).min() > -1e-12


def test_rik_single_occupied_outer_product():
    A = O.gen_ri3ao(5, 4, seed=43)
    Cm = O.gen_orthonormal(5, seed=44)
    B = O.ao2mo_occ(Cm, A, 1)
    K, _, _ = O.rik(B)
    ref = sum(np.outer(B[0, :, P], B[0, :, P]) for P in range(4))
    assert relf(K, ref) < 1e-15


# ---- packed upper ----------------------------------------------------------------
def test_upper_doc_order():
    # matrix/mod.rs:437-452: 4x4 matrix 1..16, upper order
    m = np.arange(1, 17).reshape((4, 4), order="F")
    assert list(O.upper_pack(m)) == [1, 5, 6, 9, 10, 11, 13, 14, 15, 16]


def test_upper_index_map_exhaustive():
    for n in range(0, 41):
        m = np.arange(n * n, dtype=float).reshape((n, n), order="F")
        p = O.upper_pack(m)
        assert len(p) == n * (n + 1) // 2
        for j in range(n):
            for i in range(j + 1):
                assert p[j * (j + 1) // 2 + i] == m[i, j]
        full = O.upper_unpack(p)
        up = np.triu(m)
        assert np.array_equal(full, up + np.triu(m, 1).T)
        assert O.upper_dim(len(p)) == n
    assert O.upper_unpack(np.zeros(5)) is None
    assert O.upper_dim(524800) == 1024 and O.upper_dim(524801) == -1


def test_rifull_to_matfull_symm():
    ri = O.gen_ri3ao(6, 3, seed=51)
    out = O.rifull_to_matfull_symm(ri)
    assert out.shape == (21, 3)
    assert out[5 * 6 // 2 + 2, 1] == ri[2, 5, 1]


# ---- sub-block GEMM and copies -----------------------------------------------------
def test_general_dgemm_golden(golden):
    for ci in range(int(golden["gemm_ncases"])):
        p = f"gemm{ci}_"
        ra = [tuple(r) for r in golden[p + "ra"]]
        rb = [tuple(r) for r in golden[p + "rb"]]
        rc = [tuple(r) for r in golden[p + "rc"]]
        opa, opb = (chr(int(c)) for c in golden[p + "ops"])
        alpha, beta = golden[p + "ab"]
        out = O.general_dgemm(golden[p + "A"], ra, opa, golden[p + "B"], rb, opb,
                              golden[p + "C"], rc, alpha, beta)
        assert relf(out, golden[p + "out"]) < 1e-14, ci
    # doc values (matrix_blas_lapack.rs:91-119), exact
    assert list(golden["gemm0_out"][1:3, 0:2].reshape(-1, order="F")) == [88, 101, 127, 146]
    assert list(golden["gemm1_out"][1, 0:2]) == [74, 107]
    assert list(golden["gemm2_out"][0:2, 0]) == [59, 74]


def test_copies_golden(golden):
    F, T = golden["cmm_F"], golden["cmm_T"]
    xl, yl, fxs, fys, txs, tys = golden["cmm_args"]
    assert np.array_equal(O.copy_mm(xl, yl, F, fxs, fys, T, txs, tys), golden["cmm_out"])
    F, R = golden["cr_F"], golden["cr_R"]
    for mod in range(3):
        xl, yl, fxs, fys, txs, tys, tx3 = golden[f"cmr{mod}_args"]
        assert np.array_equal(O.copy_mr(xl, yl, F, fxs, fys, R, txs, tys, tx3, mod),
                              golden[f"cmr{mod}_out"]), mod
        xl, yl, fxs, fys, fx3, txs, tys = golden[f"crm{mod}_args"]
        assert np.array_equal(O.copy_rm(xl, yl, R, fxs, fys, fx3, mod, F, txs, tys),
                              golden[f"crm{mod}_out"]), mod
    a = golden["crr_args"]
    assert np.array_equal(O.copy_rr(a[0:3], golden["crr_F"], a[3:6], R, a[6:9]),
                          golden["crr_out"])


def test_special_dgemm_identity_and_permutation():
    rng = np.random.default_rng(61)
    T = np.asfortranarray(rng.standard_normal((6, 3, 9)))
    eye = np.eye(9)
    out = O.special_dgemm_01(T, 1, 4, 2, 5, eye, 2, 5, 2, 5, 2.0, 0.5)
    assert np.allclose(out[1:5, :, 2:7], 2.5 * T[1:5, :, 2:7], rtol=0, atol=1e-15)
    assert np.array_equal(out[0], T[0]) and np.array_equal(out[:, :, :2], T[:, :, :2])
    perm = np.eye(5)[:, [4, 0, 3, 1, 2]]
    Bm = np.zeros((7, 7))
    Bm[1:6, 2:7] = perm
    out = O.special_dgemm_01(T, 0, 6, 4, 5, Bm, 1, 5, 2, 5, 1.0, 0.0)
    assert np.array_equal(out[:, :, 4:9], T[:, :, 4:9][:, :, [4, 0, 3, 1, 2]])


# ---- int8 Ozaki slicing (the reduced-precision K form) ---------------------------
def _ozaki_data(nocc=7, nmo=9, naux=11, seed=5):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((nocc, nmo, naux)) * np.exp(rng.uniform(-3, 3, (1, nmo, naux)))
    B[:, 2, :] = 0.0   # an all-zero row (exponent 0, all slices 0)
    return B


def test_ozaki_i8_slices_reconstruct():
    B = _ozaki_data()
    for gp in (1, 3, 11):
        sig, E = O.ozaki_i8_slices(B, gp)
        assert np.all(sig == np.rint(sig)) and np.abs(sig).max() <= 127
        w = np.array([254.0 ** -s for s in range(4)])
        rec = np.einsum("s,smip->mip", w, sig) / 127.0
        for g in range(E.shape[0]):
            sl = slice(g * gp, (g + 1) * gp)
            scale = np.ldexp(1.0, E[g])[None, :, None]
            err = np.abs(B[:, :, sl] - rec[:, :, sl] * scale)
            # truncation bound 2^E / 127 * 254^-3 / 2, plus FP64 rounding
            assert np.all(err <= scale * (0.5 / 127.0 * 254.0 ** -3) * (1 + 1e-6))
            mx = np.abs(B[:, :, sl]).max(axis=(0, 2))
            assert np.all((mx == 0) | ((mx < np.ldexp(1.0, E[g])) & (mx >= np.ldexp(0.5, E[g]))))


def test_ozaki_i8_k_brute_force():
    """Level sums by explicit loops over (m, P) on a tiny case."""
    B = _ozaki_data(nocc=3, nmo=4, naux=5, seed=7)
    gp = 2
    sig, E = O.ozaki_i8_slices(B, gp)
    K = np.zeros((4, 4))
    for g in range(E.shape[0]):
        for i in range(4):
            for j in range(4):
                L = [0, 0, 0, 0]
                for P in range(g * gp, min((g + 1) * gp, 5)):
                    for m in range(3):
                        for s in range(4):
                            for t in range(4 - s):
                                L[s + t] += int(sig[s, m, i, P]) * int(sig[t, m, j, P])
                v = sum(L[q] * 254.0 ** -q for q in range(4)) / 127.0 ** 2
                K[i, j] += v * 2.0 ** (E[g, i] + E[g, j])
    Ko = O.rik_ozaki_i8(B, gp)
    assert np.allclose(Ko, K, rtol=1e-14, atol=1e-14 * np.abs(K).max())


def test_ozaki_i8_k_accuracy():
    """Against the FP64 K: within 1e-8 relative Frobenius for several group
    sizes; weights 256^-q in place of 254^-q would miss the bar."""
    rng = np.random.default_rng(11)
    B = rng.standard_normal((16, 40, 64))
    K, _, _ = O.rik(B)
    for gp in (1, 4, 64):
        Ko = O.rik_ozaki_i8(B, gp)
        assert np.linalg.norm(Ko - K) / np.linalg.norm(K) < 1e-8
        assert np.array_equal(Ko, Ko.T)
    sig, E = O.ozaki_i8_slices(B, 4)
    bad = np.einsum("s,smip->mip", [256.0 ** -s for s in range(4)], sig) / 127.0
    Bb = bad * np.ldexp(1.0, np.repeat(E, 4, axis=0).T)[None, :, :64]
    Kb, _, _ = O.rik(Bb)
    assert np.linalg.norm(Kb - K) / np.linalg.norm(K) > 1e-6


# ---- level-2 / rank-k primitives against the BLAS the reference links -----------
def test_dgemv_matches_blas():
    from scipy.linalg import blas
    rng = np.random.default_rng(21)
    A = np.asfortranarray(rng.standard_normal((37, 23)))
    for trans, lx, ly in (("N", 23, 37), ("T", 37, 23)):
        for incx, incy in ((1, 1), (2, 3)):
            x = rng.standard_normal(1 + (lx - 1) * incx)
            y = rng.standard_normal(1 + (ly - 1) * incy)
            ref = blas.dgemv(0.7, A, x, beta=-1.3, y=y.copy(), incx=incx, incy=incy,
                             trans=0 if trans == "N" else 1)
            assert np.allclose(O.dgemv(A, x, y, trans, 0.7, -1.3, incx, incy), ref,
                               rtol=1e-14, atol=1e-13)
    with pytest.raises(ValueError):
        O.dgemv(A, np.zeros(22), np.zeros(37), "N", 1.0, 0.0)


def test_dsyrk_matches_blas_and_keeps_other_triangle():
    from scipy.linalg import blas
    rng = np.random.default_rng(22)
    for trans, shape in (("N", (19, 7)), ("T", (7, 19))):
        A = np.asfortranarray(rng.standard_normal(shape))
        for uplo in ("U", "L"):
            Cm = np.asfortranarray(rng.standard_normal((19, 19)))
            ref = blas.dsyrk(0.5, A, beta=2.0, c=Cm.copy(), trans=0 if trans == "N" else 1,
                             lower=1 if uplo == "L" else 0)
            out = O.dsyrk(A, Cm, uplo, trans, 0.5, 2.0)
            assert np.allclose(out, ref, rtol=1e-14, atol=1e-13)
            other = np.tril(np.ones((19, 19), bool), -1) if uplo == "U" else \
                np.triu(np.ones((19, 19), bool), 1)
            assert np.array_equal(out[other], Cm[other])


def test_einsum_ip_ip_p_and_rij_composition():
    rng = np.random.default_rng(23)
    A, B = rng.standard_normal((9, 5)), rng.standard_normal((7, 6))
    out = O.einsum_ip_ip_p(A, B)
    assert out.shape == (5,)
    assert np.allclose(out, [A[:7, p] @ B[:7, p] for p in range(5)], rtol=1e-15)
    assert np.array_equal(O.einsum_ip_ip_p(np.zeros((0, 4)), B), np.zeros(4))
    # RI-J as REST composes it on the packed ri3ao: gamma = A^T (w o d), J = A gamma
    nbas, naux = 12, 20
    A3 = O.gen_ri3ao(nbas, naux, seed=31)
    Dm = O.density(O.gen_orthonormal(nbas, seed=32), 4)
    Ap = O.rifull_to_matfull_symm(A3)
    w = 2.0 - np.eye(nbas)                           # off-diagonal pairs twice
    dw = O.upper_pack(w * Dm)
    gamma = O.dgemv(Ap, dw, np.zeros(naux), "T", 1.0, 0.0)
    Jp = O.dgemv(Ap, gamma, np.zeros(Ap.shape[0]), "N", 1.0, 0.0)
    gamma_ref, J, e_j = O.rij(A3, Dm)
    assert np.allclose(gamma, gamma_ref, rtol=1e-13)
    assert np.allclose(O.upper_unpack(Jp), J, rtol=1e-13, atol=1e-13)


# ---- FP64 aux-metric functions -------------------------------------------------
def _lib_rs_8x8():
    import json
    import os
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "matfull_inverse_and_power_8x8.json")))
    return np.array(d["data"]).reshape(d["size"], order="F")


def test_dpotrf_doc_example():
    from oracle import metric_oracle as M
    up = np.array([4.0, 12.0, 37.0, -16.0, -43.0, 98.0])
    A = O.upper_unpack(up)
    c, info = M.dpotrf(A, "U")
    assert info == 0
    assert np.array_equal(O.upper_pack(c), [2.0, 6.0, 1.0, -8.0, 5.0, 3.0])


def test_dsyev_doc_eigenvalues():
    A = O.upper_unpack(np.arange(1.0, 7.0))
    w = np.linalg.eigvalsh(A)
    assert np.allclose(w, [-1.5066326307865059, -0.05739624271478554, 11.564028873501286],
                       rtol=0, atol=1e-12)


def test_metric_functions_on_reference_test_matrix():
    from oracle import metric_oracle as M
    S = _lib_rs_8x8()
    assert np.allclose(S, S.T, atol=1e-15)
    X = M.lapack_power(S, -0.5, 1e-5)
    assert np.abs(X @ S @ X - np.eye(8)).max() < 1e-12
    assert np.abs(M.lapack_power(S, 0.5, 1e-5) @ M.lapack_power(S, 0.5, 1e-5) - S).max() < 1e-12
    for uplo in ("L", "U"):
        Mf = M.cholesky_decompose_inverse(S, uplo)
        F = Mf @ Mf.T if uplo == "L" else Mf.T @ Mf
        assert np.abs(F - np.linalg.inv(S)).max() < 1e-11
        assert np.array_equal(Mf, np.tril(Mf) if uplo == "L" else np.triu(Mf))


def test_metric_closed_forms():
    from oracle import metric_oracle as M
    D = np.diag([4.0, 9.0, 0.25])
    assert np.allclose(M.lapack_power(D, -0.5, 1e-10), np.diag([0.5, 1 / 3, 2.0]), rtol=1e-15)
    assert np.allclose(M.lapack_power(np.diag([4.0, 1e-12]), -1.0, 1e-10),
                       np.diag([0.25, 0.0]), rtol=1e-15)      # below threshold: dropped
    Mf = M.cholesky_decompose_inverse(D, "L")
    assert np.allclose(Mf, np.diag([0.5, 1 / 3, 2.0]), rtol=1e-15)
    assert M.cholesky_decompose_inverse(np.diag([1.0, -1.0]), "L") is None


# ---- RIFull transposes (ri.rs:225-290) ----------------------------------------

@pytest.mark.parametrize("size", [(3, 2, 2), (9, 3, 2), (1, 4, 3), (5, 1, 7), (4, 6, 1), (2, 3, 5)])
@pytest.mark.parametrize("order", ["jik", "jki", "kji", "ikj"])
def test_rifull_transpose_literal(size, order):
    """numpy permutation == the reference's own loops, on its test shape [3,2,2]
    (ri.rs:437-449) and others."""
    n = size[0] * size[1] * size[2]
    data = list(range(n))
    lit = O.rifull_transpose_literal(data, size, order)
    ri = np.asfortranarray(np.arange(n, dtype=float).reshape(size, order="F"))
    out = O.rifull_transpose(ri, order)
    assert out.ravel(order="F").tolist() == [float(v) for v in lit]


def test_rifull_transpose_group():
    ri = np.asfortranarray(np.random.default_rng(3).standard_normal((4, 5, 6)))
    T = O.rifull_transpose
    assert np.array_equal(T(T(ri, "jik"), "jik"), ri)
    assert np.array_equal(T(T(ri, "kji"), "kji"), ri)
    assert np.array_equal(T(T(ri, "ikj"), "ikj"), ri)
    assert np.array_equal(T(T(T(ri, "jki"), "jki"), "jki"), ri)
    assert T(ri, "jki").shape == (5, 6, 4) and T(ri, "kji").shape == (6, 5, 4)
