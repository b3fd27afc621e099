"""Writes tests/golden/ref_vectors.npz from the reference's own compiled code.

Runs only in a container that has the reference sources: `make -C oracle`
first builds oracle/_ref/librestmatr_ref.so from
src/external_libs/restmatr.f90 (compiled where it lies).  This script calls
its seven symbols through the same by-reference ABI the Rust crate binds
(src/external_libs/ffi_restmatr.rs:4-61) on small seeded inputs and stores
inputs and outputs as plain arrays (no pickles).

ri_ao2mo_f is called only with square C (num_states == num_basis): for
num_states < num_basis the Fortran writes past its ns x ns section temp
(restmatr.f90:186-188) and its output is undefined.  special_dgemm_f_01 is
not called: it aliases dgemm's A and C operands (restmatr.f90:140-150).
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(REPO, "oracle", "_ref", "librestmatr_ref.so")
OUT = os.path.join(HERE, "ref_vectors.npz")


def ip(v):
    return C.byref(C.c_int(int(v)))


def dp(a):
    return a.ctypes.data_as(C.c_void_p)


def main():
    lib = C.CDLL(REF_LIB)
    rng = np.random.default_rng(20261202)
    out = {}

    # --- ri_ao2mo_f ------------------------------------------------------------
    def ao2mo(Cm, A):
        nbas, ns = Cm.shape
        naux = A.shape[2]
        R = np.zeros((naux, ns, ns), order="F")
        lib.ri_ao2mo_f_(dp(Cm), dp(A), dp(R), ip(ns), ip(nbas), ip(naux))
        return R

    for name, nbas, naux in [("ao2mo_a", 7, 5), ("ao2mo_b", 16, 24), ("ao2mo_c", 33, 9)]:
        A = rng.standard_normal((nbas, nbas, naux))
        A = np.asfortranarray(A + A.transpose(1, 0, 2))  # symmetric like (mu nu|P)
        Cm = np.asfortranarray(rng.standard_normal((nbas, nbas)))
        out[f"{name}_C"] = Cm
        out[f"{name}_A"] = A
        out[f"{name}_R"] = ao2mo(Cm, A)
    # general (non-symmetric) ri3fn too
    A = np.asfortranarray(rng.standard_normal((12, 12, 7)))
    Cm = np.asfortranarray(rng.standard_normal((12, 12)))
    out["ao2mo_ns_C"], out["ao2mo_ns_A"], out["ao2mo_ns_R"] = Cm, A, ao2mo(Cm, A)
    # bench case: ri3fn == 2.0 [10,10,20], C == 1.0 (benches/bench_tensors.rs:5-10)
    A = np.full((10, 10, 20), 2.0, order="F")
    Cm = np.ones((10, 10), order="F")
    out["ao2mo_bench_R"] = ao2mo(Cm, A)

    # --- general_dgemm_f ---------------------------------------------------------
    def gdgemm(A, ra, opa, B, rb, opb, Cm, rc, alpha, beta):
        Cm = Cm.copy(order="F")
        lib.general_dgemm_f_(
            dp(A), ip(A.shape[0]), ip(A.shape[1]), ip(ra[0][0]), ip(ra[0][1]),
            ip(ra[1][0]), ip(ra[1][1]), C.c_char_p(opa.encode()),
            dp(B), ip(B.shape[0]), ip(B.shape[1]), ip(rb[0][0]), ip(rb[0][1]),
            ip(rb[1][0]), ip(rb[1][1]), C.c_char_p(opb.encode()),
            dp(Cm), ip(Cm.shape[0]), ip(Cm.shape[1]), ip(rc[0][0]), ip(rc[0][1]),
            ip(rc[1][0]), ip(rc[1][1]),
            C.byref(C.c_double(alpha)), C.byref(C.c_double(beta)))
        return Cm

    cases = []
    a3 = np.asfortranarray(np.arange(1, 10, dtype=float).reshape(3, 3, order="F"))
    b3 = np.asfortranarray(np.arange(6, 15, dtype=float).reshape(3, 3, order="F"))
    c3 = np.full((3, 3), 2.0, order="F")
    # the three examples of matrix_blas_lapack.rs:91-119
    cases.append((a3, ((1, 2), (1, 2)), "N", b3, ((0, 2), (0, 2)), "N", c3, ((1, 2), (0, 2)), 1.0, 1.0))
    cases.append((a3, ((1, 2), (1, 1)), "T", b3, ((0, 2), (0, 2)), "N", c3, ((1, 1), (0, 2)), 1.0, 1.0))
    cases.append((a3, ((1, 2), (0, 2)), "N", b3, ((0, 1), (0, 2)), "T", c3, ((0, 2), (0, 1)), 1.0, 1.0))
    for opa in "NT":
        for opb in "NT":
            m, n, k = rng.integers(3, 20, size=3)
            ra_rows, ra_cols = (m, k) if opa == "N" else (k, m)
            rb_rows, rb_cols = (k, n) if opb == "N" else (n, k)
            A = np.asfortranarray(rng.standard_normal((ra_rows + 5, ra_cols + 4)))
            B = np.asfortranarray(rng.standard_normal((rb_rows + 3, rb_cols + 6)))
            Cm = np.asfortranarray(rng.standard_normal((m + 7, n + 2)))
            ra = ((2, ra_rows), (1, ra_cols))
            rb = ((1, rb_rows), (3, rb_cols))
            rc = ((4, m), (1, n))
            cases.append((A, ra, opa, B, rb, opb, Cm, rc, 0.75, -1.25))
    for ci, (A, ra, opa, B, rb, opb, Cm, rc, alpha, beta) in enumerate(cases):
        p = f"gemm{ci}_"
        out[p + "A"], out[p + "B"], out[p + "C"] = A, B, Cm
        out[p + "ra"] = np.array(ra, dtype=np.int64)
        out[p + "rb"] = np.array(rb, dtype=np.int64)
        out[p + "rc"] = np.array(rc, dtype=np.int64)
        out[p + "ops"] = np.array([ord(opa), ord(opb)], dtype=np.int64)
        out[p + "ab"] = np.array([alpha, beta])
        out[p + "out"] = gdgemm(A, ra, opa, B, rb, opb, Cm, rc, alpha, beta)
    out["gemm_ncases"] = np.array(len(cases))

    # --- copy_mm / copy_mr / copy_rm / copy_rr -------------------------------------
    F = np.asfortranarray(rng.standard_normal((9, 8)))
    T = np.asfortranarray(rng.standard_normal((7, 10)))
    Tc = T.copy(order="F")
    lib.copy_mm_(ip(4), ip(3), dp(F), ip(9), ip(8), ip(2), ip(5),
                 dp(Tc), ip(7), ip(10), ip(3), ip(6))
    out["cmm_F"], out["cmm_T"], out["cmm_out"] = F, T, Tc
    out["cmm_args"] = np.array([4, 3, 2, 5, 3, 6], dtype=np.int64)

    R = np.asfortranarray(rng.standard_normal((6, 7, 8)))
    out["cr_F"], out["cr_R"] = F, R
    # (x_len, y_len, f_xs, f_ys, t_xs, t_ys, t_x3) per mod, all in range
    mr_args = {0: (4, 3, 2, 5, 1, 2, 6), 1: (4, 3, 2, 5, 1, 4, 5), 2: (3, 4, 1, 2, 2, 3, 4)}
    for mod, (xl, yl, fxs, fys, txs, tys, tx3) in mr_args.items():
        Rc = R.copy(order="F")
        lib.copy_mr_(ip(xl), ip(yl), dp(F), ip(9), ip(8), ip(fxs), ip(fys),
                     dp(Rc), ip(6), ip(7), ip(8), ip(txs), ip(tys), ip(tx3), ip(mod))
        out[f"cmr{mod}_args"] = np.array([xl, yl, fxs, fys, txs, tys, tx3], dtype=np.int64)
        out[f"cmr{mod}_out"] = Rc
    rm_args = {0: (4, 3, 1, 2, 6, 2, 5), 1: (4, 3, 1, 4, 5, 2, 5), 2: (3, 4, 2, 3, 4, 1, 2)}
    for mod, (xl, yl, fxs, fys, fx3, txs, tys) in rm_args.items():
        Fc = F.copy(order="F")
        lib.copy_rm_(ip(xl), ip(yl), dp(R), ip(6), ip(7), ip(8), ip(fxs), ip(fys),
                     ip(fx3), ip(mod), dp(Fc), ip(9), ip(8), ip(txs), ip(tys))
        out[f"crm{mod}_args"] = np.array([xl, yl, fxs, fys, fx3, txs, tys], dtype=np.int64)
        out[f"crm{mod}_out"] = Fc
    R2 = np.asfortranarray(rng.standard_normal((5, 9, 4)))
    Rc = R.copy(order="F")
    lib.copy_rr_(ip(3), ip(4), ip(2), dp(R2), ip(5), ip(9), ip(4), ip(1), ip(3), ip(1),
                 dp(Rc), ip(6), ip(7), ip(8), ip(2), ip(1), ip(5))
    out["crr_F"], out["crr_out"] = R2, Rc
    out["crr_args"] = np.array([3, 4, 2, 1, 3, 1, 2, 1, 5], dtype=np.int64)

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    if not os.path.exists(REF_LIB):
        sys.exit(f"{REF_LIB} missing: run `make -C oracle` first")
    main()
