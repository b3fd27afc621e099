"""CPU oracle for the RI hot path — TEST INFRASTRUCTURE ONLY.

This module restates the reference's algorithms in numpy (whose BLAS is
OpenBLAS, the library the reference links) so that tests, the smoke check
and the benchmark's CPU-baseline leg can check and time the GPU path.  The
product (librest_tensors_amd.so and the rest_tensors_amd package) never
imports it.

Every function cites the reference file:line it follows.  How each one is
pinned (golden vectors from the compiled reference, closed forms, brute
force) is recorded next to it; tests/test_oracle.py holds the checks.

Pinning summary:
  ri_ao2mo_ref        pinned: golden vectors of oracle/_ref (restmatr.f90
                      compiled), bench closed form (all 200.0), C = I
                      permutation, brute-force quadruple loop, ao2mo_v01.
  ao2mo_occ           pinned: equals the occupied rows of ri_ao2mo_ref for
                      square C (which is pinned), brute force.
  density             pinned: C_occ C_occ^T projector identities (D^2 = D,
                      trace = nocc for orthonormal C), brute force.
  rij                 pinned: brute-force double loop, E_J = 1/2 sum gamma^2
                      = 1/2 sum D.J, C = I reduction.
  rik / e_x           pinned: brute force, two routes for E_x, nocc = 1
                      outer-product form, symmetry and PSD.
  upper pack/unpack   pinned: doc example of matrix/mod.rs:437-452, exhaustive
                      index map, round trip.
  general_dgemm       pinned: golden vectors of oracle/_ref and the doc values
                      of matrix_blas_lapack.rs:91-119.
  copy_mm/mr/rm/rr    pinned: golden vectors of oracle/_ref.
  dgemv / dsyrk       pinned: scipy.linalg.blas (OpenBLAS, the BLAS the
                      reference links), incl. strides and the untouched triangle;
                      RI-J composition on the packed layout equals rij.
  einsum_ip_ip_p      pinned: explicit column dots, zero-row case.
  ozaki_i8 (slices, K) no reference counterpart (restates the INT8 form);
                      pinned: reconstruction bound, brute-force level sums,
                      1e-8 against rik.
  rifull_transpose    pinned: a literal restatement of the reference's
                      chunk-transpose and index-filter loops (ri.rs:225-290) on
                      its own test shape [3,2,2] and others, and the
                      permutation group (jik twice, kji twice, jki three times
                      are the identity).
  special_dgemm_01    parity unpinned against the reference (its Fortran
                      aliases dgemm's A and C, restmatr.f90:140-150); pinned
                      against the product formula it documents instead.
"""
from __future__ import annotations

import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


# ---- seeded synthetic inputs ---------------------------------------------------
def splitmix64(x):
    """SplitMix64 finalizer on uint64 arrays (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def rng_key(seed: int) -> np.uint64:
    return splitmix64(np.uint64(seed))


def normal_from_counter(key, ctr):
    """N(0,1) from counter(s): same hash and Box-Muller as the device."""
    ctr = np.asarray(ctr, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h1 = splitmix64(key + np.uint64(2) * ctr)
        h2 = splitmix64(key + np.uint64(2) * ctr + np.uint64(1))
    u1 = ((h1 >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def gen_ri3ao(nbas: int, naux: int, seed: int, decay: float = 16.0) -> np.ndarray:
    """ri3ao[i,j,P] = ri3ao[j,i,P] = z*exp(-|i-j|/decay) (Fortran order)."""
    key = rng_key(seed)
    npair = nbas * (nbas + 1) // 2
    i = np.arange(nbas)[:, None]
    j = np.arange(nbas)[None, :]
    lo = np.minimum(i, j).astype(np.uint64)
    hi = np.maximum(i, j).astype(np.uint64)
    pidx = hi * (hi + np.uint64(1)) // np.uint64(2) + lo
    env = np.exp(-(hi - lo).astype(np.float64) / decay)
    out = np.empty((nbas, nbas, naux), order="F")
    for P in range(naux):
        ctr = np.uint64(P * npair) + pidx
        out[:, :, P] = normal_from_counter(key, ctr) * env
    return out


def gen_normal(n: int, seed: int) -> np.ndarray:
    return normal_from_counter(rng_key(seed), np.arange(n, dtype=np.uint64))


def gen_orthonormal(nbas: int, seed: int) -> np.ndarray:
    """Q of the Householder QR (LAPACK dgeqrf) of a seeded Gaussian matrix,
    columns signed so that diag(R) > 0."""
    G = gen_normal(nbas * nbas, seed).reshape((nbas, nbas), order="F")
    Q, R = np.linalg.qr(G)
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return np.asfortranarray(Q * s[None, :])


# ---- ao2mo ---------------------------------------------------------------------
def ri_ao2mo_ref(C: np.ndarray, ri3fn: np.ndarray) -> np.ndarray:
    """restmatr.f90:158-194 (ri_ao2mo_f), called by ri.rs:382-408.

    For every aux slice P: tmp = A_P * C (dgemm NN, :183-185), then
    ri3mo(P,:,:) = C^T * tmp (dgemm TN, :186-188).  Returns shape
    (naux, ns, ns), Fortran order: element (P,i,j) at P + naux*i + naux*ns*j.
    The reference is defined only for ns == nbas; this is the intended result
    for every ns <= nbas.
    """
    nbas, ns = C.shape
    naux = ri3fn.shape[2]
    out = np.zeros((naux, ns, ns), order="F")
    for P in range(naux):
        tmp = ri3fn[:, :, P] @ C
        out[P, :, :] = C.T @ tmp
    return out


def ri_ao2mo_v01(C: np.ndarray, ri3fn: np.ndarray) -> np.ndarray:
    """ri.rs:360-379 (ao2mo_v01): same math with _dgemm_nn then _dgemm_tn
    (matrix_blas_lapack.rs:1185-1237) and an element-order copy into the
    [naux, nbas, ns] output (valid for square C only, ri.rs:365,375-376)."""
    nbas, ns = C.shape
    assert ns == nbas, "ao2mo_v01 is only meaningful for square C"
    naux = ri3fn.shape[2]
    out = np.zeros((naux, nbas, ns), order="F")
    for P in range(naux):
        tmp_aux = np.einsum("ik,kj->ij", ri3fn[:, :, P], C)     # _dgemm_nn
        tmp_aux2 = np.einsum("ki,kj->ij", C, tmp_aux)           # _dgemm_tn
        out[P, :, :] = tmp_aux2
    return out


def ao2mo_occ(C: np.ndarray, ri3ao: np.ndarray, nocc: int) -> np.ndarray:
    """B[m,i,P] = sum_{mu,nu} C[mu,m] A[mu,nu,P] C[nu,i], m < nocc: the occupied
    rows of ri_ao2mo_ref with the aux index slowest ([nocc, nmo, naux])."""
    nbas, nmo = C.shape
    naux = ri3ao.shape[2]
    Co = C[:, :nocc]
    out = np.empty((nocc, nmo, naux), order="F")
    for P in range(naux):
        out[:, :, P] = (Co.T @ ri3ao[:, :, P]) @ C
    return out


def density(C: np.ndarray, nocc: int) -> np.ndarray:
    """D = C_occ C_occ^T: _dgemm_full(C_occ,'N',C_occ,'T')
    (matrix_blas_lapack.rs:180-252)."""
    Co = C[:, :nocc]
    return np.asfortranarray(Co @ Co.T)


# ---- RI-J / RI-K ---------------------------------------------------------------
def rij(ri3ao: np.ndarray, D: np.ndarray):
    """gamma_P = sum_kl A[k,l,P] D[k,l];  J = sum_P A_P gamma_P;
    E_J = 1/2 sum_ij D_ij J_ij.  (_dgemv composition, matrix_blas_lapack.rs:38-70;
    definition from the RI-J configuration of the hot-path workload.)"""
    nbas = D.shape[0]
    naux = ri3ao.shape[2]
    A2 = ri3ao.reshape((nbas * nbas, naux), order="F")
    gamma = A2.T @ D.reshape(-1, order="F")
    J = (A2 @ gamma).reshape((nbas, nbas), order="F")
    e_j = 0.5 * float(np.sum(D * J))
    return gamma, J, e_j


def rik(B: np.ndarray):
    """K[i,j] = sum_P sum_{m<nocc} B[m,i,P] B[m,j,P];
    E_x = -1/2 sum_{n<nocc} K[n,n]; second route -1/2 sum B[m,n,P]^2 (m,n<nocc)."""
    nocc, nmo, naux = B.shape
    K = np.zeros((nmo, nmo))
    for P in range(naux):
        K += B[:, :, P].T @ B[:, :, P]
    e_x = -0.5 * float(np.trace(K[:nocc, :nocc]))
    e_x2 = -0.5 * float(np.sum(B[:, :nocc, :] ** 2))
    return np.asfortranarray(K), e_x, e_x2


def ozaki_i8_slices(B: np.ndarray, gp: int):
    """Int8 Ozaki slicing of B [nocc, nmo, naux] with one exponent per row i
    and aux group of gp slices: returns sigma [4, nocc, nmo, naux] (integer
    valued float64, |sigma| <= 127) and E [ngroups, nmo] with
    B = 2^E / 127 * sum_s sigma_s 254^-s up to 2^E / 127 * 254^-3 / 2.
    No reference counterpart: it restates the reduced-precision K form's
    slicing step by step (same FP64 operations, round-half-even) so that the
    device result can be checked to the last bits."""
    nocc, nmo, naux = B.shape
    ng = (naux + gp - 1) // gp
    sig = np.zeros((4, nocc, nmo, naux))
    E = np.zeros((ng, nmo), dtype=np.int64)
    for g in range(ng):
        blk = B[:, :, g * gp:(g + 1) * gp]
        mx = np.abs(blk).max(axis=(0, 2)) if blk.size else np.zeros(nmo)
        e = np.where(mx > 0, np.frexp(mx)[1], 0).astype(np.int64)
        E[g] = e
        u = np.ldexp(blk, -e[None, :, None])
        x = u * 127.0
        for s in range(4):
            sg = np.rint(x)
            sig[s, :, :, g * gp:(g + 1) * gp] = sg
            x = (x - sg) * 254.0
    return sig, E


def rik_ozaki_i8(B: np.ndarray, gp: int) -> np.ndarray:
    """K from the int8 slices: per group, the exact level sums
    L_q = sum_{s+t=q} sigma_s^T sigma_t (q <= 3; integer products summed in
    float64, exact below 2^53) folded as 2^(E_i+E_j) / 127^2 sum_q 254^-q L_q."""
    nocc, nmo, naux = B.shape
    sig, E = ozaki_i8_slices(B, gp)
    K = np.zeros((nmo, nmo))
    for g in range(E.shape[0]):
        S = [sig[s, :, :, g * gp:(g + 1) * gp].transpose(2, 0, 1).reshape(-1, nmo)
             for s in range(4)]
        L = [sum(S[s].T @ S[q - s] for s in range(q + 1)) for q in range(4)]
        v = L[3]
        for q in (2, 1, 0):
            v = L[q] + v * (1.0 / 254.0)
        K += np.ldexp(v * (1.0 / 16129.0), E[g][:, None] + E[g][None, :])
    return np.asfortranarray(K)


# ---- level-2 / rank-k primitives ------------------------------------------------
def dgemv(A, x, y, trans, alpha, beta, incx=1, incy=1):
    """_dgemv (src/matrix/matrix_blas_lapack.rs:38-70): y = alpha op(A) x + beta y
    on strided vectors, with the reference's length check
    len = 1 + (len-1)*|inc| (it panics otherwise; here ValueError)."""
    m, n = A.shape
    t = trans.lower() != "n"
    lx, ly = (m, n) if t else (n, m)
    if len(x) != 1 + (lx - 1) * abs(incx) or len(y) != 1 + (ly - 1) * abs(incy):
        raise ValueError("dgemv shape mismatch")
    out = np.array(y, dtype=np.float64, copy=True)
    prod = (A.T if t else A) @ np.asarray(x)[::incx]
    out[::incy] = alpha * prod + beta * out[::incy]
    return out


def dsyrk(A, Cm, uplo, trans, alpha, beta):
    """_dsyrk (src/matrix/matrix_blas_lapack.rs:392-413): the uplo triangle of
    C (diagonal included) becomes alpha op(A) op(A)^T + beta C with op(A) = A
    for 'N', A^T for 'T'; the other triangle is left as it was."""
    P = A @ A.T if trans.lower() == "n" else A.T @ A
    out = np.array(Cm, dtype=np.float64, copy=True, order="F")
    n = out.shape[0]
    mask = np.triu(np.ones((n, n), bool)) if uplo.lower() == "u" else np.tril(np.ones((n, n), bool))
    out[mask] = alpha * P[mask] + beta * out[mask]
    return out


def einsum_ip_ip_p(A, B):
    """einsum "ip,ip->p" (_einsum_02_general, src/matrix/matrix_blas_lapack.rs:
    1324-1349): column dots over min(rows) entries for min(cols) columns, zeros
    when a matrix has no rows."""
    ar, ac = A.shape
    br, bc = B.shape
    npp = min(ac, bc)
    if ar == 0 or br == 0:
        return np.zeros(npp)
    r = min(ar, br)
    return np.einsum("ip,ip->p", A[:r, :npp], B[:r, :npp])


# ---- packed upper ---------------------------------------------------------------
def upper_dim(packed_len: int) -> int:
    """Dimension n with n(n+1)/2 == packed_len, else -1 (integer isqrt rather
    than the f32 sqrt of matrixupper.rs:176)."""
    if packed_len < 0:
        return -1
    import math
    n = (math.isqrt(8 * packed_len + 1) - 1) // 2
    return n if n * (n + 1) // 2 == packed_len else -1


def upper_pack(full: np.ndarray) -> np.ndarray:
    """MatrixFull::to_matrixupper (matrixfull.rs:638-646) in the order of
    MatrixUpperStepBy (matrix_trait.rs:170-217): column j, rows 0..=j."""
    n = full.shape[0]
    assert full.shape == (n, n)
    return np.concatenate([full[: j + 1, j] for j in range(n)]) if n else np.zeros(0)


def upper_unpack(packed: np.ndarray) -> np.ndarray | None:
    """MatrixUpper::to_matrixfull (matrixupper.rs:330-373): fill the upper
    triangle column by column, then mirror to the lower; None if the length is
    not triangular."""
    n = upper_dim(len(packed))
    if n < 0:
        return None
    full = np.zeros((n, n), order="F")
    idx = 0
    for j in range(n):
        full[: j + 1, j] = packed[idx: idx + j + 1]
        idx += j + 1
    for j in range(n):
        for i in range(j + 1, n):
            full[i, j] = full[j, i]
    return full


def rifull_transpose(ri: np.ndarray, order: str) -> np.ndarray:
    """RIFull::transpose_{jik,jki,kji,ikj} (ri.rs:225-290) on a column-major
    (n0, n1, n2) array: the result's axes are the named permutation."""
    axes = {"jik": (1, 0, 2), "jki": (1, 2, 0), "kji": (2, 1, 0), "ikj": (0, 2, 1)}[order]
    return np.asfortranarray(np.transpose(ri, axes))


def rifull_transpose_literal(data: list, size: tuple, order: str) -> list:
    """The reference's element order, restated step by step on a flat
    column-major list (small sizes only): transpose_jik transposes each
    n0 x n1 chunk (ri.rs:227-238); transpose_jki keeps, for each ii, the
    elements whose flat index idx has |idx - ii| % n0 == 0 (ri.rs:241-256);
    transpose_kji filters transpose_jik's output by |idx - ij| % (n0 n1) == 0
    (ri.rs:259-275); transpose_ikj gathers column jj of every chunk
    (ri.rs:278-293)."""
    i, j, k = size
    if order == "jik":
        out = []
        for c in range(k):
            chunk = data[c * i * j:(c + 1) * i * j]
            out += [chunk[a + i * b] for a in range(i) for b in range(j)]
        return out
    if order == "jki":
        out = []
        for ii in range(i):
            out += [v for idx, v in enumerate(data) if abs(idx - ii) % i == 0]
        return out
    if order == "kji":
        t = rifull_transpose_literal(data, size, "jik")
        out = []
        for ij in range(i * j):
            out += [v for idx, v in enumerate(t) if abs(idx - ij) % (i * j) == 0]
        return out
    if order == "ikj":
        out = []
        for jj in range(j):
            for c in range(k):
                chunk = data[c * i * j:(c + 1) * i * j]
                out += chunk[jj * i:(jj + 1) * i]
        return out
    raise ValueError(order)


def rifull_to_matfull_symm(ri: np.ndarray) -> np.ndarray:
    """ri.rs:297-306 with richunk_reduce_by_symm (ri.rs:330-341): each aux slice's
    upper triangle, column by column, into column P of [nao(nao+1)/2, naux]."""
    nao, _, naux = ri.shape
    out = np.empty((nao * (nao + 1) // 2, naux), order="F")
    for P in range(naux):
        out[:, P] = upper_pack(ri[:, :, P])
    return out


# ---- sub-block GEMM and block copies --------------------------------------------
def general_dgemm(A, ra, opa, B, rb, opb, Cm, rc, alpha, beta):
    """general_dgemm_f (restmatr.f90:63-107) via external_libs/mod.rs:60-83:
    C[rc] = alpha*op(A[ra])*op(B[rb]) + beta*C[rc] on 0-based row/column
    ranges (tuples (start, len)); 'N' = no transpose, otherwise transpose."""
    Ab = A[ra[0][0]: ra[0][0] + ra[0][1], ra[1][0]: ra[1][0] + ra[1][1]]
    Bb = B[rb[0][0]: rb[0][0] + rb[0][1], rb[1][0]: rb[1][0] + rb[1][1]]
    if opa != "N":
        Ab = Ab.T
    if opb != "N":
        Bb = Bb.T
    out = Cm.copy(order="F")
    blk = (slice(rc[0][0], rc[0][0] + rc[0][1]), slice(rc[1][0], rc[1][0] + rc[1][1]))
    out[blk] = alpha * (Ab @ Bb) + beta * out[blk]
    return out


def copy_mm(x_len, y_len, f, f_xs, f_ys, t, t_xs, t_ys):
    """restmatr.f90:197-212."""
    out = t.copy(order="F")
    out[t_xs: t_xs + x_len, t_ys: t_ys + y_len] = f[f_xs: f_xs + x_len, f_ys: f_ys + y_len]
    return out


def copy_mr(x_len, y_len, f, f_xs, f_ys, t, t_xs, t_ys, t_x3, mod):
    """restmatr.f90:215-238 (mod 0/1/2 fixes tensor axis 2/1/0 at t_x3)."""
    out = t.copy(order="F")
    blk = f[f_xs: f_xs + x_len, f_ys: f_ys + y_len]
    a, b = slice(t_xs, t_xs + x_len), slice(t_ys, t_ys + y_len)
    if mod == 0:
        out[a, b, t_x3] = blk
    elif mod == 1:
        out[a, t_x3, b] = blk
    elif mod == 2:
        out[t_x3, a, b] = blk
    return out


def copy_rm(x_len, y_len, f, f_xs, f_ys, f_x3, mod, t, t_xs, t_ys):
    """restmatr.f90:241-264."""
    out = t.copy(order="F")
    a, b = slice(f_xs, f_xs + x_len), slice(f_ys, f_ys + y_len)
    if mod == 0:
        blk = f[a, b, f_x3]
    elif mod == 1:
        blk = f[a, f_x3, b]
    elif mod == 2:
        blk = f[f_x3, a, b]
    else:
        return out
    out[t_xs: t_xs + x_len, t_ys: t_ys + y_len] = blk
    return out


def copy_rr(lens, f, fs, t, ts):
    """restmatr.f90:266-285."""
    out = t.copy(order="F")
    x, y, z = lens
    out[ts[0]: ts[0] + x, ts[1]: ts[1] + y, ts[2]: ts[2] + z] = \
        f[fs[0]: fs[0] + x, fs[1]: fs[1] + y, fs[2]: fs[2] + z]
    return out


def special_dgemm_01(T, sx, lx, sz, lz, Bm, srb, lrb, scb, lcb, alpha, beta):
    """Documented intent of special_dgemm_f_01 (restmatr.f90:109-154): for every
    y slice, T[x-range, y, z-range] = alpha*T[...]*B[block] + beta*T[...]."""
    out = T.copy(order="F")
    Bb = Bm[srb: srb + lrb, scb: scb + lcb]
    for y in range(T.shape[1]):
        blk = T[sx: sx + lx, y, sz: sz + lz]
        out[sx: sx + lx, y, sz: sz + lz] = alpha * (blk @ Bb) + beta * blk
    return out


def rik_syrk(B: np.ndarray, chunk: int = 64) -> np.ndarray:
    """Same K as rik(), accumulated with BLAS dsyrk (upper triangle, in place)
    over chunks of aux slices stacked as [(m,P), i] rows, then mirrored: the
    CPU-baseline form of the K build."""
    from scipy.linalg.blas import dsyrk
    nocc, nmo, naux = B.shape
    K = np.zeros((nmo, nmo), order="F")
    for p0 in range(0, naux, chunk):
        blk = B[:, :, p0:p0 + chunk]
        X = np.asfortranarray(blk.transpose(0, 2, 1).reshape(-1, nmo, order="F"))
        K = dsyrk(1.0, X, beta=1.0, c=K, trans=1, lower=0, overwrite_c=1)
    iu = np.triu_indices(nmo, 1)
    K.T[iu] = K[iu]
    return K
