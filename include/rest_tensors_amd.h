/*
 * rest_tensors_amd.h — extended C ABI of librest_tensors_amd.so.
 *
 * The reference has no status channel, 32-bit dimensions and host-only
 * buffers (src/external_libs/mod.rs:6-190).  This API adds, for the RI hot
 * path: int status codes with a thread-local message (rt_last_error), 64-bit
 * sizes (ri3ao at nbas=1024, naux=4096 holds 2^32 doubles), device-resident
 * buffers, explicit streams and a precision switch.
 *
 * All "_dev" entry points take device pointers and enqueue work on `stream`
 * (NULL = the null stream); they return once the work is enqueued unless the
 * comment says otherwise.  Layouts are column-major like the reference's
 * MatrixFull / RIFull (ri.rs:18-32: element (i,j,k) at i + n0*j + n0*n1*k).
 */
#ifndef REST_TENSORS_AMD_H
#define REST_TENSORS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RT_OK 0
#define RT_ERR_ARG 1      /* bad shape / argument (the reference panics)   */
#define RT_ERR_HIP 2      /* HIP runtime failure                           */
#define RT_ERR_NOMEM 3    /* device allocation failed                      */
#define RT_ERR_NOTCONV 4  /* iterative solver did not converge             */
#define RT_ERR_NUMERIC 5  /* numerical breakdown (e.g. linear dependence)  */

/* arithmetic of the GEMM-shaped stages (ao2mo, RI-K) */
#define RT_PREC_FP64 0        /* FP64 MFMA                                 */
#define RT_PREC_SPLIT_FP32 1  /* hi/lo FP32 split, 3 FP32-MFMA products,
                                 FP64 fold per 32-deep k-chunk             */
#define RT_PREC_BF16X3 2      /* hi/lo BF16 split, 3 BF16-MFMA products    */
#define RT_PREC_OZAKI_BF16 3  /* exponent-aligned BF16 slices, exact FP32
                                 partial sums, FP64 accumulation           */
#define RT_PREC_OZAKI_INT8 4  /* Ozaki-scheme INT8 slices (4 per value,
                                 exponent per row and aux group), exact
                                 int32 level sums, FP64 folding            */

/* ---- runtime plumbing -------------------------------------------------- */
const char* rt_last_error(void);
/* Empties this thread's last-error message (so that a caller of the void
 * drop-in symbols, which have no status channel, can tell whether a call
 * failed: the message is non-empty afterwards). */
void rt_clear_error(void);
int rt_version(void);
int rt_device_count(int* count);
int rt_set_device(int device);
int rt_device_name(char* buf, int buflen);
int rt_malloc(void** dptr, size_t bytes);
int rt_free(void* dptr);
int rt_memcpy_h2d(void* dst, const void* src, size_t bytes);
int rt_memcpy_d2h(void* dst, const void* src, size_t bytes);
int rt_memcpy_d2d(void* dst, const void* src, size_t bytes);
int rt_memset(void* dst, int value, size_t bytes);
int rt_synchronize(void);
int rt_stream_create(void** stream);
int rt_stream_destroy(void* stream);
int rt_stream_synchronize(void* stream);
int rt_event_create(void** event);
int rt_event_destroy(void* event);
int rt_event_record(void* event, void* stream);
int rt_event_elapsed_ms(void* start, void* stop, float* ms);

/* Per-kernel timing: when enabled, the library brackets each launch of its
 * named kernels with HIP events on the launch stream; rt_kernel_timer_get
 * synchronizes and returns the summed duration and launch count of every
 * kernel whose name starts with `prefix`. */
int rt_kernel_timer_enable(int on);
int rt_kernel_timer_reset(void);
int rt_kernel_timer_get(const char* prefix, double* total_ms, int64_t* count);

/* ---- seeded synthetic inputs ------------------------------------------ */
/* ri3ao[i,j,P] = ri3ao[j,i,P] = z * exp(-|i-j|/decay), z ~ N(0,1) from a
 * SplitMix64 counter hash of (seed, P, min(i,j), max(i,j)) and Box-Muller. */
int rt_gen_ri3ao_dev(double* ri3ao, int64_t nbas, int64_t naux, uint64_t seed,
                     double decay, void* stream);
/* out[k] = N(0,1) draw for counter k, same hash as above (for C, eps, ...) */
int rt_gen_normal_dev(double* out, int64_t n, uint64_t seed, void* stream);

/* ---- dense FP64 GEMM on device (A6: _dgemm / general_dgemm_f) ---------- */
/* C = alpha*op(A)*op(B) + beta*C, column-major, op in {'N','T'}.
 * Replaces _dgemm_full (matrix_blas_lapack.rs:180-252). */
int rt_dgemm_dev(char opa, char opb, int64_t m, int64_t n, int64_t k,
                 double alpha, const double* A, int64_t lda, const double* B,
                 int64_t ldb, double beta, double* C, int64_t ldc,
                 void* stream);
/* batched: batch b uses A + b*stride_a, B + b*stride_b, C + b*stride_c */
int rt_dgemm_strided_batched_dev(char opa, char opb, int64_t m, int64_t n,
                                 int64_t k, double alpha, const double* A,
                                 int64_t lda, int64_t stride_a,
                                 const double* B, int64_t ldb,
                                 int64_t stride_b, double beta, double* C,
                                 int64_t ldc, int64_t stride_c, int64_t batch,
                                 void* stream);
/* Sub-block GEMM with the reference's argument meaning (general_dgemm_f,
 * external_libs/mod.rs:60-83): blocks given by 0-based starts and lengths in
 * full column-major matrices of the stated sizes; shape mismatch -> RT_ERR_ARG
 * (the reference's _dgemm panics, matrix_blas_lapack.rs:132-164). */
int rt_general_dgemm_dev(const double* A, int64_t rows_a, int64_t cols_a,
                         int64_t start_row_a, int64_t len_row_a,
                         int64_t start_col_a, int64_t len_col_a, char opa,
                         const double* B, int64_t rows_b, int64_t cols_b,
                         int64_t start_row_b, int64_t len_row_b,
                         int64_t start_col_b, int64_t len_col_b, char opb,
                         double* C, int64_t rows_c, int64_t cols_c,
                         int64_t start_row_c, int64_t len_row_c,
                         int64_t start_col_c, int64_t len_col_c, double alpha,
                         double beta, void* stream);

/* ---- RIFull ao2mo (A3/A4) ---------------------------------------------- */
/* Reference layout: ri3mo[P + naux*i + naux*ns*j] = (C^T A_P C)[i,j].
 * ri3ao [nbas,nbas,naux], C [nbas,ns].  Replaces RIFull::ao2mo
 * (ri.rs:356-358, 382-408) -> ri_ao2mo_f (restmatr.f90:158-194). */
int rt_ao2mo_dev(const double* C, const double* ri3ao, double* ri3mo,
                 int64_t ns, int64_t nbas, int64_t naux, void* stream);
/* Occupied-row transform: B[m + nocc*i + nocc*nmo*P] =
 *   sum_{mu,nu} C[mu,m] ri3ao[mu,nu,P] C[nu,i],  m < nocc, i < nmo.
 * C [nbas,nmo]; the first nocc columns are the occupied orbitals.
 * precision: RT_PREC_*.  The half-transformed intermediate lives in library
 * scratch that grows on demand and is kept until the process ends. */
int rt_ao2mo_occ_dev(const double* C, int64_t nbas, int64_t nmo, int64_t nocc,
                     const double* ri3ao, int64_t naux, double* B,
                     int precision, void* stream);

/* ---- density, RI-J, RI-K (A6, A13, A14) -------------------------------- */
/* D = C_occ C_occ^T, C_occ = first nocc columns of C [nbas, *] (ld = nbas);
 * _dgemm_full(C_occ,'N',C_occ,'T') of matrix_blas_lapack.rs:180-252. */
int rt_density_dev(const double* C, int64_t nbas, int64_t nocc, double* D,
                   void* stream);
/* gamma_P = sum_{kl} ri3ao[k,l,P] D[k,l];  J[i,j] = sum_P ri3ao[i,j,P] gamma_P;
 * e_j (host pointer, may be NULL) = 1/2 sum_ij D_ij J_ij = 1/2 sum_P gamma_P^2.
 * ri3ao must be symmetric in its first two indices (it is (mu nu|P)); only
 * its upper triangle is read.  gamma: device [naux] (may be NULL).
 * Synchronizes `stream` when e_j != NULL. */
int rt_rij_dev(const double* ri3ao, int64_t nbas, int64_t naux,
               const double* D, double* J, double* gamma, double* e_j,
               void* stream);
/* K[i,j] = sum_P sum_{m<nocc} B[m,i,P] B[m,j,P]  (B from rt_ao2mo_occ_dev);
 * e_x (host, may be NULL) = -1/2 sum_{n<nocc} K[n,n].  K is [nmo,nmo] full
 * (upper computed, mirrored).  precision: RT_PREC_*.
 * Synchronizes `stream` when e_x != NULL. */
int rt_rik_dev(const double* B, int64_t nocc, int64_t nmo, int64_t naux,
               double* K, double* e_x, int precision, void* stream);
/* Aux slices that share one exponent per row in RT_PREC_OZAKI_INT8 on the
 * current device (the slicing is reproducible from this number). */
int64_t rt_ozaki_i8_group_size(int64_t nocc, int64_t nmo, int64_t naux);
/* E_x by the second route: -1/2 sum_P sum_{m,n<nocc} B[m,n,P]^2, FP64
 * accumulation (host result; synchronizes). */
int rt_exchange_energy_dev(const double* B, int64_t nocc, int64_t nmo,
                           int64_t naux, double* e_x, void* stream);

/* ---- level-2 / rank-k primitives REST composes on RI tensors (A7) ------- */
/* y = alpha op(A) x + beta y, A [m, n] column-major (lda >= m); op 'N' or
 * 'T'/'C'; positive increments.  Replaces _dgemv
 * (src/matrix/matrix_blas_lapack.rs:38-70); e.g. gamma = A^T d and J = A gamma
 * on the packed ri3ao of rt_rifull_to_matfull_symm_dev. */
int rt_dgemv_dev(char trans, int64_t m, int64_t n, double alpha, const double* A,
                 int64_t lda, const double* x, int64_t incx, double beta, double* y,
                 int64_t incy, void* stream);
/* C = alpha A A^T + beta C ('N', A [n, k]) or alpha A^T A + beta C ('T',
 * A [k, n]); only the uplo ('U'/'L') triangle of C is read and written.
 * Replaces _dsyrk (src/matrix/matrix_blas_lapack.rs:392-413). */
int rt_dsyrk_dev(char uplo, char trans, int64_t n, int64_t k, double alpha,
                 const double* A, int64_t lda, double beta, double* C, int64_t ldc,
                 void* stream);
/* out[p] = sum_i A[i, p] B[i, p] for p < min(a_cols, b_cols), i <
 * min(a_rows, b_rows) (zeros when a row count is 0).  Replaces the einsum
 * "ip,ip->p" of _einsum_02_* (src/matrix/matrix_blas_lapack.rs:1293-1349). */
int rt_einsum_ip_ip_p_dev(const double* A, int64_t a_rows, int64_t a_cols,
                          const double* B, int64_t b_rows, int64_t b_cols,
                          double* out, void* stream);

/* ---- FP64 aux-metric functions (A11) ------------------------------------ */
/* Cholesky factorization of the SPD matrix whose uplo ('U'/'L') triangle A
 * holds: A = U^T U or L L^T, the factor written over that triangle, the other
 * strict triangle untouched; *info (host) = 0, or k > 0 when the leading
 * minor of order k is not positive definite (the reference panics then).
 * Replaces _dpotrf (src/matrix/matrix_blas_lapack.rs:518-534).  Synchronizes. */
int rt_dpotrf_dev(char uplo, int64_t n, double* A, int64_t lda, int* info,
                  void* stream);
/* out = the Cholesky factor of A^-1 (lower for 'L': A^-1 = M M^T, upper for
 * 'U': A^-1 = U^T U), the other strict triangle zero; A read from its uplo
 * triangle.  RT_ERR_NUMERIC when A is not positive definite (the reference
 * returns None).  Replaces MatrixFull::cholesky_decompose_inverse
 * (src/matrix/matrix_blas_lapack.rs:931-1002).  Synchronizes. */
int rt_cholesky_decompose_inverse_dev(char uplo, int64_t n, const double* A,
                                      int64_t lda, double* out, int64_t ldo,
                                      void* stream);
/* out = A^p in the reference's sense (MatrixFull::lapack_power,
 * src/matrix/matrix_blas_lapack.rs:1004-1064; _power, :599-653): the lower
 * triangle of A is eigen-decomposed (rt_dsyev_dev on -A, so eigenvalues come
 * largest first), eigenvalues below `threshold` are dropped, and
 * out = sum_{lambda >= threshold} sqrt(lambda)^p sqrt(lambda)^p u u^T.
 * Any real p.  Synchronizes. */
int rt_matrix_power_dev(int64_t n, const double* A, int64_t lda, double p,
                        double threshold, double* out, int64_t ldo, void* stream);
/* All eigenvalues (ascending, w: n doubles on the device) and, for
 * jobz 'V', the orthonormal eigenvectors (overwriting A's columns) of the
 * symmetric matrix whose lower triangle A holds: LAPACK DSYEV semantics with
 * uplo 'L', as _dsyev calls it (src/matrix/matrix_blas_lapack.rs:319-348).
 * info as LAPACK: 0, or the number of off-diagonal entries of the
 * tridiagonal form that did not converge (the reference panics on it).
 * Tridiagonal reduction, rotation application and back-transformation run
 * on the device; the QL/QR iteration on the tridiagonal runs on the host.
 * Synchronizes. */
int rt_dsyev_dev(char jobz, int64_t n, double* A, int64_t lda, double* w, int* info,
                 void* stream);
/* Generalized eigenproblem A x = lambda B x (itype 1) on packed-upper A, B
 * (j(j+1)/2 + i): the lowest num_orb eigenpairs, Z [n, num_orb] with
 * Z^T B Z = I, w[n] ascending with w[num_orb..n) = 0 (the reference's w).
 * info: 0, n + k when B's leading minor of order k is not positive definite
 * (the reference panics), else the eigensolver's.  Replaces _dspgvx /
 * lapack_dspgvx (src/matrix/matrix_blas_lapack.rs:2123-2186, 1096-1147).
 * Synchronizes. */
int rt_dspgvx_dev(int64_t n, const double* Ap, const double* Bp, int64_t num_orb, double* Z,
                  double* w, int* info, void* stream);
/* All eigenpairs of a packed-upper symmetric matrix: Z [n, n], w[n]
 * ascending, *n_found = n.  Replaces lapack_dspevx
 * (src/matrix/matrix_blas_lapack.rs:1075-1095).  Synchronizes. */
int rt_dspevx_dev(int64_t n, const double* Ap, double* Z, double* w, int* n_found,
                  void* stream);
/* Packed RHF Fock matrix F = H + J - K/2 (upper triangle) from full n x n
 * H, J, K.  No counterpart in the reference crate (the SCF driver that
 * assembles it lives outside); it feeds rt_dspgvx_dev. */
int rt_fock_rhf_packed_dev(int64_t n, const double* H, const double* J, const double* K,
                           double* F, void* stream);
/* Host-only: LAPACK DSTEQR on host arrays (compz 'N', 'I' or 'V'; d[n],
 * e[n-1], z[n x n] with ld ldz), the tridiagonal stage of rt_dsyev_dev. */
int rt_dsteqr_host(char compz, int64_t n, double* d, double* e, double* z, int64_t ldz,
                   int* info);

/* ---- packed-upper MatrixUpper (A8/A9) ---------------------------------- */
/* Packed index of (i,j), i<=j: j*(j+1)/2 + i (index.rs:209-233).
 * unpack: full[n,n] from packed[len]; len must be triangular (else
 * RT_ERR_ARG, the reference returns None, matrixupper.rs:330-373). */
int rt_upper_unpack_dev(const double* packed, int64_t packed_len,
                        double* full, void* stream);
/* pack: packed[n(n+1)/2] from the upper triangle of full[n,n]
 * (MatrixFull::to_matrixupper, matrixfull.rs:638-646). */
int rt_upper_pack_dev(const double* full, int64_t n, double* packed,
                      void* stream);
/* RIFull [nao,nao,naux] -> MatrixFull [nao(nao+1)/2, naux] upper-triangle
 * pack per aux slice (rifull_to_matfull_symm, ri.rs:297-306, 330-341). */
int rt_rifull_to_matfull_symm_dev(const double* ri, int64_t nao, int64_t naux,
                                  double* out, void* stream);
/* special_dgemm_f_01 (restmatr.f90:111-154; the O_V V^-1/2 step of the ri3fn
 * build) on device buffers, with the drop-in symbol's argument meaning: for
 * every y, ten3_a[x-range, y, z-range] = alpha * ten3_a[x-range, y, z-range]
 * * B[row-range, col-range] + beta * ten3_a[...], in place; the B block is
 * len_z x len_z.  0-based starts. */
int rt_special_dgemm_dev(double* ten3_a, int64_t x_a, int64_t y_a, int64_t z_a,
                         int64_t start_x_a, int64_t len_x_a, int64_t start_z_a, int64_t len_z_a,
                         const double* matr_b, int64_t rows_b, int64_t columns_b,
                         int64_t start_row_b, int64_t len_row_b, int64_t start_column_b,
                         int64_t len_column_b, double alpha, double beta, void* stream);
/* Index permutations of a column-major [n0, n1, n2] RIFull (ri.rs:225-290):
 * order "jik" -> [n1, n0, n2] (transpose_jik), "jki" -> [n1, n2, n0]
 * (transpose_jki), "kji" -> [n2, n1, n0] (transpose_kji), "ikj" ->
 * [n0, n2, n1] (transpose_ikj).  Bit-exact copies; src and out must not
 * overlap. */
int rt_rifull_transpose_dev(const double* src, int64_t n0, int64_t n1, int64_t n2,
                            const char* order, double* out, void* stream);
/* integer dimension of a packed length (exact isqrt; -1 if not triangular) */
int64_t rt_upper_dim(int64_t packed_len);

/* ---- device block movers (copy_mm/mr/rm/rr on device buffers) ---------- */
int rt_copy_mm_dev(int64_t x_len, int64_t y_len, const double* f,
                   int64_t f_x, int64_t f_y, int64_t f_xs, int64_t f_ys,
                   double* t, int64_t t_x, int64_t t_y, int64_t t_xs,
                   int64_t t_ys, void* stream);
int rt_copy_mr_dev(int64_t x_len, int64_t y_len, const double* f,
                   int64_t f_x, int64_t f_y, int64_t f_xs, int64_t f_ys,
                   double* t, int64_t t_x, int64_t t_y, int64_t t_z,
                   int64_t t_xs, int64_t t_ys, int64_t t_x3, int mod,
                   void* stream);
int rt_copy_rm_dev(int64_t x_len, int64_t y_len, const double* f,
                   int64_t f_x, int64_t f_y, int64_t f_z, int64_t f_xs,
                   int64_t f_ys, int64_t f_x3, int mod, double* t,
                   int64_t t_x, int64_t t_y, int64_t t_xs, int64_t t_ys,
                   void* stream);
int rt_copy_rr_dev(int64_t x_len, int64_t y_len, int64_t z_len,
                   const double* f, int64_t f_x, int64_t f_y, int64_t f_z,
                   int64_t f_xs, int64_t f_ys, int64_t f_zs, double* t,
                   int64_t t_x, int64_t t_y, int64_t t_z, int64_t t_xs,
                   int64_t t_ys, int64_t t_zs, void* stream);

/* ---- Davidson eigensolver (A12) ------------------------------------------ */
/* Replaces davidson_solve (src/davidson.rs:48-266).  The reference's closure
 * a_x(&Vec<f64>) -> Vec<f64> becomes a block operator on device vectors:
 * Y[:, c] = A X[:, c] for c < k (column-major, leading dim n).  Return RT_OK
 * on success. */
typedef int (*rt_matvec_fn)(void* ctx, const double* X, double* Y, int64_t n,
                            int64_t k, void* stream);

#define RT_PRECOND_DIAG_REF 0  /* r / (hdiag - e[0]) for every root (davidson.rs:224) */
#define RT_PRECOND_DIAG_ROOT 1 /* r / (hdiag - e[k]) */
#define RT_PRECOND_MLP 2       /* learned MLP correction (rt_mlp_*) */
#define RT_PRECOND_AUX 3       /* the auxiliary operator's output (params.aux_op) */

/* Optional recorder of (residual, Ritz vector, e_k, e_0, root, cycle) per
 * root per cycle, the training data of the learned preconditioner.  With
 * xs != NULL it also keeps a copy of the subspace each cycle's corrections
 * are orthogonalized against (while xs_cap allows): sample q's subspace is
 * the xs_cnt[q] vectors at xs + xs_off[q] (xs_off[q] = -1: not kept). */
typedef struct rt_davidson_record {
  double* r;     /* device, n * cap */
  double* x;     /* device, n * cap */
  double* e_k;   /* host, cap */
  double* e_0;   /* host, cap */
  int* root;     /* host, cap */
  int* iter;     /* host, cap */
  int64_t cap;
  int64_t count; /* in/out: samples recorded so far */
  double* xs;       /* device store of subspace copies, or NULL */
  int64_t xs_cap;   /* doubles */
  int64_t xs_used;  /* in/out: doubles used */
  int64_t* xs_off;  /* host, cap */
  int* xs_cnt;      /* host, cap */
} rt_davidson_record;

/* DavidsonParams (davidson.rs:28-46) plus the preconditioner choice. */
typedef struct rt_davidson_params {
  double tol;      /* |de| < tol and |r| < sqrt(tol) (davidson.rs:189) */
  int maxcyc;
  int maxspace;    /* grown by 3*(nroots-1) as in davidson.rs:74 */
  double lindep;
  int nroots;
  int precond;     /* RT_PRECOND_* */
  const void* mlp; /* network for RT_PRECOND_MLP */
  rt_davidson_record* record; /* may be NULL */
  int verbose;
  /* auxiliary correction operator S = aux_op(aux_ctx, R) (e.g. a shift-invert
   * of A): the correction itself for RT_PRECOND_AUX, an extra basis vector of
   * the MLP for RT_PRECOND_MLP.  May be NULL. */
  rt_matvec_fn aux_op;
  void* aux_ctx;
} rt_davidson_params;

typedef struct rt_davidson_stats {
  int iterations;
  int matvecs;
  int restarts;
  int nroots_found;
  int mlp_fallbacks;
  double seconds;      /* wall time of the solve */
  double max_residual; /* max |r| of the last cycle */
} rt_davidson_stats;

/* x0: device initial guess (one vector, as davidson.rs:73); hdiag: device
 * diagonal of A; conv, e: host [nroots]; x_out: device n x nroots (may be
 * NULL).  Returns RT_ERR_NOTCONV if maxcyc is reached (outputs still set). */
int rt_davidson_solve_dev(rt_matvec_fn a_x, void* ctx, const double* x0,
                          const double* hdiag, int64_t n,
                          const rt_davidson_params* params, int* conv,
                          double* e, double* x_out, rt_davidson_stats* stats,
                          void* stream);

/* ---- RI linear-response operator of the Davidson workload (A15) --------- */
/* A = diag(d) + 2 scale^2 B B^T, B: device n x naux block with leading dim
 * ldb (e.g. the occ-vir block of ri3mo(occ,all,aux): B = ri3mo + nocc^2,
 * ldb = nocc*nmo).  The handle does not own B or d. */
int rt_casida_op_create(const double* B, int64_t ldb, int64_t n, int64_t naux,
                        double scale, const double* d, void** op);
int rt_casida_op_destroy(void* op);
int rt_casida_matvec(void* op, const double* X, double* Y, int64_t n,
                     int64_t k, void* stream); /* an rt_matvec_fn */
int rt_casida_hdiag(void* op, double* hdiag, void* stream);
/* Shift-invert of the operator, T = (A - sigma)^-1 R with sigma = min(d) -
 * shift (shift > 0), by the Woodbury identity on the low-rank part; the
 * naux x naux capacitance inverse is built once (FP64 MFMA GEMM, then a
 * device Cholesky inverse).  rt_casida_si_apply is an rt_matvec_fn (aux_op
 * of Davidson); rt_casida_si_info reports sigma and |I - C C^-1|_F.
 * Not in the reference: the structured correction of the learned
 * preconditioner. */
int rt_casida_si_create(void* op, double shift, void** si, void* stream);
int rt_casida_si_destroy(void* si);
int rt_casida_si_apply(void* si, const double* R, double* T, int64_t n,
                       int64_t k, void* stream);
int rt_casida_si_info(void* si, double* sigma, double* err);

/* ---- learned preconditioner: 2-layer FP32 MLP (hidden 64) ------------------ */
int rt_mlp_num_params(void);
int rt_mlp_create(uint64_t seed, void** net);
int rt_mlp_destroy(void* net);
int rt_mlp_get_params(void* net, float* out);
int rt_mlp_set_params(void* net, const float* in);
/* Inputs of sample s (columns s of the device arrays, n rows): residual
 * R[:, s], Ritz vector X[:, s], auxiliary correction S[:, s] (S may be NULL:
 * that basis vector is then zero) and the shifts ek[s], e0[s] (host).
 * T[:, s] = the network's correction. */
int rt_mlp_apply_dev(void* net, const double* R, const double* X, const double* S,
                     const double* hdiag, const double* ek, const double* e0,
                     int64_t n, int ns, double* T, void* stream);
/* Loss (host scalar) and gradient (host [rt_mlp_num_params()]) over ns
 * samples with device targets tau (n x ns) and host weights wts[ns].  When Q
 * is given, sample s's correction is compared with tau after projecting out
 * the q_cnt[s] orthonormal vectors at Q + q_off[s] (host arrays; q_off[s] < 0:
 * no projection), as Davidson projects it before use. */
int rt_mlp_loss_grad_dev(void* net, const double* R, const double* X,
                         const double* S, const double* hdiag, const double* ek,
                         const double* e0, const double* tau, const double* wts,
                         int64_t n, int ns, const double* Q, const int64_t* q_off,
                         const int* q_cnt, double* loss, double* grad,
                         void* stream);
/* Adam over `steps` minibatches of `batch` samples (seeded draw);
 * loss_hist: host [steps] or NULL. */
int rt_mlp_train_dev(void* net, const double* R, const double* X, const double* S,
                     const double* hdiag, const double* ek, const double* e0,
                     const double* tau, const double* wts, int64_t n, int ns,
                     const double* Q, const int64_t* q_off, const int* q_cnt,
                     int steps, int batch, double lr, uint64_t seed,
                     double* loss_hist, void* stream);
/* targets tau[:, s] = normalize(x* - Q_s Q_s^T x*) with x* = Xfinal[:, root[s]]
 * and Q_s sample s's subspace (without Q: normalize(x* - (x*.x) x));
 * wts[s] = 0 where |x*.x| < min_overlap (root mismatch) or no subspace was kept. */
int rt_mlp_targets_dev(const double* X, const int* root, int ns,
                       const double* Xfinal, int64_t n, double min_overlap,
                       const double* Q, const int64_t* q_off, const int* q_cnt,
                       double* tau, double* wts, void* stream);

/* ---- batched pipeline for many molecules (grouped launches + HIP graph) ---- */
/* Device buffers of one molecule: ri3ao [nbas,nbas,naux], C [nbas,nbas];
 * outputs D, J, K [nbas,nbas] and B [nocc,nbas,naux] (ri3mo(occ,all,aux)). */
typedef struct rt_mol_desc {
  int64_t nbas, naux, nocc;
  const double* ri3ao;
  const double* C;
  double* D;
  double* J;
  double* K;
  double* B;
} rt_mol_desc;

/* Builds the grouped work lists (one persistent launch per pipeline stage
 * covering all molecules) and allocates the intermediates. */
int rt_batch_create(const rt_mol_desc* mols, int nmol, void** handle);
/* Times every candidate tile edge (64, 128) x persistent grid (1, 2, 4 per
 * CU) of each grouped GEMM stage on this molecule set and keeps the fastest. */
int rt_batch_autotune(void* handle, int reps, void* stream);
/* Refresh the batch's padded copy of every molecule's ri3ao (made at
 * rt_batch_create) after the caller changed ri3ao; ordered on `stream`. */
int rt_batch_reload(void* handle, void* stream);
/* Fix one stage's tile (0 density, 1 ao2mo s1, 2 ao2mo s2, 3 K partials):
 * bm x bn with waves_m of the 4 waves along M, grid_mul persistent
 * workgroups per CU; the shape must be one the autotuner considers
 * (RT_ERR_ARG otherwise).  Drops a captured graph. */
int rt_batch_set_stage(void* handle, int stage, int bm, int bn, int waves_m, int grid_mul);
/* Runs density, RI-J, ao2mo (occupied rows), RI-K and the energies of every
 * molecule (8 launches). */
int rt_batch_run(void* handle, void* stream);
/* Captures rt_batch_run's launches on `stream` (non-NULL) into a HIP graph;
 * rt_batch_graph_launch replays it. */
int rt_batch_graph_capture(void* handle, void* stream);
int rt_batch_graph_launch(void* handle, void* stream);
/* Host arrays [nmol]: E_J = 1/2 sum_P gamma_P^2, E_x = -1/2 sum_{n<nocc} K[n,n].
 * Synchronizes `stream`. */
int rt_batch_energies(void* handle, double* e_j, double* e_x, void* stream);
/* JSON text of the chosen tile configuration per stage. */
int rt_batch_info(void* handle, char* buf, int len);
int rt_batch_destroy(void* handle);

#ifdef __cplusplus
}
#endif

#endif /* REST_TENSORS_AMD_H */
