// FP64 aux-metric functions kept in FP64 on the RI path (A11):
//   _dpotrf                     src/matrix/matrix_blas_lapack.rs:518-534
//   cholesky_decompose_inverse  src/matrix/matrix_blas_lapack.rs:931-1002
// Blocked right-looking Cholesky with CH_NB-wide panels: the diagonal block is
// factored by one workgroup in LDS, the panel below by a row-parallel
// triangular solve, the trailing matrix by the FP64 MFMA GEMM.  Inverses go
// through the triangular inverse (blocked forward substitution) and one TN
// GEMM.  (lapack_power, an eigen-decomposition, lives with the eigensolver.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "metric.hpp"
#include "timer.hpp"

namespace rta {

// Panel width: the panel kernels are one chain of dependent steps per row, so
// their latency (not the GPU's throughput) sets the pace; 32-column panels
// halve each chain (the shift-invert setup at m = 2048: 8.6 -> 7.2-8.1 ms).
constexpr int CH_NB = 32;

static unsigned mt_grid(int64_t len) { return (unsigned)((len + 255) / 256); }

// Row r0 + t of the CH_NB-column block of M (ld) at columns c0.. into out[CH_NB]:
// every address is clamped into the nr x nc valid part so that all 64 loads
// are issued before the first is waited for (a bounds branch per load would
// serialize them); values outside the valid part are replaced by the caller.
__device__ __forceinline__ void panel_row(const double* __restrict__ M, int64_t ld, int64_t r0,
                                          int nr, int64_t c0, int nc, int t, double* out) {
  const double* row = M + r0 + (t < nr ? t : nr - 1);
#pragma unroll
  for (int j = 0; j < CH_NB; ++j) out[j] = row[ld * (c0 + (j < nc ? j : nc - 1))];
}

// One workgroup of CH_NB threads factors W[k0:k0+kb, k0:k0+kb] (lower),
// right-looking: thread t keeps row t in registers; at step j thread j takes
// the pivot, every row below scales its entry j by the pivot's reciprocal (as
// LAPACK's DPOTF2 does), the column is shared through LDS and each row updates
// its entries c > j, all of them: the entries c > t (the upper triangle)
// take garbage that is never stored, which saves a divergent branch per
// entry.  The steps are unrolled, so the row stays in registers.  info (device): first failing
// pivot (1-based) if still 0.
__global__ __launch_bounds__(CH_NB) void potf2_kernel(double* __restrict__ W, int64_t ld,
                                                      int64_t k0, int kb,
                                                      int* __restrict__ info) {
  __shared__ double col[CH_NB];
  __shared__ double piv;
  const int t = threadIdx.x;
  double r[CH_NB];
  {
    double v[CH_NB];
    panel_row(W, ld, k0, kb, k0, kb, t, v);
#pragma unroll
    for (int j = 0; j < CH_NB; ++j) r[j] = (t < kb && j < kb) ? v[j] : (t == j ? 1.0 : 0.0);
  }
#pragma unroll
  for (int j = 0; j < CH_NB; ++j) {
    if (t == j) {
      const double d = r[j];
      if (j < kb && !(d > 0.0) && *info == 0) *info = (int)(k0 + j + 1);
      r[j] = sqrt(d);
      piv = 1.0 / r[j];
    }
    __syncthreads();
    if (t > j) r[j] *= piv;
    const double lj = r[j];
    col[t] = lj;
    __syncthreads();
    if (t > j) {
#pragma unroll
      for (int c = j + 1; c < CH_NB; ++c) r[c] -= lj * col[c];
    }
    __syncthreads();  // col and piv are rewritten by the next step
  }
  if (t < kb)
#pragma unroll
    for (int j = 0; j < CH_NB; ++j)
      if (j <= t && j < kb) W[(k0 + t) + ld * (k0 + j)] = r[j];
}

// Rows r in [r0, r1): W[r, k0:k0+kb] <- W[r, k0:k0+kb] L11^-T (L11 = the
// factored diagonal block), i.e. x_j = (a_j - sum_{i<j} x_i L11[j][i]) / L11[j][j].
// CH_NB threads, one row each; the block's rows and L11 live in LDS.
__global__ __launch_bounds__(CH_NB) void trsm_rows_kernel(double* __restrict__ W, int64_t ld,
                                                          int64_t k0, int kb, int64_t r0,
                                                          int64_t r1) {
  __shared__ double L[CH_NB][CH_NB + 1];
  __shared__ double xs[CH_NB][CH_NB + 1];
  const int t = threadIdx.x;
  const int64_t rb = r0 + blockIdx.x * (int64_t)CH_NB;
  {
    double v[CH_NB];
    panel_row(W, ld, k0, kb, k0, kb, t, v);
#pragma unroll
    for (int j = 0; j < CH_NB; ++j) L[t][j] = (t < kb && j < kb) ? v[j] : (t == j ? 1.0 : 0.0);
    const int nr = (int)((r1 - rb < CH_NB) ? r1 - rb : CH_NB);
    panel_row(W, ld, rb, nr, k0, kb, t, v);
#pragma unroll
    for (int j = 0; j < CH_NB; ++j) xs[t][j] = (t < nr && j < kb) ? v[j] : 0.0;
  }
  __syncthreads();
  if (rb + t >= r1) return;
  for (int j = 0; j < kb; ++j) {
    // four partial sums in flight (fixed order)
    double v0 = xs[t][j], v1 = 0.0, v2 = 0.0, v3 = 0.0;
    int i = 0;
    for (; i + 3 < j; i += 4) {
      v0 -= xs[t][i] * L[j][i];
      v1 -= xs[t][i + 1] * L[j][i + 1];
      v2 -= xs[t][i + 2] * L[j][i + 2];
      v3 -= xs[t][i + 3] * L[j][i + 3];
    }
    for (; i < j; ++i) v0 -= xs[t][i] * L[j][i];
    xs[t][j] = ((v0 + v1) + (v2 + v3)) / L[j][j];
  }
#pragma unroll 8
  for (int j = 0; j < kb; ++j) W[(rb + t) + ld * (k0 + j)] = xs[t][j];
}

// S (n x n, ld n) = the symmetric matrix whose uplo triangle is A's
__global__ void mt_symmetrize_kernel(const double* __restrict__ A, int64_t lda, int64_t n,
                                     bool upper, double* __restrict__ S) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= n * n) return;
  const int64_t i = e % n, j = e / n;
  const bool in = upper ? (i <= j) : (i >= j);
  S[e] = in ? A[i + lda * j] : A[j + lda * i];
}

// out (ld ldo) <- the lower triangle of W (ld n) or its transpose into the
// upper triangle; zero_other: write zeros in the other strict triangle
__global__ void mt_store_tri_kernel(const double* __restrict__ W, int64_t n, bool upper,
                                    bool zero_other, double* __restrict__ out, int64_t ldo) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= n * n) return;
  const int64_t i = e % n, j = e / n;
  if (upper) {
    if (i <= j) out[i + ldo * j] = W[j + n * i];
    else if (zero_other) out[i + ldo * j] = 0.0;
  } else {
    if (i >= j) out[i + ldo * j] = W[i + n * j];
    else if (zero_other) out[i + ldo * j] = 0.0;
  }
}

__global__ void mt_fill_kernel(double* __restrict__ X, int64_t len, double v) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e < len) X[e] = v;
}

// W (ld n) <- lower Cholesky factor of W in its lower triangle (the strict
// upper triangle is left holding intermediate values); info (host) = first
// non-positive pivot (1-based) or 0.
// Two levels of blocking: CH_NB-column panels (potf2 + trsm_rows) inside
// 512-column blocks; a panel's update reaches only the rest of its block, and
// each block ends with one rank-512 update of the trailing matrix (one GEMM of
// depth 512 instead of one per panel, each re-reading and re-writing the
// trailing matrix; at n = 4096 blocks of 256 measured 2-4 % slower).
int chol_lower(double* W, int64_t n, int* info_h, int* info_d, hipStream_t s) {
  constexpr int64_t NB2 = 512;
  RTA_HIP(hipMemsetAsync(info_d, 0, sizeof(int), s));
  for (int64_t b0 = 0; b0 < n; b0 += NB2) {
    const int64_t b1 = std::min<int64_t>(b0 + NB2, n);
    for (int64_t k0 = b0; k0 < b1; k0 += CH_NB) {
      const int kb = (int)std::min<int64_t>(CH_NB, b1 - k0);
      hipLaunchKernelGGL(potf2_kernel, dim3(1), dim3(CH_NB), 0, s, W, n, k0, kb, info_d);
      const int64_t r0 = k0 + kb;
      if (r0 < n) {
        hipLaunchKernelGGL(trsm_rows_kernel, dim3((unsigned)((n - r0 + CH_NB - 1) / CH_NB)),
                           dim3(CH_NB), 0, s, W, n, k0, kb, r0, n);
        if (r0 < b1)  // the block's remaining columns, all rows below
          RTA_TRY(rt_dgemm_dev('N', 'T', n - r0, b1 - r0, kb, -1.0, W + r0 + n * k0, n,
                               W + r0 + n * k0, n, 1.0, W + r0 + n * r0, n, s));
      }
      RTA_HIP(hipGetLastError());
    }
    if (b1 < n) {
      const int64_t m2 = n - b1;
      RTA_TRY(rt_dgemm_dev('N', 'T', m2, m2, b1 - b0, -1.0, W + b1 + n * b0, n, W + b1 + n * b0,
                           n, 1.0, W + b1 + n * b1, n, s));
    }
  }
  RTA_HIP(hipMemcpyAsync(info_h, info_d, sizeof(int), hipMemcpyDeviceToHost, s));
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}

// Diagonal CH_NB blocks of X (ld n) <- their inverses L_bb^-1, one
// workgroup per block: thread t forward-substitutes column t of the identity
// (y_j = 0 for j < t; four partial sums in flight), the block is written row
// by row.
__global__ __launch_bounds__(CH_NB) void trinv_diag_kernel(const double* __restrict__ Lf,
                                                           int64_t n, double* __restrict__ X) {
  __shared__ double L[CH_NB][CH_NB + 1];
  __shared__ double ys[CH_NB][CH_NB + 1];
  const int t = threadIdx.x;
  const int64_t k0 = (int64_t)blockIdx.x * CH_NB;
  const int kb = (int)((n - k0 < CH_NB) ? n - k0 : CH_NB);
  {
    double v[CH_NB];
    panel_row(Lf, n, k0, kb, k0, kb, t, v);
#pragma unroll
    for (int j = 0; j < CH_NB; ++j) L[t][j] = (t < kb && j < kb) ? v[j] : (t == j ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    double v0 = (j == t) ? 1.0 : 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    int i = 0;
    for (; i + 3 < j; i += 4) {
      v0 -= L[j][i] * ys[t][i];
      v1 -= L[j][i + 1] * ys[t][i + 1];
      v2 -= L[j][i + 2] * ys[t][i + 2];
      v3 -= L[j][i + 3] * ys[t][i + 3];
    }
    for (; i < j; ++i) v0 -= L[j][i] * ys[t][i];
    ys[t][j] = (j < t) ? 0.0 : ((v0 + v1) + (v2 + v3)) / L[j][j];
  }
  __syncthreads();
  if (t < kb)
    for (int c = 0; c < kb; ++c) X[(k0 + t) + n * (k0 + c)] = ys[c][t];
}

// X (ld n) <- L^-1 for the lower factor in Lf (ld n), by recursive doubling:
// the diagonal CH_NB blocks are inverted in one launch; then for h = CH_NB,
// 2 CH_NB, ... every pair of inverted h-blocks on the diagonal is joined,
//   [L11 0; L21 L22]^-1 = [X11 0; -X22 (L21 X11) X22],
// the pairs of one level as two strided-batched GEMMs (the last, shorter
// pair on its own).  The strict upper triangle of X is zero.
int tri_inverse_lower(const double* Lf, int64_t n, double* X, hipStream_t s) {
  hipLaunchKernelGGL(mt_fill_kernel, dim3(mt_grid(n * n)), dim3(256), 0, s, X, n * n, 0.0);
  hipLaunchKernelGGL(trinv_diag_kernel, dim3((unsigned)((n + CH_NB - 1) / CH_NB)), dim3(CH_NB),
                     0, s, Lf, n, X);
  RTA_HIP(hipGetLastError());
  if (n <= CH_NB) return RT_OK;
  // T: L21 X11 of every pair of a level, h x h each (ld h), the shorter last
  // pair included
  int64_t tlen = 0;
  for (int64_t h = CH_NB; h < n; h *= 2) {
    const int64_t np = n / (2 * h);
    tlen = std::max(tlen, np * h * h);
    if (np * 2 * h + h < n) tlen = std::max(tlen, h * h);
  }
  double* T = nullptr;
  if (hipMalloc(&T, sizeof(double) * (size_t)tlen) != hipSuccess) {
    (void)hipGetLastError();
    return fail(RT_ERR_NOMEM, "triangular inverse: scratch allocation failed");
  }
  int st = RT_OK;
  for (int64_t h = CH_NB; h < n && st == RT_OK; h *= 2) {
    const int64_t np = n / (2 * h);  // full pairs
    auto join = [&](int64_t off, int64_t h2, int64_t batch) -> int {
      GemmDesc d1;  // T_p = L21 X11 (h2 x h)
      d1.M = h2;
      d1.N = h;
      d1.K = h;
      d1.A = Lf + (off + h) + n * off;
      d1.lda = n;
      d1.sA = 2 * h * (n + 1);
      d1.B = X + off + n * off;
      d1.ldb = n;
      d1.sB = 2 * h * (n + 1);
      d1.C = T;
      d1.c_rs = 1;
      d1.c_cs = h;
      d1.sC = h * h;
      d1.batch = batch;
      RTA_TRY(gemm_batched(d1, s, "tri_inv"));
      GemmDesc d2;  // X21 = -X22 T_p (h2 x h)
      d2.M = h2;
      d2.N = h;
      d2.K = h2;
      d2.A = X + (off + h) + n * (off + h);
      d2.lda = n;
      d2.sA = 2 * h * (n + 1);
      d2.B = T;
      d2.ldb = h;
      d2.sB = h * h;
      d2.C = X + (off + h) + n * off;
      d2.c_rs = 1;
      d2.c_cs = n;
      d2.sC = 2 * h * (n + 1);
      d2.alpha = -1.0;
      d2.batch = batch;
      return gemm_batched(d2, s, "tri_inv");
    };
    if (np > 0) st = join(0, h, np);
    const int64_t off = np * 2 * h;
    if (st == RT_OK && off + h < n) st = join(off, n - off - h, 1);
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(T);
  return st;
}

struct MtBufs {
  std::vector<double*> p;
  int* info = nullptr;
  ~MtBufs() {
    for (double* q : p) (void)hipFree(q);
    if (info) (void)hipFree(info);
  }
  int take(size_t count, double** out) {
    double* q = nullptr;
    if (hipMalloc(&q, sizeof(double) * std::max<size_t>(count, 1)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(RT_ERR_NOMEM, "metric: allocation of %zu doubles failed", count);
    }
    p.push_back(q);
    *out = q;
    return RT_OK;
  }
};

// Sinv = S^-1 for the SPD matrix S (ld n) through its Cholesky factor
static int spd_inverse(const double* S, int64_t n, double* Sinv, MtBufs& b, hipStream_t s,
                       int* info_h) {
  double *W, *X;
  RTA_TRY(b.take((size_t)n * n, &W));
  RTA_TRY(b.take((size_t)n * n, &X));
  RTA_HIP(hipMemcpyAsync(W, S, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  RTA_TRY(chol_lower(W, n, info_h, b.info, s));
  if (*info_h != 0) return RT_OK;
  RTA_TRY(tri_inverse_lower(W, n, X, s));
  // S^-1 = L^-T L^-1 = X^T X
  return rt_dgemm_dev('T', 'N', n, n, n, 1.0, X, n, X, n, 0.0, Sinv, n, s);
}

static bool is_upper(char uplo) { return uplo == 'U' || uplo == 'u'; }

}  // namespace rta

using namespace rta;

extern "C" int rt_dpotrf_dev(char uplo, int64_t n, double* A, int64_t lda, int* info,
                             void* stream) {
  RTA_CHECK(uplo == 'U' || uplo == 'u' || uplo == 'L' || uplo == 'l',
            "uplo must be 'U' or 'L'");
  RTA_CHECK(n >= 0 && lda >= std::max<int64_t>(1, n), "bad dimensions");
  RTA_CHECK(info, "info is NULL");
  *info = 0;
  if (n == 0) return RT_OK;
  RTA_CHECK(A, "A is NULL");
  hipStream_t s = (hipStream_t)stream;
  MtBufs b;
  RTA_HIP(hipMalloc(&b.info, sizeof(int)));
  double* W;
  RTA_TRY(b.take((size_t)n * n, &W));
  KernelTimer kt("dpotrf", s);
  hipLaunchKernelGGL(mt_symmetrize_kernel, dim3(mt_grid(n * n)), dim3(256), 0, s, A, lda, n,
                     is_upper(uplo), W);
  RTA_TRY(chol_lower(W, n, info, b.info, s));
  // the factor replaces the uplo triangle; the other strict triangle is untouched
  hipLaunchKernelGGL(mt_store_tri_kernel, dim3(mt_grid(n * n)), dim3(256), 0, s, W, n,
                     is_upper(uplo), false, A, lda);
  RTA_HIP(hipGetLastError());
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}

extern "C" int rt_cholesky_decompose_inverse_dev(char uplo, int64_t n, const double* A,
                                                 int64_t lda, double* out, int64_t ldo,
                                                 void* stream) {
  RTA_CHECK(uplo == 'U' || uplo == 'u' || uplo == 'L' || uplo == 'l',
            "uplo must be 'U' or 'L'");
  RTA_CHECK(n >= 0 && lda >= std::max<int64_t>(1, n) && ldo >= std::max<int64_t>(1, n),
            "bad dimensions");
  if (n == 0) return RT_OK;
  RTA_CHECK(A && out, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  MtBufs b;
  RTA_HIP(hipMalloc(&b.info, sizeof(int)));
  double *S, *Sinv;
  RTA_TRY(b.take((size_t)n * n, &S));
  RTA_TRY(b.take((size_t)n * n, &Sinv));
  KernelTimer kt("chol_inv", s);
  hipLaunchKernelGGL(mt_symmetrize_kernel, dim3(mt_grid(n * n)), dim3(256), 0, s, A, lda, n,
                     is_upper(uplo), S);
  int info = 0;
  RTA_TRY(spd_inverse(S, n, Sinv, b, s, &info));
  if (info != 0)
    return fail(RT_ERR_NUMERIC, "DPOTRF failed with %d (leading minor not positive definite)",
                info);
  // the factor of the inverse: S^-1 = M M^T (lower) or U^T U (upper: U = M^T);
  // like the reference, a failure of this second factorization is not reported
  RTA_TRY(chol_lower(Sinv, n, &info, b.info, s));
  hipLaunchKernelGGL(mt_store_tri_kernel, dim3(mt_grid(n * n)), dim3(256), 0, s, Sinv, n,
                     is_upper(uplo), true, out, ldo);
  RTA_HIP(hipGetLastError());
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}
