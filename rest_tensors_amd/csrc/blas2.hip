// Level-2 / rank-k primitives that REST composes on RI tensors (A7):
//   _dgemv  (matrix_blas_lapack.rs:38-70):  y = alpha op(A) x + beta y
//   _dsyrk  (matrix_blas_lapack.rs:392-413): C = alpha A A^T + beta C  ('N')
//                                            C = alpha A^T A + beta C  ('T'),
//                                            only the uplo triangle of C touched
//   einsum "ip,ip->p" (_einsum_02_*, matrix_blas_lapack.rs:1293-1349): column dots
// For the packed ri3ao [nbas(nbas+1)/2, naux] of rifull_to_matfull_symm these
// are the two HBM-bound passes of RI-J (gamma = A^T d, J = A gamma) and the
// batched SYRK of RI-K.  Every kernel sums in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
#include "timer.hpp"

namespace rta {

__device__ __forceinline__ double b2_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double b2_block_sum(double v, double* red) {
  v = b2_wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    red[16] = s;
  }
  __syncthreads();
  s = red[16];
  __syncthreads();
  return s;
}

// 'T': part[c * nsplit + s] = sum over rows of slice s of A[i, c] x[i]
__global__ __launch_bounds__(256) void gemv_t_kernel(const double* __restrict__ A,
                                                     int64_t lda, int64_t m,
                                                     const double* __restrict__ x,
                                                     int64_t incx, int64_t slice,
                                                     double* __restrict__ part) {
  __shared__ double red[17];
  const int64_t c = blockIdx.x, s = blockIdx.y;
  const int64_t i0 = s * slice, i1 = (i0 + slice < m) ? i0 + slice : m;
  const double* a = A + c * lda;
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc += a[i] * x[i * incx];
  acc = b2_block_sum(acc, red);
  if (threadIdx.x == 0) part[c * gridDim.y + s] = acc;
}

// 'T' with 16-byte loads of A (lda even, A 16-byte aligned, slice even): a
// lane takes row pairs (i, i+1), 4 pairs in flight; an odd last row alone
__global__ __launch_bounds__(256) void gemv_t_x2_kernel(const double* __restrict__ A,
                                                        int64_t lda, int64_t m,
                                                        const double* __restrict__ x,
                                                        int64_t incx, int64_t slice,
                                                        double* __restrict__ part) {
  __shared__ double red[17];
  const int64_t c = blockIdx.x, s = blockIdx.y;
  const int64_t i0 = s * slice, i1 = (i0 + slice < m) ? i0 + slice : m;
  const double2* a2 = reinterpret_cast<const double2*>(A + c * lda);
  double acc = 0.0, acc2 = 0.0;
  int64_t i = i0 + 2 * threadIdx.x;
  for (; i + 3 * 512 + 1 < i1; i += 4 * 512) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = a2[(i + u * 512) >> 1];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t r = i + u * 512;
      acc += v[u].x * x[r * incx];
      acc2 += v[u].y * x[(r + 1) * incx];
    }
  }
  for (; i + 1 < i1; i += 512) {
    const double2 v = a2[i >> 1];
    acc += v.x * x[i * incx];
    acc2 += v.y * x[(i + 1) * incx];
  }
  if (i < i1) acc += A[c * lda + i] * x[i * incx];
  acc = b2_block_sum(acc + acc2, red);
  if (threadIdx.x == 0) part[c * gridDim.y + s] = acc;
}

// y[c] = alpha * sum_s part[c * nsplit + s] + beta * y[c]
__global__ void gemv_t_finish(const double* __restrict__ part, int nsplit, int64_t n,
                              double alpha, double beta, double* __restrict__ y,
                              int64_t incy) {
  const int64_t c = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (c >= n) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[c * nsplit + k];
  double* yc = y + c * incy;
  *yc = alpha * s + (beta == 0.0 ? 0.0 : beta * *yc);
}

// 'N': part[s * m + i] = sum over columns of slice s of A[i, c] x[c]
__global__ __launch_bounds__(256) void gemv_n_kernel(const double* __restrict__ A,
                                                     int64_t lda, int64_t m, int64_t n,
                                                     const double* __restrict__ x,
                                                     int64_t incx, int64_t slice,
                                                     double* __restrict__ part) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  const int64_t s = blockIdx.y;
  if (i >= m) return;
  const int64_t c0 = s * slice, c1 = (c0 + slice < n) ? c0 + slice : n;
  double acc = 0.0;
  for (int64_t c = c0; c < c1; ++c) acc += A[i + c * lda] * x[c * incx];
  part[s * m + i] = acc;
}

// 'N' with 16-byte loads of A (lda even, A 16-byte aligned): a lane owns the
// row pair (i, i+1) (an odd last row alone), 8 columns in flight
__global__ __launch_bounds__(256) void gemv_n_x2_kernel(const double* __restrict__ A,
                                                        int64_t lda, int64_t m, int64_t n,
                                                        const double* __restrict__ x,
                                                        int64_t incx, int64_t slice,
                                                        double* __restrict__ part) {
  const int64_t i = 2 * (blockIdx.x * (int64_t)256 + threadIdx.x);
  const int64_t s = blockIdx.y;
  if (i >= m) return;
  const int64_t c0 = s * slice, c1 = (c0 + slice < n) ? c0 + slice : n;
  const double* a = A + i;
  int64_t c = c0;
  if (i + 1 < m) {
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
    for (; c + 7 < c1; c += 8) {
      double2 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const double2*>(a + (c + u) * lda);
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        const double xa = x[(c + u) * incx], xb = x[(c + u + 1) * incx];
        ax += v[u].x * xa;
        ay += v[u].y * xa;
        bx += v[u + 1].x * xb;
        by += v[u + 1].y * xb;
      }
    }
    for (; c < c1; ++c) {
      const double2 v = *reinterpret_cast<const double2*>(a + c * lda);
      const double xc = x[c * incx];
      ax += v.x * xc;
      ay += v.y * xc;
    }
    part[s * m + i] = ax + bx;
    part[s * m + i + 1] = ay + by;
  } else {
    double acc = 0.0;
    for (; c < c1; ++c) acc += a[c * lda] * x[c * incx];
    part[s * m + i] = acc;
  }
}

__global__ void gemv_n_finish(const double* __restrict__ part, int nsplit, int64_t m,
                              double alpha, double beta, double* __restrict__ y,
                              int64_t incy) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[(int64_t)k * m + i];
  double* yi = y + i * incy;
  *yi = alpha * s + (beta == 0.0 ? 0.0 : beta * *yi);
}

// C[tri] = alpha * T[tri] + beta * C[tri] on the uplo triangle (diagonal included)
__global__ void syrk_merge_kernel(const double* __restrict__ T, int64_t n, bool upper,
                                  double alpha, double beta, double* __restrict__ C,
                                  int64_t ldc) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= n * n) return;
  const int64_t i = e % n, j = e / n;
  if (upper ? (i > j) : (i < j)) return;
  double* c = C + i + j * ldc;
  *c = alpha * T[e] + (beta == 0.0 ? 0.0 : beta * *c);
}

// out[p] = sum_{i < rows} A[i + a_rows p] B[i + b_rows p], rows = min(a_rows, b_rows)
__global__ __launch_bounds__(256) void coldot_kernel(const double* __restrict__ A,
                                                     int64_t a_rows,
                                                     const double* __restrict__ B,
                                                     int64_t b_rows, int64_t rows,
                                                     double* __restrict__ out) {
  __shared__ double red[17];
  const int64_t p = blockIdx.x;
  const double* a = A + p * a_rows;
  const double* b = B + p * b_rows;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < rows; i += 256) acc += a[i] * b[i];
  acc = b2_block_sum(acc, red);
  if (threadIdx.x == 0) out[p] = acc;
}

static bool op_t(char c) { return !(c == 'N' || c == 'n'); }

}  // namespace rta

using namespace rta;

extern "C" int rt_dgemv_dev(char trans, int64_t m, int64_t n, double alpha,
                            const double* A, int64_t lda, const double* x, int64_t incx,
                            double beta, double* y, int64_t incy, void* stream) {
  RTA_CHECK(m >= 0 && n >= 0, "negative dimension");
  RTA_CHECK(lda >= std::max<int64_t>(m, 1), "lda %lld < m %lld", (long long)lda,
            (long long)m);
  RTA_CHECK(incx >= 1 && incy >= 1, "increments must be positive");
  const bool t = op_t(trans);
  const int64_t lx = t ? m : n, ly = t ? n : m;
  if (ly == 0) return RT_OK;
  RTA_CHECK(y && (x || lx == 0) && (A || m * n == 0), "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  KernelTimer kt("dgemv", s);
  // 16-byte loads of A: every column starts 16-byte aligned
  const bool x2 = lda % 2 == 0 && ((uintptr_t)A & 15) == 0;
  if (t) {
    // columns of A are the outputs: split long columns so that >= 2048 blocks run
    const int nsplit = (int)std::max<int64_t>(
        1, std::min<int64_t>((m + 4095) / 4096, (2048 + n - 1) / n));
    const int64_t slice = ((m + nsplit - 1) / nsplit + 1) / 2 * 2;  // even: x2 row pairs
    double* part = nullptr;
    RTA_TRY(scratch_get(SCRATCH_JK, sizeof(double) * n * nsplit, (void**)&part));
    if (m > 0 && x2)
      hipLaunchKernelGGL(gemv_t_x2_kernel, dim3((unsigned)n, nsplit), dim3(256), 0, s, A, lda,
                         m, x, incx, slice, part);
    else if (m > 0)
      hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)n, nsplit), dim3(256), 0, s, A, lda,
                         m, x, incx, slice, part);
    hipLaunchKernelGGL(gemv_t_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       part, m > 0 ? nsplit : 0, n, alpha, beta, y, incy);
  } else {
    const int64_t rows_per_block = x2 ? 512 : 256;
    const int64_t rblocks = (m + rows_per_block - 1) / rows_per_block;
    const int nsplit = (int)std::max<int64_t>(
        1, std::min<int64_t>(std::min<int64_t>(n, 64), (2048 + rblocks - 1) / rblocks));
    const int64_t slice = (n + nsplit - 1) / nsplit;
    double* part = nullptr;
    RTA_TRY(scratch_get(SCRATCH_JK, sizeof(double) * m * nsplit, (void**)&part));
    if (n > 0 && x2)
      hipLaunchKernelGGL(gemv_n_x2_kernel, dim3((unsigned)rblocks, nsplit), dim3(256), 0, s, A,
                         lda, m, n, x, incx, slice, part);
    else if (n > 0)
      hipLaunchKernelGGL(gemv_n_kernel, dim3((unsigned)rblocks, nsplit), dim3(256), 0, s, A,
                         lda, m, n, x, incx, slice, part);
    hipLaunchKernelGGL(gemv_n_finish, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, part,
                       n > 0 ? nsplit : 0, m, alpha, beta, y, incy);
  }
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

extern "C" int rt_dsyrk_dev(char uplo, char trans, int64_t n, int64_t k, double alpha,
                            const double* A, int64_t lda, double beta, double* C,
                            int64_t ldc, void* stream) {
  RTA_CHECK(n >= 0 && k >= 0, "negative dimension");
  RTA_CHECK(uplo == 'U' || uplo == 'u' || uplo == 'L' || uplo == 'l',
            "uplo must be 'U' or 'L'");
  const bool t = op_t(trans);
  RTA_CHECK(lda >= std::max<int64_t>(1, t ? k : n), "lda too small");
  RTA_CHECK(ldc >= std::max<int64_t>(1, n), "ldc too small");
  if (n == 0) return RT_OK;
  RTA_CHECK(C && (A || k == 0), "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  double* T = nullptr;
  RTA_TRY(scratch_get(SCRATCH_DROPIN_D, sizeof(double) * n * n, (void**)&T));
  // T = op(A) op(A)^T as a full product ('T': A^T A runs on the k-major TN path)
  if (t)
    RTA_TRY(rt_dgemm_dev('T', 'N', n, n, k, 1.0, A, lda, A, lda, 0.0, T, n, stream));
  else
    RTA_TRY(rt_dgemm_dev('N', 'T', n, n, k, 1.0, A, lda, A, lda, 0.0, T, n, stream));
  hipLaunchKernelGGL(syrk_merge_kernel, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0,
                     s, T, n, uplo == 'U' || uplo == 'u', alpha, beta, C, ldc);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

extern "C" int rt_einsum_ip_ip_p_dev(const double* A, int64_t a_rows, int64_t a_cols,
                                     const double* B, int64_t b_rows, int64_t b_cols,
                                     double* out, void* stream) {
  RTA_CHECK(a_rows >= 0 && a_cols >= 0 && b_rows >= 0 && b_cols >= 0,
            "negative dimension");
  const int64_t np = std::min(a_cols, b_cols);
  if (np == 0) return RT_OK;
  RTA_CHECK(out, "NULL output");
  const int64_t rows = std::min(a_rows, b_rows);
  RTA_CHECK((A && B) || rows == 0, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(coldot_kernel, dim3((unsigned)np), dim3(256), 0, s, A, a_rows, B,
                     b_rows, rows, out);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}
