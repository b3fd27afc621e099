// Device kernels of the Davidson solver (A12/A15): vector algebra on
// device-resident subspaces, the diagonal preconditioners, and the skinny
// products of the RI linear-response operator A = diag(d) + 2 s^2 B B^T.
//
// Vectors are columns of length n in column-major blocks (leading dim n).
// All reductions run in a fixed order (deterministic results).
#include <hip/hip_runtime.h>

#include <vector>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "dvkernels.hpp"
#include "gemm.hpp"
#include "metric.hpp"
#include "timer.hpp"

namespace rta {

__device__ __forceinline__ double dv_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of one value per thread (blockDim.x multiple of 64, <= 1024).
__device__ double dv_block_sum(double v, double* red) {
  v = dv_wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < nw; ++i) s += red[i];
    red[32] = s;
  }
  __syncthreads();
  s = red[32];
  return s;
}

// ---- Gram blocks: G[i + ldg*j] = U[:, i] . W[:, j] ---------------------------
// Block (slice of rows, 8 columns of U, 8 columns of W): every thread sums its
// rows into an 8 x 8 register tile; the 64 values are reduce-scattered across
// the wave (6 butterfly steps that halve the values each lane holds: lane l
// ends with the wave's sum of value l, 63 shuffles instead of 64 x 6; 1024-row
// slices measured faster than 256-row ones, whose slice sum costs more than
// the shorter row loop saves), then summed across the 4 waves in a fixed
// order; partial[s][i + nu*j] per slice s, then a fixed-order sum over the
// slices.
// rows per slice: at least kGramRows, at most kGramSlices slices
constexpr int64_t kGramRows = 1024, kGramSlices = 512;

__global__ __launch_bounds__(256) void gram_tile_kernel(
    const double* __restrict__ U, int64_t ldu, int nu,
    const double* __restrict__ W, int64_t ldw, int nw, int64_t n, int64_t slice,
    double* __restrict__ partial) {
  constexpr int TI = 8, TJ = 8, NV = TI * TJ;
  __shared__ double red[4][NV];
  const int64_t x0 = (int64_t)blockIdx.x * slice;
  const int64_t x1 = (x0 + slice < n) ? x0 + slice : n;
  const int i0 = blockIdx.y * TI, j0 = blockIdx.z * TJ;
  const int ni = (nu - i0 < TI) ? nu - i0 : TI, nj = (nw - j0 < TJ) ? nw - j0 : TJ;
  double v[NV];  // v[i + TI j]
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  for (int64_t x = x0 + threadIdx.x; x < x1; x += 256) {
    double u[TI], w[TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i) u[i] = (i < ni) ? U[x + (int64_t)(i0 + i) * ldu] : 0.0;
#pragma unroll
    for (int j = 0; j < TJ; ++j) w[j] = (j < nj) ? W[x + (int64_t)(j0 + j) * ldw] : 0.0;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) v[i + TI * j] += u[i] * w[j];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // step with offset o: a lane with bit o set keeps the upper half of its
  // values and sends the lower half to its partner, which does the opposite
#pragma unroll
  for (int o = 32, h = NV / 2; o >= 1; o >>= 1, h >>= 1) {
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int q = 0; q < h; ++q) {
      const double send = up ? v[q] : v[q + h];
      const double keep = up ? v[q + h] : v[q];
      v[q] = keep + __shfl_xor(send, o, 64);
    }
  }
  red[wave][lane] = v[0];
  __syncthreads();
  const int t = threadIdx.x;
  if (t < NV) {
    const int i = t % TI, j = t / TI;
    if (i < ni && j < nj)
      partial[(int64_t)blockIdx.x * nu * nw + (i0 + i) + (int64_t)nu * (j0 + j)] =
          (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
  }
}

__global__ void gram_reduce_kernel(const double* __restrict__ partial,
                                   int nsplit, int nu, int nw,
                                   double* __restrict__ G, int64_t ldg) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nu * nw) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += partial[(int64_t)k * nu * nw + p];
  G[(p % nu) + ldg * (p / nu)] = s;
}

static int gram_nsplit(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kGramSlices, (n + kGramRows - 1) / kGramRows));
}

int dv_gram(const double* U, int64_t ldu, int nu, const double* W, int64_t ldw,
            int nw, int64_t n, double* G, int64_t ldg, double* scratch,
            hipStream_t s) {
  if (nu <= 0 || nw <= 0) return RT_OK;
  const int ns = gram_nsplit(n);
  const int64_t slice = (n + ns - 1) / ns;
  KernelTimer kt("dv_gram", s);
  hipLaunchKernelGGL(gram_tile_kernel, dim3(ns, (nu + 7) / 8, (nw + 7) / 8), dim3(256), 0, s,
                     U, ldu, nu, W, ldw, nw, n, slice, scratch);
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((nu * nw + 255) / 256),
                     dim3(256), 0, s, scratch, ns, nu, nw, G, ldg);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

int64_t dv_gram_scratch(int64_t n, int nu, int nw) {
  return (int64_t)gram_nsplit(n) * nu * nw;
}

// ---- Y[:, c] = sum_s V[:, s] coef[s + ldc*c]  (tall times small) ---------------
__global__ __launch_bounds__(256) void combine_kernel(
    const double* __restrict__ V, int64_t ldv, int ns,
    const double* __restrict__ coef, int ldc, int nc, int64_t n,
    double* __restrict__ Y, int64_t ldy) {
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (x >= n || c >= nc) return;
  double acc = 0.0;
  for (int s = 0; s < ns; ++s) acc += V[x + s * ldv] * coef[s + ldc * c];
  Y[x + c * ldy] = acc;
}

int dv_combine(const double* V, int64_t ldv, int ns, const double* coef,
               int ldc, int nc, int64_t n, double* Y, int64_t ldy,
               hipStream_t s) {
  if (nc <= 0 || n <= 0) return RT_OK;
  hipLaunchKernelGGL(combine_kernel, dim3((unsigned)((n + 255) / 256), nc),
                     dim3(256), 0, s, V, ldv, ns, coef, ldc, nc, n, Y, ldy);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// ---- Ritz vectors, residuals and column norms ------------------------------------
// sum of squares of columns: out[k] = sum_s partial[k*ns + s]
__global__ void colsum_kernel(const double* __restrict__ partial, int ns,
                              int nk, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  double s = 0.0;
  for (int i = 0; i < ns; ++i) s += partial[k * ns + i];
  out[k] = s;
}


// Ritz vectors and residuals in one pass over the subspace blocks:
//   X0[:, k] = XS C[:, k],  R[:, k] = AX C[:, k] - e_k X0[:, k]
// (the sums over the subspace in the order of combine_kernel), with the
// squared norm of each residual per row slice.
__global__ __launch_bounds__(256) void ritz_residual_kernel(
    const double* __restrict__ XS, const double* __restrict__ AX, int64_t ld, int ns,
    const double* __restrict__ coef, int ldc, const double* __restrict__ e, int64_t n,
    int64_t slice, double* __restrict__ X0, double* __restrict__ R,
    double* __restrict__ sq_partial) {
  __shared__ double red[40];
  const int k = blockIdx.y;
  const int64_t x0 = (int64_t)blockIdx.x * slice;
  const int64_t x1 = (x0 + slice < n) ? x0 + slice : n;
  const double ek = e[k];
  const double* c = coef + (int64_t)ldc * k;
  double acc = 0.0;
  for (int64_t x = x0 + threadIdx.x; x < x1; x += 256) {
    double xv = 0.0, av = 0.0;
    for (int q = 0; q < ns; ++q) {
      const double cq = c[q];
      xv += XS[x + q * ld] * cq;
      av += AX[x + q * ld] * cq;
    }
    X0[x + k * n] = xv;
    const double r = av - ek * xv;
    R[x + k * n] = r;
    acc += r * r;
  }
  acc = dv_block_sum(acc, red);
  if (threadIdx.x == 0) sq_partial[k * gridDim.x + blockIdx.x] = acc;
}

int dv_ritz_residuals(const double* XS, const double* AX, int64_t ld, int ns,
                      const double* coef, int ldc, int nk, const double* e_dev, int64_t n,
                      double* X0, double* R, double* sq_out, double* scratch, hipStream_t s) {
  if (nk <= 0) return RT_OK;
  // about one row per thread: 4 slices per Gram slice (scratch >= 4 nsplit nk)
  const int nsl = 4 * gram_nsplit(n);
  const int64_t slice = (n + nsl - 1) / nsl;
  hipLaunchKernelGGL(ritz_residual_kernel, dim3(nsl, nk), dim3(256), 0, s, XS, AX, ld, ns, coef,
                     ldc, e_dev, n, slice, X0, R, scratch);
  hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(64), 0, s, scratch, nsl, nk, sq_out);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// squared norms of nk columns (ld n)
__global__ __launch_bounds__(256) void sqnorm_kernel(const double* __restrict__ X,
                                                     int64_t n, int64_t slice,
                                                     double* __restrict__ part) {
  __shared__ double red[40];
  const int k = blockIdx.y;
  const int64_t x0 = (int64_t)blockIdx.x * slice;
  const int64_t x1 = (x0 + slice < n) ? x0 + slice : n;
  double acc = 0.0;
  for (int64_t x = x0 + threadIdx.x; x < x1; x += 256) {
    double v = X[x + k * n];
    acc += v * v;
  }
  acc = dv_block_sum(acc, red);
  if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = acc;
}

int dv_sqnorms(const double* X, int nk, int64_t n, double* sq_out,
               double* scratch, hipStream_t s) {
  if (nk <= 0) return RT_OK;
  const int ns = gram_nsplit(n);
  const int64_t slice = (n + ns - 1) / ns;
  hipLaunchKernelGGL(sqnorm_kernel, dim3(ns, nk), dim3(256), 0, s, X, n, slice,
                     scratch);
  hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(64), 0, s, scratch, ns, nk,
                     sq_out);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// ---- preconditioners --------------------------------------------------------------
// T[:,k] = R[:,k] / (hdiag - shift[k]) with |h| < 1e-8 -> 1e-8 (davidson.rs:63-70)
__global__ __launch_bounds__(256) void precond_diag_kernel(
    const double* __restrict__ R, const double* __restrict__ hdiag,
    const double* __restrict__ shift, int64_t n, double* __restrict__ T) {
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (x >= n) return;
  double h = hdiag[x] - shift[k];
  if (fabs(h) < 1e-8) h = 1e-8;
  T[x + k * n] = R[x + k * n] / h;
}

int dv_precond_diag(const double* R, const double* hdiag, const double* shift,
                    int nk, int64_t n, double* T, hipStream_t s) {
  if (nk <= 0) return RT_OK;
  hipLaunchKernelGGL(precond_diag_kernel, dim3((unsigned)((n + 255) / 256), nk),
                     dim3(256), 0, s, R, hdiag, shift, n, T);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// X[:, k] *= 1 / sqrt(sq[k])   (sq on the device)
__global__ void rnorm_cols_kernel(double* __restrict__ X, const double* __restrict__ sq,
                                  int64_t n) {
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (x < n) X[x + k * n] *= 1.0 / sqrt(sq[k]);
}

int dv_normalize_cols(double* X, int nk, int64_t n, double* sq, double* scratch,
                      hipStream_t s) {
  if (nk <= 0) return RT_OK;
  RTA_TRY(dv_sqnorms(X, nk, n, sq, scratch, s));
  hipLaunchKernelGGL(rnorm_cols_kernel, dim3((unsigned)((n + 255) / 256), nk), dim3(256), 0, s,
                     X, sq, n);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// T[:, c] -= sum_s Q[:, s] G[s + ns*c]
__global__ __launch_bounds__(256) void sub_combine_kernel(const double* __restrict__ Q,
                                                          int64_t ldq, int ns,
                                                          const double* __restrict__ G,
                                                          double* __restrict__ T, int64_t n) {
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (x >= n) return;
  double acc = 0.0;
  for (int q = 0; q < ns; ++q) acc += Q[x + q * ldq] * G[q + (int64_t)ns * c];
  T[x + c * n] -= acc;
}

int dv_project_out_block(const double* Q, int64_t ldq, int ns, double* T, int nt, int64_t n,
                         double* G, double* scratch, hipStream_t s) {
  if (ns <= 0 || nt <= 0) return RT_OK;
  KernelTimer kt("dv_project", s);
  for (int pass = 0; pass < 2; ++pass) {
    RTA_TRY(dv_gram(Q, ldq, ns, T, n, nt, n, G, ns, scratch, s));
    hipLaunchKernelGGL(sub_combine_kernel, dim3((unsigned)((n + 255) / 256), nt), dim3(256), 0,
                       s, Q, ldq, ns, G, T, n);
    RTA_HIP(hipGetLastError());
  }
  return RT_OK;
}

// Modified Gram-Schmidt with drops on the Gram matrix G = X^T X (host,
// k x k): column i becomes v = e_i - sum_j (c_j^T G v) c_j over the accepted
// coefficient columns c_j in order, kept as v / |v|_G when v^T G v > lindep.
// The accepted columns go to C (k x nv, ld k); returns nv.  *min_frac: the
// smallest share v^T G v / G_ii of its squared norm an accepted column kept.
static int gs_gram_host(const double* G, int k, double lindep, double* C,
                        double* min_frac) {
  std::vector<double> v(k), Gv(k);
  int nv = 0;
  for (int i = 0; i < k; ++i) {
    for (int r = 0; r < k; ++r) v[r] = (r == i) ? 1.0 : 0.0;
    for (int j = 0; j < nv; ++j) {
      const double* cj = C + (size_t)k * j;
      double prod = 0.0;
      for (int c = 0; c < k; ++c) {
        double gv = 0.0;
        for (int r = 0; r < k; ++r) gv += G[c + (size_t)k * r] * v[r];
        prod += cj[c] * gv;
      }
      for (int r = 0; r < k; ++r) v[r] -= prod * cj[r];
    }
    double nn = 0.0;
    for (int c = 0; c < k; ++c) {
      double gv = 0.0;
      for (int r = 0; r < k; ++r) gv += G[c + (size_t)k * r] * v[r];
      nn += v[c] * gv;
    }
    const double norm = std::sqrt(std::max(nn, 0.0));
    if (norm * norm > lindep) {
      double* cv = C + (size_t)k * nv;
      for (int r = 0; r < k; ++r) cv[r] = v[r] / norm;
      ++nv;
      const double gii = G[i + (size_t)k * i];
      if (gii > 0.0) *min_frac = std::min(*min_frac, nn / gii);
    }
  }
  return nv;
}

int dv_qr(double* X, int nx, int64_t n, double lindep, bool prescale, double* Q, int* nout,
          double* G, double* coef, double* scratch, double* host, hipStream_t s) {
  KernelTimer kt("dv_mgs", s);
  *nout = 0;
  int k = nx;
  double* Gh = host;            // k x k Gram matrix
  double* Ch = host + (size_t)nx * nx;  // k x nv coefficients
  // a second pass re-orthogonalizes only when the first cancelled heavily: a
  // column that kept at least 1 % of its squared norm loses orthogonality to
  // about eps / 0.01, which the Rayleigh-Ritz step does not notice
  double min_frac = 1.0;
  for (int pass = 0; pass < 2 && k > 0; ++pass) {
    if (pass == 1 && min_frac >= 1e-2) break;
    // pass 0: X -> Q; pass 1: Q -> X, copied back to Q
    const double* src = pass == 0 ? X : Q;
    double* dst = pass == 0 ? Q : X;
    RTA_TRY(dv_gram(src, n, k, src, n, k, n, G, k, scratch, s));
    RTA_HIP(hipMemcpyAsync(Gh, G, sizeof(double) * k * k, hipMemcpyDeviceToHost, s));
    RTA_HIP(hipStreamSynchronize(s));
    // the synchronization above also retired the previous pass's upload and
    // combine, so host and coef are free again
    int nv;
    if (pass == 0 && prescale) {
      // unit-normalize first: column i is dropped when G_ii <= lindep, the
      // others enter as x_i / |x_i| (G_ab -> G_ab / (|x_a| |x_b|)); C is
      // scaled back to the unnormalized columns, zero rows for the dropped
      std::vector<int> idx;
      std::vector<double> sc;
      for (int i = 0; i < k; ++i) {
        const double gii = Gh[i + (size_t)k * i];
        if (gii > lindep) {
          idx.push_back(i);
          sc.push_back(1.0 / std::sqrt(gii));
        }
      }
      const int m = (int)idx.size();
      std::vector<double> Gs((size_t)m * m), Cs((size_t)m * std::max(m, 1));
      for (int b = 0; b < m; ++b)
        for (int a = 0; a < m; ++a)
          Gs[a + (size_t)m * b] = sc[a] * Gh[idx[a] + (size_t)k * idx[b]] * sc[b];
      nv = m > 0 ? gs_gram_host(Gs.data(), m, lindep, Cs.data(), &min_frac) : 0;
      for (int j = 0; j < nv; ++j) {
        double* cj = Ch + (size_t)k * j;
        for (int r = 0; r < k; ++r) cj[r] = 0.0;
        for (int a = 0; a < m; ++a) cj[idx[a]] = sc[a] * Cs[a + (size_t)m * j];
      }
    } else {
      nv = gs_gram_host(Gh, k, lindep, Ch, &min_frac);
    }
    if (nv > 0) {
      RTA_HIP(hipMemcpyAsync(coef, Ch, sizeof(double) * k * nv, hipMemcpyHostToDevice, s));
      RTA_TRY(dv_combine(src, n, k, coef, k, nv, n, dst, n, s));
      if (pass == 1)
        RTA_HIP(hipMemcpyAsync(Q, X, sizeof(double) * n * nv, hipMemcpyDeviceToDevice, s));
    }
    k = nv;
  }
  *nout = k;
  return RT_OK;
}

// ---- RI linear-response operator --------------------------------------------------
// Z[P + naux*c] = sum_x B[x + ldb*P] X[x + n*c]   (8 aux columns per block)
template <int KMAX>
__global__ __launch_bounds__(256) void skinny_bt_x_kernel(
    const double* __restrict__ B, int64_t ldb, int64_t naux,
    const double* __restrict__ X, int64_t n, int k, double* __restrict__ Z) {
  constexpr int PG = 8;
  __shared__ double red[PG * KMAX][5];
  const int64_t P0 = (int64_t)blockIdx.x * PG;
  double acc[PG][KMAX];
#pragma unroll
  for (int p = 0; p < PG; ++p)
#pragma unroll
    for (int c = 0; c < KMAX; ++c) acc[p][c] = 0.0;
  for (int64_t x = threadIdx.x; x < n; x += 256) {
    double xv[KMAX];
#pragma unroll
    for (int c = 0; c < KMAX; ++c) xv[c] = (c < k) ? X[x + n * c] : 0.0;
#pragma unroll
    for (int p = 0; p < PG; ++p) {
      const double b = (P0 + p < naux) ? B[x + ldb * (P0 + p)] : 0.0;
#pragma unroll
      for (int c = 0; c < KMAX; ++c) acc[p][c] += b * xv[c];
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < PG; ++p)
#pragma unroll
    for (int c = 0; c < KMAX; ++c) {
      double v = dv_wave_sum(acc[p][c]);
      if (lane == 0) red[p * KMAX + c][wave] = v;
    }
  __syncthreads();
  for (int q = threadIdx.x; q < PG * KMAX; q += 256) {
    const int p = q / KMAX, c = q % KMAX;
    if (P0 + p < naux && c < k)
      Z[(P0 + p) + naux * c] = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
  }
}

// Ypart[s][x + n*c] = sum_{P in slice s} B[x + ldb*P] Z[P + naux*c]
template <int KMAX>
__global__ __launch_bounds__(256) void skinny_b_z_kernel(
    const double* __restrict__ B, int64_t ldb, int64_t naux,
    const double* __restrict__ Z, int64_t n, int k, int64_t pslice,
    double* __restrict__ Ypart) {
  const int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  const int64_t s = blockIdx.y;
  const int64_t P0 = s * pslice;
  const int64_t P1 = (P0 + pslice < naux) ? P0 + pslice : naux;
  double acc[KMAX];
#pragma unroll
  for (int c = 0; c < KMAX; ++c) acc[c] = 0.0;
  if (x < n) {
    const double* b = B + x;
    int64_t P = P0;
    for (; P + 3 < P1; P += 4) {
      const double b0 = b[ldb * P], b1 = b[ldb * (P + 1)], b2 = b[ldb * (P + 2)],
                   b3 = b[ldb * (P + 3)];
#pragma unroll
      for (int c = 0; c < KMAX; ++c) {
        if (c < k) {
          const double* z = Z + naux * c + P;
          acc[c] += b0 * z[0];
          acc[c] += b1 * z[1];
          acc[c] += b2 * z[2];
          acc[c] += b3 * z[3];
        }
      }
    }
    for (; P < P1; ++P) {
      const double bv = b[ldb * P];
#pragma unroll
      for (int c = 0; c < KMAX; ++c)
        if (c < k) acc[c] += bv * Z[naux * c + P];
    }
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
      if (c < k) Ypart[s * n * k + x + n * c] = acc[c];
  }
}

// 16-byte forms of the two passes (ldb and n even, B and the vectors 16-byte
// aligned): a lane takes the row pair (x, x+1).
// Zpart[r][P + naux*c] = sum over rows x of slice r of B[x + ldb*P] X[x + n*c]
// (8 aux columns per block, rows split into gridDim.y slices of `rows`)
template <int KMAX>
__global__ __launch_bounds__(256) void skinny_bt_x2_kernel(
    const double* __restrict__ B, int64_t ldb, int64_t naux,
    const double* __restrict__ X, int64_t n, int k, int64_t rows,
    double* __restrict__ Zpart) {
  constexpr int PG = 8;
  __shared__ double red[PG * KMAX][5];
  const int64_t P0 = (int64_t)blockIdx.x * PG;
  const int64_t r0 = (int64_t)blockIdx.y * rows;
  const int64_t r1 = (r0 + rows < n) ? r0 + rows : n;
  double acc[PG][KMAX];
#pragma unroll
  for (int p = 0; p < PG; ++p)
#pragma unroll
    for (int c = 0; c < KMAX; ++c) acc[p][c] = 0.0;
  for (int64_t x = r0 + 2 * threadIdx.x; x < r1; x += 512) {
    double2 xv[KMAX];
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
      xv[c] = (c < k) ? *reinterpret_cast<const double2*>(X + x + n * c) : make_double2(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < PG; ++p) {
      const double2 b = (P0 + p < naux)
                            ? *reinterpret_cast<const double2*>(B + x + ldb * (P0 + p))
                            : make_double2(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < KMAX; ++c) acc[p][c] += b.x * xv[c].x + b.y * xv[c].y;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < PG; ++p)
#pragma unroll
    for (int c = 0; c < KMAX; ++c) {
      double v = dv_wave_sum(acc[p][c]);
      if (lane == 0) red[p * KMAX + c][wave] = v;
    }
  __syncthreads();
  for (int q = threadIdx.x; q < PG * KMAX; q += 256) {
    const int p = q / KMAX, c = q % KMAX;
    if (P0 + p < naux && c < k)
      Zpart[(int64_t)blockIdx.y * naux * k + (P0 + p) + naux * c] =
          (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
  }
}

// Z[e] = sum_r Zpart[r][e], e < naux*k, in slice order
__global__ void skinny_zsum_kernel(const double* __restrict__ Zpart, int nr, int64_t m,
                                   double* __restrict__ Z) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= m) return;
  double s = 0.0;
  for (int r = 0; r < nr; ++r) s += Zpart[r * m + e];
  Z[e] = s;
}

// Ypart[s][x + n*c] = sum_{P in slice s} B[x + ldb*P] Z[P + naux*c], 8 slices
// of B in flight per lane
template <int KMAX>
__global__ __launch_bounds__(256) void skinny_b_z2_kernel(
    const double* __restrict__ B, int64_t ldb, int64_t naux,
    const double* __restrict__ Z, int64_t n, int k, int64_t pslice,
    double* __restrict__ Ypart) {
  const int64_t x = 2 * (blockIdx.x * (int64_t)256 + threadIdx.x);
  const int64_t s = blockIdx.y;
  const int64_t P0 = s * pslice;
  const int64_t P1 = (P0 + pslice < naux) ? P0 + pslice : naux;
  if (x >= n) return;
  double2 acc[KMAX];
#pragma unroll
  for (int c = 0; c < KMAX; ++c) acc[c] = make_double2(0.0, 0.0);
  const double* b = B + x;
  int64_t P = P0;
  for (; P + 7 < P1; P += 8) {
    double2 bv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) bv[u] = *reinterpret_cast<const double2*>(b + ldb * (P + u));
#pragma unroll
    for (int c = 0; c < KMAX; ++c) {
      if (c < k) {
        const double* z = Z + naux * c + P;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          acc[c].x += bv[u].x * z[u];
          acc[c].y += bv[u].y * z[u];
        }
      }
    }
  }
  for (; P < P1; ++P) {
    const double2 bv = *reinterpret_cast<const double2*>(b + ldb * P);
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
      if (c < k) {
        acc[c].x += bv.x * Z[naux * c + P];
        acc[c].y += bv.y * Z[naux * c + P];
      }
  }
#pragma unroll
  for (int c = 0; c < KMAX; ++c)
    if (c < k) *reinterpret_cast<double2*>(Ypart + s * n * k + x + n * c) = acc[c];
}

// The two passes over B for a block of kc <= 8 vectors: Z = B^T X, and the
// kBzSlices partial sums Ypart[s] of B Z over slices s of the aux index.
constexpr int kBzSlices = 16;   // partial slices of the second pass
constexpr int kBtxRowSlices = 4;

static bool skinny_x2(const double* B, int64_t ldb, const double* X, int64_t n) {
  return ldb % 2 == 0 && n % 2 == 0 && ((uintptr_t)B & 15) == 0 && ((uintptr_t)X & 15) == 0;
}

static int skinny_btx(const double* B, int64_t ldb, int64_t naux, const double* X, int64_t n,
                      int kc, double* Z, hipStream_t s) {
  KernelTimer kt("casida_btx", s);
  if (skinny_x2(B, ldb, X, n)) {
    double* Zpart;
    RTA_TRY(scratch_get(SCRATCH_DV_ZPART, sizeof(double) * kBtxRowSlices * naux * 8,
                        (void**)&Zpart));
    const int64_t rows = ((n + kBtxRowSlices - 1) / kBtxRowSlices + 1) / 2 * 2;
    hipLaunchKernelGGL(skinny_bt_x2_kernel<8>,
                       dim3((unsigned)((naux + 7) / 8), kBtxRowSlices), dim3(256), 0, s, B,
                       ldb, naux, X, n, kc, rows, Zpart);
    hipLaunchKernelGGL(skinny_zsum_kernel, dim3((unsigned)((naux * kc + 255) / 256)),
                       dim3(256), 0, s, Zpart, kBtxRowSlices, naux * kc, Z);
  } else {
    hipLaunchKernelGGL(skinny_bt_x_kernel<8>, dim3((unsigned)((naux + 7) / 8)), dim3(256), 0,
                       s, B, ldb, naux, X, n, kc, Z);
  }
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

static int skinny_bz(const double* B, int64_t ldb, int64_t naux, const double* Z, int64_t n,
                     int kc, double* Ypart, hipStream_t s) {
  KernelTimer kt("casida_bz", s);
  const int64_t pslice = (naux + kBzSlices - 1) / kBzSlices;
  if (skinny_x2(B, ldb, Ypart, n))
    hipLaunchKernelGGL(skinny_b_z2_kernel<8>, dim3((unsigned)((n + 511) / 512), kBzSlices),
                       dim3(256), 0, s, B, ldb, naux, Z, n, kc, pslice, Ypart);
  else
    hipLaunchKernelGGL(skinny_b_z_kernel<8>, dim3((unsigned)((n + 255) / 256), kBzSlices),
                       dim3(256), 0, s, B, ldb, naux, Z, n, kc, pslice, Ypart);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// Y[x + n*c] = d[x] X[x + n*c] + alpha * sum_s Ypart[s][x + n*c]
__global__ __launch_bounds__(256) void skinny_finish_kernel(
    const double* __restrict__ Ypart, int nslice, const double* __restrict__ d,
    const double* __restrict__ X, double alpha, int64_t n, int k,
    double* __restrict__ Y) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= n * k) return;
  const int64_t x = e % n;
  double s = 0.0;
  for (int i = 0; i < nslice; ++i) s += Ypart[(int64_t)i * n * k + e];
  Y[e] = d[x] * X[e] + alpha * s;
}

// hdiag[x] = d[x] + alpha * sum_P B[x + ldb*P]^2
__global__ __launch_bounds__(256) void casida_hdiag_kernel(
    const double* __restrict__ B, int64_t ldb, int64_t naux,
    const double* __restrict__ d, double alpha, int64_t n,
    double* __restrict__ hdiag) {
  const int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x >= n) return;
  double s = 0.0;
  for (int64_t P = 0; P < naux; ++P) {
    const double b = B[x + ldb * P];
    s += b * b;
  }
  hdiag[x] = d[x] + alpha * s;
}

int casida_apply(const CasidaOp* op, const double* X, double* Y, int64_t n,
                 int64_t k, hipStream_t s) {
  RTA_CHECK(n == op->n, "vector length %lld != operator dimension %lld",
            (long long)n, (long long)op->n);
  if (k <= 0) return RT_OK;
  const double alpha = 2.0 * op->scale * op->scale;
  for (int64_t c0 = 0; c0 < k; c0 += 8) {
    const int kc = (int)std::min<int64_t>(8, k - c0);
    const double* Xc = X + c0 * n;
    double* Yc = Y + c0 * n;
    double *Z, *Ypart;
    RTA_TRY(scratch_get(SCRATCH_DV_Z, sizeof(double) * op->naux * 8, (void**)&Z));
    RTA_TRY(scratch_get(SCRATCH_DV_YPART, sizeof(double) * kBzSlices * n * 8,
                        (void**)&Ypart));
    RTA_TRY(skinny_btx(op->B, op->ldb, op->naux, Xc, n, kc, Z, s));
    RTA_TRY(skinny_bz(op->B, op->ldb, op->naux, Z, n, kc, Ypart, s));
    hipLaunchKernelGGL(skinny_finish_kernel, dim3((unsigned)((n * kc + 255) / 256)),
                       dim3(256), 0, s, Ypart, kBzSlices, op->d, Xc, alpha, n, kc,
                       Yc);
    RTA_HIP(hipGetLastError());
  }
  return RT_OK;
}

int casida_hdiag(const CasidaOp* op, double* hdiag, hipStream_t s) {
  hipLaunchKernelGGL(casida_hdiag_kernel, dim3((unsigned)((op->n + 255) / 256)),
                     dim3(256), 0, s, op->B, op->ldb, op->naux, op->d,
                     2.0 * op->scale * op->scale, op->n, hdiag);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// ---- shift-invert preconditioner (Woodbury) -------------------------------------
__global__ void si_dm_kernel(const double* __restrict__ d, double sigma, int64_t n,
                             double* __restrict__ Dm) {
  const int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x < n) Dm[x] = 1.0 / (d[x] - sigma);
}

// Bs[x + n*P] = sqrt(Dm[x]) B[x + ldb*P]
__global__ void si_scale_rows_kernel(const double* __restrict__ B, int64_t ldb,
                                     const double* __restrict__ Dm, int64_t n,
                                     int64_t naux, double* __restrict__ Bs) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= n * naux) return;
  const int64_t x = e % n, P = e / n;
  Bs[e] = sqrt(Dm[x]) * B[x + ldb * P];
}

// M[i + m*i] += v
__global__ void si_add_diag_kernel(double* __restrict__ M, int64_t m, double v) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i < m) M[i + m * i] += v;
}

// E = I - T
__global__ void si_eye_minus_kernel(const double* __restrict__ T, int64_t m,
                                    double* __restrict__ E) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= m * m) return;
  E[e] = ((e % m == e / m) ? 1.0 : 0.0) - T[e];
}

// G = Dm o R (n x k)
__global__ void si_diag_apply_kernel(const double* __restrict__ Dm,
                                     const double* __restrict__ R, int64_t n,
                                     int64_t k, double* __restrict__ G) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e < n * k) G[e] = Dm[e % n] * R[e];
}

// T = G - Dm o sum_i Ypart[i]
__global__ void si_finish_kernel(const double* __restrict__ Ypart, int nslice,
                                 const double* __restrict__ Dm,
                                 const double* __restrict__ G, int64_t n, int k,
                                 double* __restrict__ T) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= n * k) return;
  double acc = 0.0;
  for (int i = 0; i < nslice; ++i) acc += Ypart[(int64_t)i * n * k + e];
  T[e] = G[e] - Dm[e % n] * acc;
}

static unsigned si_grid(int64_t len) { return (unsigned)((len + 255) / 256); }

int casida_si_setup(CasidaSI* si, double shift, hipStream_t s) {
  const CasidaOp& op = si->op;
  const int64_t n = op.n, m = op.naux;
  RTA_CHECK(shift > 0.0, "shift-invert shift must be positive");
  RTA_CHECK(op.scale != 0.0, "operator without a low-rank part");
  std::vector<double> dh(n);
  RTA_HIP(hipMemcpyAsync(dh.data(), op.d, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  RTA_HIP(hipStreamSynchronize(s));
  double dmin = dh[0];
  for (int64_t x = 1; x < n; ++x) dmin = std::min(dmin, dh[x]);
  si->sigma = dmin - shift;
  const double c = 2.0 * op.scale * op.scale;
  RTA_HIP(hipMalloc(&si->Dm, sizeof(double) * n));
  RTA_HIP(hipMalloc(&si->Ci, sizeof(double) * m * m));
  hipLaunchKernelGGL(si_dm_kernel, dim3(si_grid(n)), dim3(256), 0, s, op.d, si->sigma,
                     n, si->Dm);
  double *Bs = nullptr, *Cap = nullptr, *Xn = nullptr, *T = nullptr, *sq = nullptr,
         *part = nullptr;
  int st = RT_OK;
  auto alloc = [&](double** p, size_t cnt) {
    if (hipMalloc(p, sizeof(double) * cnt) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return false;
    }
    return true;
  };
  if (!alloc(&Bs, (size_t)n * m) || !alloc(&Cap, (size_t)m * m) ||
      !alloc(&Xn, (size_t)m * m) || !alloc(&T, (size_t)m * m) ||
      !alloc(&sq, 64 * 1024)) {
    st = fail(RT_ERR_NOMEM, "shift-invert setup: allocation failed");
  }
  double* X = si->Ci;
  double err = 0.0;
  if (st == RT_OK) {
    // Cap = I/c + Bs^T Bs, Bs = sqrt(Dm) o B
    hipLaunchKernelGGL(si_scale_rows_kernel, dim3(si_grid(n * m)), dim3(256), 0, s, op.B,
                       op.ldb, si->Dm, n, m, Bs);
    if (n % 16 == 0) {
      // symmetric: the sum over 16-row blocks b of Bs_b^T Bs_b on upper tiles
      // only, partial slots summed in a fixed order and mirrored
      GemmDesc d;
      d.ta = true;
      d.tb = false;
      d.M = d.N = m;
      d.K = 16;
      d.A = d.B = Bs;
      d.lda = d.ldb = n;
      d.sA = d.sB = 16;
      d.batch = n / 16;
      const int64_t bchunk = reduce_bchunk(m, 16, d.batch, 128);
      const int64_t nch = (d.batch + bchunk - 1) / bchunk;
      if (!alloc(&part, (size_t)nch * m * m))
        st = fail(RT_ERR_NOMEM, "shift-invert setup: allocation failed");
      else
        st = gemm_reduce_sym(d, bchunk, part, Cap, m, s, "si_cap", 128);
    } else {
      st = rt_dgemm_dev('T', 'N', m, m, n, 1.0, Bs, n, Bs, n, 0.0, Cap, m, s);
    }
  }
  if (st == RT_OK) {
    hipLaunchKernelGGL(si_add_diag_kernel, dim3(si_grid(m)), dim3(256), 0, s, Cap, m,
                       1.0 / c);
    // Ci = Cap^-1 by Cholesky (Cap >= I/c is positive definite): Cap = L L^T
    // in Xn, T = L^-1, Ci = T^T T
    RTA_HIP(hipMemcpyAsync(Xn, Cap, sizeof(double) * m * m, hipMemcpyDeviceToDevice, s));
    int info = 0;
    st = chol_lower(Xn, m, &info, reinterpret_cast<int*>(sq), s);
    if (st == RT_OK && info != 0)
      st = fail(RT_ERR_NUMERIC, "shift-invert: Cholesky of I/c + B^T Dm B failed at %d", info);
    if (st == RT_OK) st = tri_inverse_lower(Xn, m, T, s);
    if (st == RT_OK) st = rt_dgemm_dev('T', 'N', m, m, m, 1.0, T, m, T, m, 0.0, X, m, s);
  }
  if (st == RT_OK) {
    // check: |I - Cap Ci|_F
    st = rt_dgemm_dev('T', 'N', m, m, m, 1.0, Cap, m, X, m, 0.0, T, m, s);
    if (st == RT_OK) {
      hipLaunchKernelGGL(si_eye_minus_kernel, dim3(si_grid(m * m)), dim3(256), 0, s, T, m,
                         Xn);
      st = dv_sqnorms(Xn, 1, m * m, sq, sq + 1, s);
    }
    if (st == RT_OK) {
      RTA_HIP(hipMemcpyAsync(&err, sq, sizeof(double), hipMemcpyDeviceToHost, s));
      RTA_HIP(hipStreamSynchronize(s));
      err = std::sqrt(err);
      if (!(err < 1e-6)) st = fail(RT_ERR_NUMERIC, "shift-invert: |I - C Ci| = %.3e", err);
    }
  }
  (void)hipStreamSynchronize(s);
  if (Bs) (void)hipFree(Bs);
  if (Cap) (void)hipFree(Cap);
  if (Xn) (void)hipFree(Xn);
  if (T) (void)hipFree(T);
  if (sq) (void)hipFree(sq);
  if (part) (void)hipFree(part);
  si->inv_err = err;
  return st;
}

int casida_si_apply(const CasidaSI* si, const double* R, double* T, int64_t n,
                    int64_t k, hipStream_t s) {
  const CasidaOp& op = si->op;
  RTA_CHECK(n == op.n, "vector length %lld != operator dimension %lld", (long long)n,
            (long long)op.n);
  if (k <= 0) return RT_OK;
  const int64_t m = op.naux;
  double *G, *Z, *Ypart;
  // G (n x 8) first: 16-byte aligned like the scratch base, for the x2 passes
  RTA_TRY(scratch_get(SCRATCH_DV_Z, sizeof(double) * (n * 8 + 2 * m * 8), (void**)&G));
  Z = G + n * 8;
  double* W = Z + m * 8;
  RTA_TRY(scratch_get(SCRATCH_DV_YPART, sizeof(double) * kBzSlices * n * 8, (void**)&Ypart));
  for (int64_t c0 = 0; c0 < k; c0 += 8) {
    const int kc = (int)std::min<int64_t>(8, k - c0);
    hipLaunchKernelGGL(si_diag_apply_kernel, dim3(si_grid(n * kc)), dim3(256), 0, s,
                       si->Dm, R + c0 * n, n, (int64_t)kc, G);
    RTA_TRY(skinny_btx(op.B, op.ldb, m, G, n, kc, Z, s));
    RTA_TRY(skinny_btx(si->Ci, m, m, Z, m, kc, W, s));  // W = Ci^T Z (Ci: m x m)
    RTA_TRY(skinny_bz(op.B, op.ldb, m, W, n, kc, Ypart, s));
    hipLaunchKernelGGL(si_finish_kernel, dim3(si_grid(n * kc)), dim3(256), 0, s, Ypart,
                       kBzSlices, si->Dm, G, n, kc, T + c0 * n);
    RTA_HIP(hipGetLastError());
  }
  return RT_OK;
}

}  // namespace rta
