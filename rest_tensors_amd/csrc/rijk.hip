// RI-J (A13) and RI-K / E_x (A14).
//
// RI-J is two HBM-bound streams over the upper triangle of ri3ao
// (ri3ao is (mu nu|P), symmetric in mu,nu):
//   gamma_P = sum_{k<=l} w_kl A[k,l,P] D[k,l]      (w = 2 off-diagonal, 1 on it)
//   J[i,j]  = sum_P A[i,j,P] gamma_P               (i <= j, then mirrored)
// Algorithmic traffic: 8*naux*nbas(nbas+1) bytes (two upper-triangle passes).
//
// RI-K is an MFMA-bound batch-reducing GEMM:
//   K = sum_P B_P^T B_P,  B_P = ri3mo(occ, all, P) (nocc x nmo, contiguous)
// computed on upper tiles only, summed over aux chunks in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "precision.hpp"
#include "timer.hpp"

namespace rta {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One block per pair of aux slices (P, P+1); wave w walks columns j = w,
// w+4, ...; lanes walk rows of that column.  Each D element read serves both
// slices.  Fixed order -> deterministic.
__global__ __launch_bounds__(256) void rij_gamma_kernel(
    const double* __restrict__ A, const double* __restrict__ D, int64_t nbas,
    int64_t naux, double* __restrict__ gamma) {
  __shared__ double red[2][4];
  const int64_t P = 2 * (int64_t)blockIdx.x;
  const int64_t n2 = nbas * nbas;
  const bool two = P + 1 < naux;
  const double* Ap = A + P * n2;
  const double* Aq = two ? Ap + n2 : Ap;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc = 0.0, acc2 = 0.0;
  for (int64_t j = wave; j < nbas; j += 4) {
    const double* cp = Ap + j * nbas;
    const double* cq = Aq + j * nbas;
    const double* dcol = D + j * nbas;
    for (int64_t i = lane; i <= j; i += 64) {
      const double dw = (i == j) ? dcol[i] : 2.0 * dcol[i];
      acc += cp[i] * dw;
      acc2 += cq[i] * dw;
    }
  }
  acc = wave_sum(acc);
  acc2 = wave_sum(acc2);
  if (lane == 0) {
    red[0][wave] = acc;
    red[1][wave] = acc2;
  }
  __syncthreads();
  if (threadIdx.x == 0) gamma[P] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
  if (threadIdx.x == 1 && two)
    gamma[P + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
}

// The same sum for an even nbas and 16-byte aligned A and D: one block per SP
// aux slices, each D element read once for all SP.  A lane takes the pair of
// rows (2t, 2t+1) of a column with one 16-byte load per slice (weights 2 above
// the diagonal, 1 on it, 0 for row j+1 of an even column j, which the even
// nbas keeps inside the column); wave w walks columns j = w, w+4, ... in
// chunks of 64 pairs, the loads of the next chunk issued before the current
// one is summed.  Fixed order -> deterministic.
template <int SP>
__global__ __launch_bounds__(256) void rij_gamma_x2_kernel(
    const double* __restrict__ A, const double* __restrict__ D, int64_t nbas,
    int64_t naux, double* __restrict__ gamma) {
  __shared__ double red[SP][4];
  const int64_t P0 = (int64_t)blockIdx.x * SP;
  const int64_t n2 = nbas * nbas, ldh = nbas / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double2* Ab[SP];
#pragma unroll
  for (int q = 0; q < SP; ++q)
    Ab[q] = reinterpret_cast<const double2*>(A + (P0 + q < naux ? P0 + q : P0) * n2);
  const double2* D2 = reinterpret_cast<const double2*>(D);
  double acc[SP];
#pragma unroll
  for (int q = 0; q < SP; ++q) acc[q] = 0.0;
  auto load = [&](int64_t j, int64_t u, double2* a, double2& d) {
    if (2 * (u + lane) <= j) {
      const int64_t off = j * ldh + u + lane;
      d = D2[off];
#pragma unroll
      for (int q = 0; q < SP; ++q) a[q] = Ab[q][off];
    } else {
      d = make_double2(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < SP; ++q) a[q] = make_double2(0.0, 0.0);
    }
  };
  int64_t j = wave, u = 0;
  double2 a[SP], d;
  if (j < nbas) load(j, u, a, d);
  while (j < nbas) {
    int64_t jn = j, un = u + 64;
    if (2 * un > jn) {
      un = 0;
      jn += 4;
    }
    double2 an[SP], dn = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < SP; ++q) an[q] = dn;
    if (jn < nbas) load(jn, un, an, dn);
    const int64_t r = 2 * (u + lane);
    const double d0 = (r < j ? 2.0 : (r == j ? 1.0 : 0.0)) * d.x;
    const double d1 = (r + 1 < j ? 2.0 : (r + 1 == j ? 1.0 : 0.0)) * d.y;
#pragma unroll
    for (int q = 0; q < SP; ++q) acc[q] += a[q].x * d0 + a[q].y * d1;
    j = jn;
    u = un;
    d = dn;
#pragma unroll
    for (int q = 0; q < SP; ++q) a[q] = an[q];
  }
#pragma unroll
  for (int q = 0; q < SP; ++q) acc[q] = wave_sum(acc[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < SP; ++q) red[q][wave] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < SP && P0 + threadIdx.x < naux) {
    const int q = threadIdx.x;
    gamma[P0 + q] = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
  }
}

// Block = (column j, 64-row segment s of rows 0..j); its 4 waves take every
// 4th aux slice (8 loads in flight per lane; the wave-uniform gamma index
// lets gamma come through scalar loads) and are combined through LDS.
// Writes J[i,j] and J[j,i].
__global__ __launch_bounds__(256) void rij_j_kernel(
    const double* __restrict__ A, const double* __restrict__ gamma,
    int64_t nbas, int64_t naux, double* __restrict__ J) {
  __shared__ double red[4][64];
  const int64_t j = blockIdx.x;
  const int64_t i0 = (int64_t)blockIdx.y * 64;
  if (i0 > j) return;  // segment entirely below the diagonal
  const double* gm = gamma;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t i = i0 + lane;
  const int64_t n2 = nbas * nbas;
  double acc = 0.0, acc2 = 0.0;
  if (i <= j) {
    const double* p = A + i + j * nbas;
    int64_t P = wave;
    for (; P + 28 < naux; P += 32) {
      double a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = p[(P + 4 * u) * n2];
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        acc += a[u] * gm[P + 4 * u];
        acc2 += a[u + 1] * gm[P + 4 * u + 4];
      }
    }
    for (; P < naux; P += 4) acc += p[P * n2] * gm[P];
  }
  red[wave][lane] = acc + acc2;
  __syncthreads();
  if (wave == 0 && i <= j) {
    double v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    J[i + j * nbas] = v;
    if (i != j) J[j + i * nbas] = v;
  }
}

// out[0] = sum_{n<nocc} K[n,n]  (single block, fixed order)
__global__ __launch_bounds__(256) void diag_sum_kernel(
    const double* __restrict__ K, int64_t ld, int64_t n,
    double* __restrict__ out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += 256) acc += K[k + k * ld];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = (red[0] + red[1]) + (red[2] + red[3]);
}

// partial[P] = sum_{m,n<nocc} B[m,n,P]^2 over the occ-occ block of slice P
__global__ __launch_bounds__(256) void occ_sumsq_kernel(
    const double* __restrict__ B, int64_t nocc, int64_t nmo,
    double* __restrict__ partial) {
  __shared__ double red[4];
  const int64_t P = blockIdx.x;
  const double* b = B + P * nocc * nmo;  // occ-occ block is the first nocc^2
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < nocc * nocc; e += 256) acc += b[e] * b[e];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[P] = (red[0] + red[1]) + (red[2] + red[3]);
}

int rik_fp64(const double* B, int64_t nocc, int64_t nmo, int64_t naux,
             double* K, hipStream_t s) {
  const int tile = 128;
  int64_t bchunk = reduce_bchunk(nmo, nocc, naux, tile);
  int64_t nchunks = (naux + bchunk - 1) / bchunk;
  double* part = nullptr;
  RTA_TRY(scratch_get(SCRATCH_GEMM_PARTIAL,
                      sizeof(double) * std::max<int64_t>(nchunks, 1) * nmo * nmo,
                      (void**)&part));
  GemmDesc d;
  d.ta = true;  // op(A)(i,m) = B_P[m + i*nocc]
  d.tb = false; // op(B)(m,j) = B_P[m + j*nocc]
  d.M = nmo;
  d.N = nmo;
  d.K = nocc;
  d.A = B;
  d.lda = nocc;
  d.sA = nocc * nmo;
  d.B = B;
  d.ldb = nocc;
  d.sB = nocc * nmo;
  d.alpha = 1.0;
  d.batch = naux;
  return gemm_reduce_sym(d, bchunk, part, K, nmo, s, "rik_mfma", tile);
}

}  // namespace rta

using namespace rta;

extern "C" int rt_rij_dev(const double* ri3ao, int64_t nbas, int64_t naux,
                          const double* D, double* J, double* gamma,
                          double* e_j, void* stream) {
  RTA_CHECK(nbas >= 0 && naux >= 0, "negative dimension");
  RTA_CHECK(nbas < (int64_t(1) << 31) / 64 && naux < (int64_t(1) << 31),
            "dimension too large");
  hipStream_t s = (hipStream_t)stream;
  if (nbas == 0) {
    if (e_j) *e_j = 0.0;
    return RT_OK;
  }
  RTA_CHECK(D && J, "NULL buffer");
  RTA_CHECK(naux == 0 || ri3ao, "ri3ao is NULL");
  double* g = gamma;
  if (!g && naux > 0)
    RTA_TRY(scratch_get(SCRATCH_JK, sizeof(double) * naux, (void**)&g));
  if (naux > 0) {
    KernelTimer kt("rij_gamma", s);
    const bool x2 = nbas % 2 == 0 && ((uintptr_t)ri3ao & 15) == 0 && ((uintptr_t)D & 15) == 0;
    if (x2)
      hipLaunchKernelGGL(rij_gamma_x2_kernel<8>, dim3((unsigned)((naux + 7) / 8)), dim3(256), 0,
                         s, ri3ao, D, nbas, naux, g);
    else
      hipLaunchKernelGGL(rij_gamma_kernel, dim3((unsigned)((naux + 1) / 2)), dim3(256), 0, s,
                         ri3ao, D, nbas, naux, g);
    RTA_HIP(hipGetLastError());
  }
  {
    KernelTimer kt("rij_j", s);
    hipLaunchKernelGGL(rij_j_kernel,
                       dim3((unsigned)nbas, (unsigned)((nbas + 63) / 64)),
                       dim3(256), 0, s, ri3ao, g, nbas, naux, J);
    RTA_HIP(hipGetLastError());
  }
  if (e_j) {
    std::vector<double> h(naux);
    if (naux > 0) {
      RTA_HIP(hipMemcpyAsync(h.data(), g, sizeof(double) * naux,
                             hipMemcpyDeviceToHost, s));
      RTA_HIP(hipStreamSynchronize(s));
    }
    double acc = 0.0;
    for (int64_t P = 0; P < naux; ++P) acc += h[P] * h[P];
    *e_j = 0.5 * acc;
  }
  return RT_OK;
}

extern "C" int rt_rik_dev(const double* B, int64_t nocc, int64_t nmo,
                          int64_t naux, double* K, double* e_x, int precision,
                          void* stream) {
  RTA_CHECK(nocc >= 0 && nmo >= 0 && naux >= 0, "negative dimension");
  RTA_CHECK(nocc <= nmo, "nocc > nmo");
  hipStream_t s = (hipStream_t)stream;
  if (nmo == 0) {
    if (e_x) *e_x = 0.0;
    return RT_OK;
  }
  RTA_CHECK(K, "K is NULL");
  RTA_CHECK(B || nocc == 0 || naux == 0, "B is NULL");
  if (precision == RT_PREC_FP64)
    RTA_TRY(rik_fp64(B, nocc, nmo, naux, K, s));
  else
    RTA_TRY(rik_lowp(B, nocc, nmo, naux, K, precision, s));
  if (e_x) {
    double* dsum = nullptr;
    RTA_TRY(scratch_get(SCRATCH_JK, sizeof(double), (void**)&dsum));
    hipLaunchKernelGGL(diag_sum_kernel, dim3(1), dim3(256), 0, s, K, nmo, nocc,
                       dsum);
    RTA_HIP(hipGetLastError());
    double h = 0.0;
    RTA_HIP(hipMemcpyAsync(&h, dsum, sizeof(double), hipMemcpyDeviceToHost, s));
    RTA_HIP(hipStreamSynchronize(s));
    *e_x = -0.5 * h;
  }
  return RT_OK;
}

extern "C" int rt_exchange_energy_dev(const double* B, int64_t nocc,
                                      int64_t nmo, int64_t naux, double* e_x,
                                      void* stream) {
  RTA_CHECK(nocc >= 0 && nmo >= 0 && naux >= 0 && nocc <= nmo,
            "bad dimensions");
  RTA_CHECK(e_x, "e_x is NULL");
  hipStream_t s = (hipStream_t)stream;
  if (nocc == 0 || naux == 0) {
    *e_x = 0.0;
    return RT_OK;
  }
  double* part = nullptr;
  RTA_TRY(scratch_get(SCRATCH_JK, sizeof(double) * naux, (void**)&part));
  hipLaunchKernelGGL(occ_sumsq_kernel, dim3((unsigned)naux), dim3(256), 0, s,
                     B, nocc, nmo, part);
  RTA_HIP(hipGetLastError());
  std::vector<double> h(naux);
  RTA_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * naux,
                         hipMemcpyDeviceToHost, s));
  RTA_HIP(hipStreamSynchronize(s));
  double acc = 0.0;
  for (int64_t P = 0; P < naux; ++P) acc += h[P];
  *e_x = -0.5 * acc;
  return RT_OK;
}
