// RIFull ao2mo (A3/A4, A14), density (A6) and the seeded input generators.
//
// ao2mo (reference layout), per aux slice P (restmatr.f90:180-189):
//   T_P = A_P * C            (nbas x ns)      stage 1, batched over P
//   ri3mo(P,:,:) = C^T T_P   (ns x ns)        stage 2, stored aux-fastest
// ao2mo (occupied rows), per P:
//   X_P = C_occ^T A_P        (nocc x nbas)    stage 1
//   B_P = X_P C              (nocc x nmo)     stage 2, B[m + nocc*i + nocc*nmo*P]
// Both stages are FP64 MFMA batched GEMMs; P is processed in chunks so the
// intermediate stays within a fixed scratch budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "gemm.hpp"
#include "precision.hpp"
#include "rng.hpp"
#include "timer.hpp"

namespace rta {

static const size_t kAo2moScratchBytes = size_t(2) << 30;  // 2 GiB

__global__ void gen_ri3ao_kernel(double* __restrict__ out, int64_t nbas,
                                 int64_t naux, uint64_t key, double inv_decay) {
  const int64_t n2 = nbas * nbas, total = n2 * naux;
  const int64_t npair = nbas * (nbas + 1) / 2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t P = e / n2, r = e - P * n2;
    int64_t j = r / nbas, i = r - j * nbas;
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    uint64_t ctr = (uint64_t)(P * npair + hi * (hi + 1) / 2 + lo);
    double z = normal_from_counter(key, ctr);
    out[e] = z * exp(-(double)(hi - lo) * inv_decay);
  }
}

__global__ void gen_normal_kernel(double* __restrict__ out, int64_t n,
                                  uint64_t key) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = normal_from_counter(key, (uint64_t)e);
}

static unsigned gen_grid(int64_t total) {
  return (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((total + 255) / 256, 256 * 1024));
}

}  // namespace rta

using namespace rta;

extern "C" int rt_gen_ri3ao_dev(double* ri3ao, int64_t nbas, int64_t naux,
                                uint64_t seed, double decay, void* stream) {
  RTA_CHECK(nbas >= 0 && naux >= 0, "negative dimension");
  RTA_CHECK(decay > 0.0, "decay must be positive");
  if (nbas == 0 || naux == 0) return RT_OK;
  RTA_CHECK(ri3ao, "ri3ao is NULL");
  KernelTimer kt("gen_ri3ao", (hipStream_t)stream);
  hipLaunchKernelGGL(gen_ri3ao_kernel, dim3(gen_grid(nbas * nbas * naux)),
                     dim3(256), 0, (hipStream_t)stream, ri3ao, nbas, naux,
                     rng_key(seed), 1.0 / decay);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

extern "C" int rt_gen_normal_dev(double* out, int64_t n, uint64_t seed,
                                 void* stream) {
  RTA_CHECK(n >= 0, "negative length");
  if (n == 0) return RT_OK;
  RTA_CHECK(out, "out is NULL");
  hipLaunchKernelGGL(gen_normal_kernel, dim3(gen_grid(n)), dim3(256), 0,
                     (hipStream_t)stream, out, n, rng_key(seed));
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

extern "C" int rt_ao2mo_dev(const double* C, const double* ri3ao,
                            double* ri3mo, int64_t ns, int64_t nbas,
                            int64_t naux, void* stream) {
  RTA_CHECK(ns >= 0 && nbas >= 0 && naux >= 0, "negative dimension");
  RTA_CHECK(ns <= nbas, "num_states (%lld) > num_basis (%lld)", (long long)ns,
            (long long)nbas);
  if (ns == 0 || naux == 0) return RT_OK;
  RTA_CHECK(C && ri3ao && ri3mo, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  const int64_t tslice = nbas * ns;
  int64_t pc = std::max<int64_t>(
      1, std::min<int64_t>(naux, (int64_t)(kAo2moScratchBytes /
                                             (sizeof(double) * tslice))));
  double* T = nullptr;
  RTA_TRY(scratch_get(SCRATCH_AO2MO_T, sizeof(double) * tslice * pc,
                      (void**)&T));
  for (int64_t p0 = 0; p0 < naux; p0 += pc) {
    int64_t np = std::min(pc, naux - p0);
    // T_P = A_P^T C (k-major operands), so that ri3mo(P,:,:) = T_P^T C
    // equals C^T A_P C of restmatr.f90:183-188 for any (also non-symmetric) A_P
    GemmDesc s1;
    s1.ta = true;
    s1.tb = false;
    s1.M = nbas;
    s1.N = ns;
    s1.K = nbas;
    s1.A = ri3ao + p0 * nbas * nbas;
    s1.lda = nbas;
    s1.sA = nbas * nbas;
    s1.B = C;
    s1.ldb = nbas;
    s1.sB = 0;
    s1.C = T;
    s1.c_rs = 1;
    s1.c_cs = nbas;
    s1.sC = tslice;
    s1.batch = np;
    RTA_TRY(gemm_batched(s1, s, "ao2mo_s1"));
    GemmDesc s2;  // ri3mo(P,:,:) = T_P^T C, aux index fastest
    s2.ta = true;
    s2.tb = false;
    s2.M = ns;
    s2.N = ns;
    s2.K = nbas;
    s2.A = T;
    s2.lda = nbas;
    s2.sA = tslice;
    s2.B = C;
    s2.ldb = nbas;
    s2.sB = 0;
    s2.C = ri3mo + p0;
    s2.c_rs = naux;
    s2.c_cs = naux * ns;
    s2.sC = 1;
    s2.batch = np;
    RTA_TRY(gemm_batched(s2, s, "ao2mo_s2"));
  }
  return RT_OK;
}

extern "C" int rt_ao2mo_occ_dev(const double* C, int64_t nbas, int64_t nmo,
                                int64_t nocc, const double* ri3ao,
                                int64_t naux, double* B, int precision,
                                void* stream) {
  RTA_CHECK(nbas >= 0 && nmo >= 0 && nocc >= 0 && naux >= 0,
            "negative dimension");
  RTA_CHECK(nocc <= nmo && nmo <= nbas,
            "need nocc <= nmo <= nbas (got %lld, %lld, %lld)", (long long)nocc,
            (long long)nmo, (long long)nbas);
  if (nocc == 0 || nmo == 0 || naux == 0) return RT_OK;
  RTA_CHECK(C && ri3ao && B, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  if (precision != RT_PREC_FP64)
    return ao2mo_occ_lowp(C, nbas, nmo, nocc, ri3ao, naux, B, precision, s);
  const int64_t xslice = nocc * nbas;
  int64_t pc = std::max<int64_t>(
      1, std::min<int64_t>(naux, (int64_t)(kAo2moScratchBytes /
                                             (sizeof(double) * xslice))));
  double* X = nullptr;
  RTA_TRY(scratch_get(SCRATCH_AO2MO_T, sizeof(double) * xslice * pc,
                      (void**)&X));
  for (int64_t p0 = 0; p0 < naux; p0 += pc) {
    int64_t np = std::min(pc, naux - p0);
    // X^T_P = A_P^T C_occ  (nbas x nocc, k-major for the second stage)
    GemmDesc s1;
    s1.ta = true;
    s1.tb = false;
    s1.M = nbas;
    s1.N = nocc;
    s1.K = nbas;
    s1.A = ri3ao + p0 * nbas * nbas;
    s1.lda = nbas;
    s1.sA = nbas * nbas;
    s1.B = C;
    s1.ldb = nbas;
    s1.sB = 0;
    s1.C = X;
    s1.c_rs = 1;
    s1.c_cs = nbas;
    s1.sC = xslice;
    s1.batch = np;
    RTA_TRY(gemm_batched(s1, s, "ao2mo_occ_s1"));
    GemmDesc s2;  // B_P = (X^T_P)^T C
    s2.ta = true;
    s2.tb = false;
    s2.M = nocc;
    s2.N = nmo;
    s2.K = nbas;
    s2.A = X;
    s2.lda = nbas;
    s2.sA = xslice;
    s2.B = C;
    s2.ldb = nbas;
    s2.sB = 0;
    s2.C = B + p0 * nocc * nmo;
    s2.c_rs = 1;
    s2.c_cs = nocc;
    s2.sC = nocc * nmo;
    s2.batch = np;
    RTA_TRY(gemm_batched(s2, s, "ao2mo_occ_s2"));
  }
  return RT_OK;
}

extern "C" int rt_density_dev(const double* C, int64_t nbas, int64_t nocc,
                              double* D, void* stream) {
  RTA_CHECK(nbas >= 0 && nocc >= 0, "negative dimension");
  RTA_CHECK(nocc <= nbas, "nocc > nbas");
  if (nbas == 0) return RT_OK;
  RTA_CHECK(C && D, "NULL buffer");
  GemmDesc d;  // D = C_occ C_occ^T
  d.ta = false;
  d.tb = true;
  d.M = nbas;
  d.N = nbas;
  d.K = nocc;
  d.A = C;
  d.lda = nbas;
  d.B = C;
  d.ldb = nbas;
  d.C = D;
  d.c_rs = 1;
  d.c_cs = nbas;
  d.batch = 1;
  return gemm_batched(d, (hipStream_t)stream, "density");
}
