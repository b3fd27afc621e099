// Sliced-precision K build: K = sum_P B_P^T B_P from low-precision slices
// of B on BF16 MFMA (v_mfma_f32_16x16x32_bf16), FP32 accumulation per aux
// slice, FP64 accumulation across aux slices.
//
// Slicing of one column v = B[:, i, P] (nocc values):
//   aligned (RT_PREC_OZAKI_BF16): e = exponent with max|v| < 2^e,
//     u = v 2^-e, sigma_0 = rint(128 u), then sigma_s = rint(256 * remainder)
//     for s >= 1, so  v = 2^(e-7) sum_s 2^(-8s) sigma_s  up to 2^-31 relative;
//     every |sigma_s| <= 128 is an integer, exact in BF16.  A product level
//     q = s + t sums (q+1) slice products over nocc <= 256 terms, each
//     |sigma_s sigma_t| <= 2^14: the FP32 sum stays below 2^24 and is exact.
//     The per-P fold  sum_q 2^(-8q) acc_q  is exact in FP64 and is scaled by
//     the exact power of two 2^(e_i + e_j - 14).  Levels q <= NS-1 are kept.
//   unaligned (RT_PREC_BF16X3): h = bf16(v), l = bf16(v - h); levels
//     hh and hl + lh, FP32 accumulation (rounded), fold with scale 1.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

// Slice layout: S[((P*nmo + i)*NS + s)*nocc + m] (bf16 bits), exponents
// E[P*nmo + i] (int32, 0 for unaligned slicing).
int slice_b(const double* B, int64_t nocc, int64_t nmo, int64_t naux, int ns,
            bool aligned, uint16_t* S, int32_t* E, hipStream_t s);

int rik_sliced(const uint16_t* S, const int32_t* E, int64_t nocc, int64_t nmo,
               int64_t naux, int ns, bool aligned, double* K, hipStream_t s);

// K[m,n] = K[n,m] = [K[m,n] +] sum_{k<nslot} part[k*M*M + m + n*M] for m <= n
// (fixed order), reading only the upper triangle of each M x M slot.
int reduce_partials_sym(const double* part, int64_t nslot, int64_t M, double* K,
                        bool accumulate, hipStream_t s);

}  // namespace rta
