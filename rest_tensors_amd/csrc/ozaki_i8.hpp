// RI-K on INT8 MFMA with Ozaki-scheme slicing (RT_PREC_OZAKI_INT8); the
// arithmetic is described at the top of ozaki_i8.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

// Largest number of contraction terms (m, P) that share one exponent and one
// int32 accumulation: 4 products x 8192 terms x 127^2 < 2^29.
constexpr int64_t kOzMaxGroupK = 8192;

// Upper bound of the FP64 partial slots of one launch (one nmo x nmo slot per
// exponent group); larger problems take several launches.
constexpr int64_t kOzMaxSlotBytes = (int64_t)2 << 30;

// Aux slices per exponent group: as many as the int32 bound allows, fewer when
// that would leave CUs idle.
int64_t oz_group_size(int64_t nocc, int64_t nmo, int64_t naux, int64_t ncu);

// Int8 slices S (byte ((P*nmo + i)*nkc + kc)*256 + s*64 + m%64, nkc =
// ceil(nocc/64)) and exponents E[g*nmo + i] of B [nocc, nmo, naux].
int oz_slice(const double* B, int64_t nocc, int64_t nmo, int64_t naux, int64_t gp,
             int8_t* S, int32_t* E, hipStream_t s);

// K = sum_P B_P^T B_P (symmetric, both triangles written).
int rik_ozaki_i8(const double* B, int64_t nocc, int64_t nmo, int64_t naux, double* K,
                 hipStream_t s);

}  // namespace rta
