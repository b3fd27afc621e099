// RI-K in split FP32 (RT_PREC_SPLIT_FP32); the arithmetic is described at the
// top of split32.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

// K = sum_P B_P^T B_P (symmetric, both triangles written).
int rik_split32(const double* B, int64_t nocc, int64_t nmo, int64_t naux, double* K,
                hipStream_t s);

}  // namespace rta
