// Reduced-precision forms of the GEMM-shaped stages (ao2mo, RI-K).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

// ao2mo (occupied rows) with precision in {SPLIT_FP32, BF16X3, OZAKI_BF16}.
int ao2mo_occ_lowp(const double* C, int64_t nbas, int64_t nmo, int64_t nocc,
                   const double* ri3ao, int64_t naux, double* B, int precision,
                   hipStream_t stream);

// K = sum_P B_P^T B_P (upper, mirrored) with precision != FP64.
int rik_lowp(const double* B, int64_t nocc, int64_t nmo, int64_t naux,
             double* K, int precision, hipStream_t stream);

}  // namespace rta
