// Reduced-precision forms of the RI-K build.  RT_PREC_OZAKI_BF16 and
// RT_PREC_BF16X3 slice B into BF16 planes (sliced.hpp) and run the K
// contraction on BF16 MFMA with FP64 accumulation across aux slices;
// RT_PREC_OZAKI_INT8 slices it into int8 planes for INT8 MFMA (ozaki_i8.hpp),
// RT_PREC_SPLIT_FP32 into hi/lo FP32 planes for FP32 MFMA (split32.hpp).
#include "common.hpp"
#include "ozaki_i8.hpp"
#include "precision.hpp"
#include "split32.hpp"
#include "sliced.hpp"

namespace rta {

int ao2mo_occ_lowp(const double*, int64_t, int64_t, int64_t, const double*,
                   int64_t, double*, int precision, hipStream_t) {
  return fail(RT_ERR_ARG, "ao2mo precision %d not available (FP64 only)",
              precision);
}

int rik_lowp(const double* B, int64_t nocc, int64_t nmo, int64_t naux,
             double* K, int precision, hipStream_t s) {
  if (precision == RT_PREC_OZAKI_INT8) return rik_ozaki_i8(B, nocc, nmo, naux, K, s);
  if (precision == RT_PREC_SPLIT_FP32) return rik_split32(B, nocc, nmo, naux, K, s);
  int ns;
  bool aligned;
  switch (precision) {
    case RT_PREC_OZAKI_BF16:
      ns = 4;
      aligned = true;
      break;
    case RT_PREC_BF16X3:
      ns = 2;
      aligned = false;
      break;
    default:
      return fail(RT_ERR_ARG, "RI-K precision %d not available", precision);
  }
  const int64_t noccp = (nocc + 31) / 32 * 32;
  uint16_t* S = nullptr;
  int32_t* E = nullptr;
  RTA_TRY(scratch_get(SCRATCH_LOWP_A,
                      sizeof(uint16_t) * (size_t)ns * noccp * nmo * naux,
                      (void**)&S));
  RTA_TRY(scratch_get(SCRATCH_LOWP_B, sizeof(int32_t) * (size_t)nmo * naux,
                      (void**)&E));
  RTA_TRY(slice_b(B, nocc, nmo, naux, ns, aligned, S, E, s));
  return rik_sliced(S, E, nocc, nmo, naux, ns, aligned, K, s);
}

}  // namespace rta
