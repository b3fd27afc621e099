// Reduced-precision forms of ao2mo and RI-K (filled in below).
#include "common.hpp"
#include "precision.hpp"

namespace rta {

int ao2mo_occ_lowp(const double*, int64_t, int64_t, int64_t, const double*,
                   int64_t, double*, int precision, hipStream_t) {
  return fail(RT_ERR_ARG, "ao2mo precision %d not available", precision);
}

int rik_lowp(const double*, int64_t, int64_t, int64_t, double*, int precision,
             hipStream_t) {
  return fail(RT_ERR_ARG, "RI-K precision %d not available", precision);
}

}  // namespace rta
