// Shared host/device plumbing for librest_tensors_amd: status codes, the
// thread-local last-error string, HIP error checking and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/rest_tensors_amd.h"

namespace rta {

// Records a printf-style message as this thread's last error (read back with
// rt_last_error()).  Returns the status code so callers can `return fail(...)`.
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Prints the last error to stderr (used by the void drop-in symbols, which
// have no status channel).
void report_dropin_error(const char* symbol);

// Device-side scratch that grows on demand and is kept for the process
// lifetime (one per slot).  Slots let one entry point hold several buffers.
enum ScratchSlot {
  SCRATCH_GEMM_PARTIAL = 0,
  SCRATCH_AO2MO_T,
  SCRATCH_DROPIN_A,
  SCRATCH_DROPIN_B,
  SCRATCH_DROPIN_C,
  SCRATCH_DROPIN_D,
  SCRATCH_JK,
  SCRATCH_DV_Z,
  SCRATCH_DV_YPART,
  SCRATCH_DV_ZPART,
  SCRATCH_DV_PROJ,
  SCRATCH_LOWP_A,
  SCRATCH_LOWP_B,
  SCRATCH_MLP_A,
  SCRATCH_MLP_B,
  SCRATCH_NSLOTS
};
int scratch_get(int slot, size_t bytes, void** ptr);

}  // namespace rta

#define RTA_HIP(call)                                                         \
  do {                                                                        \
    hipError_t rta_e_ = (call);                                               \
    if (rta_e_ != hipSuccess)                                                 \
      return ::rta::fail(RT_ERR_HIP, "%s failed: %s (%s:%d)", #call,          \
                         hipGetErrorString(rta_e_), __FILE__, __LINE__);      \
  } while (0)

#define RTA_CHECK(cond, ...)                                                  \
  do {                                                                        \
    if (!(cond)) return ::rta::fail(RT_ERR_ARG, __VA_ARGS__);                 \
  } while (0)

#define RTA_TRY(expr)                                                         \
  do {                                                                        \
    int rta_s_ = (expr);                                                      \
    if (rta_s_ != RT_OK) return rta_s_;                                       \
  } while (0)

static inline unsigned rta_blocks(int64_t n, int per) {
  return (unsigned)((n + per - 1) / per);
}
