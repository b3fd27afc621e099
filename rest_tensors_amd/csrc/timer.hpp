// Scoped per-launch timer: records HIP events around a kernel launch on its
// stream when rt_kernel_timer_enable(1) is in effect; otherwise does nothing.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace rta {

class KernelTimer {
 public:
  KernelTimer(const char* name, hipStream_t stream);
  ~KernelTimer();
  KernelTimer(const KernelTimer&) = delete;
  KernelTimer& operator=(const KernelTimer&) = delete;

 private:
  bool active_;
  hipStream_t stream_;
  hipEvent_t start_ = nullptr, stop_ = nullptr;
  std::string name_;
};

}  // namespace rta
