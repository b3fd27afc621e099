// Runtime plumbing of librest_tensors_amd: error strings, device memory,
// streams, events, the per-kernel timer and grow-on-demand scratch buffers.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "timer.hpp"

namespace rta {

static thread_local std::string g_last_error;

int fail(int status, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

void report_dropin_error(const char* symbol) {
  fprintf(stderr, "librest_tensors_amd: %s: %s\n", symbol, g_last_error.c_str());
  fflush(stderr);
}

// ---- scratch -------------------------------------------------------------
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};
static std::mutex g_scratch_mu;
static std::map<int, std::vector<Scratch>> g_scratch;  // per device

int scratch_get(int slot, size_t bytes, void** ptr) {
  int dev = 0;
  RTA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto& v = g_scratch[dev];
  if (v.empty()) v.resize(SCRATCH_NSLOTS);
  Scratch& s = v[slot];
  if (s.bytes < bytes) {
    if (s.ptr) {
      // the old buffer may still be in use by queued work on any stream
      RTA_HIP(hipDeviceSynchronize());
      RTA_HIP(hipFree(s.ptr));
      s.ptr = nullptr;
      s.bytes = 0;
    }
    size_t want = bytes + bytes / 8;  // headroom against repeated regrowth
    hipError_t e = hipMalloc(&s.ptr, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      s.ptr = nullptr;
      return fail(RT_ERR_NOMEM, "scratch allocation of %zu bytes failed: %s",
                  want, hipGetErrorString(e));
    }
    s.bytes = want;
  }
  *ptr = s.ptr;
  return RT_OK;
}

// ---- kernel timer ----------------------------------------------------------
struct TimerRec {
  std::string name;
  hipEvent_t start, stop;
};
static std::mutex g_timer_mu;
static bool g_timer_on = false;
static std::vector<TimerRec> g_timer_recs;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t pool_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

KernelTimer::KernelTimer(const char* name, hipStream_t stream)
    : active_(false), stream_(stream) {
  std::lock_guard<std::mutex> lk(g_timer_mu);
  if (!g_timer_on) return;
  start_ = pool_event();
  stop_ = pool_event();
  if (!start_ || !stop_) return;
  name_ = name;
  active_ = true;
  (void)hipEventRecord(start_, stream_);
}

KernelTimer::~KernelTimer() {
  if (!active_) return;
  (void)hipEventRecord(stop_, stream_);
  std::lock_guard<std::mutex> lk(g_timer_mu);
  g_timer_recs.push_back({name_, start_, stop_});
}

}  // namespace rta

using namespace rta;

extern "C" {

const char* rt_last_error(void) { return g_last_error.c_str(); }
void rt_clear_error(void) { g_last_error.clear(); }

int rt_version(void) { return 100; }

int rt_device_count(int* count) {
  RTA_CHECK(count, "count is NULL");
  RTA_HIP(hipGetDeviceCount(count));
  return RT_OK;
}

int rt_set_device(int device) {
  RTA_HIP(hipSetDevice(device));
  return RT_OK;
}

int rt_device_name(char* buf, int buflen) {
  RTA_CHECK(buf && buflen > 0, "bad buffer");
  int dev = 0;
  RTA_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  RTA_HIP(hipGetDeviceProperties(&p, dev));
  snprintf(buf, buflen, "%s (%s, %d CUs)", p.name, p.gcnArchName,
           p.multiProcessorCount);
  return RT_OK;
}

int rt_malloc(void** dptr, size_t bytes) {
  RTA_CHECK(dptr, "dptr is NULL");
  *dptr = nullptr;
  if (bytes == 0) return RT_OK;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return fail(RT_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes,
                hipGetErrorString(e));
  }
  return RT_OK;
}

int rt_free(void* dptr) {
  if (dptr) RTA_HIP(hipFree(dptr));
  return RT_OK;
}

int rt_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  if (bytes) RTA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return RT_OK;
}

int rt_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  if (bytes) RTA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return RT_OK;
}

int rt_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  if (bytes) RTA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return RT_OK;
}

int rt_memset(void* dst, int value, size_t bytes) {
  if (bytes) RTA_HIP(hipMemset(dst, value, bytes));
  return RT_OK;
}

int rt_synchronize(void) {
  RTA_HIP(hipDeviceSynchronize());
  return RT_OK;
}

int rt_stream_create(void** stream) {
  RTA_CHECK(stream, "stream is NULL");
  hipStream_t s;
  RTA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return RT_OK;
}

int rt_stream_destroy(void* stream) {
  if (stream) RTA_HIP(hipStreamDestroy((hipStream_t)stream));
  return RT_OK;
}

int rt_stream_synchronize(void* stream) {
  RTA_HIP(hipStreamSynchronize((hipStream_t)stream));
  return RT_OK;
}

int rt_event_create(void** event) {
  RTA_CHECK(event, "event is NULL");
  hipEvent_t e;
  RTA_HIP(hipEventCreate(&e));
  *event = (void*)e;
  return RT_OK;
}

int rt_event_destroy(void* event) {
  if (event) RTA_HIP(hipEventDestroy((hipEvent_t)event));
  return RT_OK;
}

int rt_event_record(void* event, void* stream) {
  RTA_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return RT_OK;
}

int rt_event_elapsed_ms(void* start, void* stop, float* ms) {
  RTA_CHECK(ms, "ms is NULL");
  RTA_HIP(hipEventSynchronize((hipEvent_t)stop));
  RTA_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return RT_OK;
}

int rt_kernel_timer_enable(int on) {
  std::lock_guard<std::mutex> lk(g_timer_mu);
  g_timer_on = on != 0;
  return RT_OK;
}

int rt_kernel_timer_reset(void) {
  std::lock_guard<std::mutex> lk(g_timer_mu);
  for (auto& r : g_timer_recs) {
    (void)hipEventSynchronize(r.stop);
    g_event_pool.push_back(r.start);
    g_event_pool.push_back(r.stop);
  }
  g_timer_recs.clear();
  return RT_OK;
}

int rt_kernel_timer_get(const char* prefix, double* total_ms, int64_t* count) {
  RTA_CHECK(prefix && total_ms && count, "NULL argument");
  std::lock_guard<std::mutex> lk(g_timer_mu);
  double tot = 0.0;
  int64_t n = 0;
  size_t plen = strlen(prefix);
  for (auto& r : g_timer_recs) {
    if (r.name.compare(0, plen, prefix) != 0) continue;
    RTA_HIP(hipEventSynchronize(r.stop));
    float ms = 0.f;
    RTA_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    tot += ms;
    ++n;
  }
  *total_ms = tot;
  *count = n;
  return RT_OK;
}

}  // extern "C"
