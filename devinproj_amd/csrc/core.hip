// Library-wide state: error text, reduction workspace, version and device info.
#include "common.h"
#include "../../include/accusim_hip.h"

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <utility>

namespace accu {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

int Workspace::grow_error() {
  set_error("workspace allocation failed: %s", hipGetErrorString(hipGetLastError()));
  return ACCU_ERUNTIME;
}

int Workspace::ensure(size_t ndoubles) {
  if (!counter) {
    HIP_TRY(hipMalloc(&counter, 64));
    HIP_TRY(hipMemset(counter, 0, 64));
    HIP_TRY(hipDeviceSynchronize());  // the null-stream fill completes before any caller stream runs
    HIP_TRY(hipMalloc(&out, 16 * sizeof(double)));
  }
  if (ndoubles > cap) {
    size_t c = ndoubles < 4096 ? 4096 : ndoubles;
    if (grow(partials, cap, c)) return ACCU_ERUNTIME;
  }
  return ACCU_OK;
}

Workspace* workspace_for(hipStream_t stream) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, std::unique_ptr<Workspace>> all;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("no current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  std::unique_ptr<Workspace>& w = all[std::make_pair(dev, stream)];
  if (!w) {
    w.reset(new Workspace());
    w->dev = dev;
    w->st = stream;
  }
  return w.get();
}

}  // namespace accu

extern "C" {

const char* accu_last_error(void) { return accu::g_err; }

int accu_version(void) { return ACCU_ABI_VERSION; }

int accu_device_info(char* name, int len, int* cus, long long* hbm_bytes) {
  hipDeviceProp_t p;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipGetDeviceProperties(&p, dev));
  if (name && len > 0) {
    strncpy(name, p.gcnArchName, len - 1);
    name[len - 1] = 0;
  }
  if (cus) *cus = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (long long)p.totalGlobalMem;
  return ACCU_OK;
}

int accu_sync(void* stream) {
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return ACCU_OK;
}

}  // extern "C"
