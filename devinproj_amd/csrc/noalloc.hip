// accu_gravity_direct without device allocation: the scratch of one call (energy or
// split partials and the grid-sum counter) is a fixed array in device memory.
#include "common.h"
#include "gravity.h"
#include "../../include/accusim_hip.h"

#include <mutex>

namespace accu {

// Largest scratch one allocation-free call may use, in doubles (8 MiB). The first
// kScratchHead doubles hold the grid-sum counter; the partials follow.
constexpr size_t kScratchDoubles = size_t(1) << 20;
constexpr size_t kScratchHead = 32;  // 256 bytes
static __device__ __align__(256) double scratch[kScratchHead + kScratchDoubles];

// Clears the grid-sum counter (a kernel rather than a memset, so that a caller may
// capture the call in a graph).
__global__ void k_noalloc_zero(unsigned* p) {
  if (threadIdx.x == 0) p[0] = 0u;
}

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device. The
// array's name is a host-side stand-in: only this address may be given to kernels.
static int scratch_on_current_device(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  ARG_CHECK(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ARG_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the gravity scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return ACCU_OK;
}

}  // namespace accu

using namespace accu;

extern "C" {

int accu_gravity_direct_noalloc(const double* pos, const double* mass, int n, const long long* targets, int ntgt,
                                double G, double eps2, const double* box, int periodic, int precision,
                                const double* origin, double* F, double* U, void* stream) {
  int rc = check_gravity_args(pos, mass, F, precision, periodic);
  if (rc) return rc;
  if (!targets) ntgt = n;
  if (n <= 0 || ntgt <= 0) return ACCU_OK;
  const GravPlan g = plan_gravity(n, ntgt, targets != nullptr, U != nullptr, precision);
  ARG_CHECK(g.need <= kScratchDoubles,
            "accu_gravity_direct_noalloc: %d targets of %d sources need %zu doubles of scratch, more than the "
            "fixed %zu", ntgt, n, g.need, kScratchDoubles);
  double* base = nullptr;
  rc = scratch_on_current_device(&base);
  if (rc) return rc;
  unsigned* counter = reinterpret_cast<unsigned*>(base);
  hipStream_t st = (hipStream_t)stream;
  if (U && g.nsplit == 0) {
    // the counter must be zero on entry
    hipLaunchKernelGGL(k_noalloc_zero, dim3(1), dim3(64), 0, st, counter);
    LAUNCH_CHECK();
  }
  return launch_gravity(g, pos, mass, n, targets, ntgt, G, eps2, box, periodic, precision, origin, F, U,
                        base + kScratchHead, counter, st);
}

}  // extern "C"
