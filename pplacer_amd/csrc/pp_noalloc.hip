// Allocation-free variant of the ten_masked_logdot_c twin (pp_ten_masked_logdot)
// for callers that cannot allocate device memory: code under stream capture,
// tight serving loops.  Inputs and the output are the caller's device
// buffers; the per-site util vector the kernels need lives in `scratch`, a
// file-scope device array of kNoAllocMaxSites doubles that every GPU the
// library runs on holds until the process ends.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pplacer_hip.h"
#include "pp_internal.h"
#include "pp_twin_kernels.h"

namespace {

constexpr int kNoAllocMaxSites = 1 << 18;  // 2 MiB of doubles per device
constexpr int kMaxDevices = 64;

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

}  // namespace

// Per-site util of the latest call on this device (calls that overlap on one
// device race for it).
static __device__ __align__(256) double scratch[kNoAllocMaxSites];

namespace {

// The address of the current device's copy of `scratch`, looked up on the
// first call on that device (the name itself is a host-side stand-in).
int scratch_of_current_device(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    pp::set_error(std::string("hipGetDevice failed: ") + hipGetErrorString(e));
    return PP_EDEVICE;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    pp::set_error("device index beyond the allocation-free scratch table");
    return PP_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      pp::set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return PP_EDEVICE;
    }
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return PP_OK;
}

}  // namespace

extern "C" int pp_ten_masked_logdot_noalloc(const double *x, const double *y, const uint16_t *mask, int n_rates,
                                            int n_sites, int n_states, double *out, void *stream) {
  if (!x || !y || !mask || !out || n_rates <= 0 || n_sites < 0 || n_states <= 0) {
    pp::set_error("bad arguments");
    return PP_EINVAL;
  }
  if (n_sites > kNoAllocMaxSites) {
    pp::set_error("pp_ten_masked_logdot_noalloc: n_sites " + std::to_string(n_sites) +
                  " exceeds the allocation-free limit of " + std::to_string(kNoAllocMaxSites) + " sites");
    return PP_EINVAL;
  }
  double *util = nullptr;
  const int st = scratch_of_current_device(&util);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n_sites > 0)
    hipLaunchKernelGGL(pp::util_kernel, dim3((n_sites + 255) / 256), dim3(256), 0, s, util, nullptr, x, y, nullptr,
                       mask, 0, n_sites, n_rates, n_sites, n_states);
  hipLaunchKernelGGL(pp::logsum_kernel, dim3(1), dim3(64), 0, s, util, mask, 0, n_sites, n_rates, 1, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pp::set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return PP_EDEVICE;
  }
  return PP_OK;
}
