// The element evaluation without a caller workspace and without any device
// allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// sparlab_elements_evaluate_device passes its kernels a caller workspace
// (each point's stress and tangent of its final local pass, each element's
// K_aa factors, the schedule of the local solve). Here that workspace is the
// file-scope device array `scratch`, one per GPU, held until the process
// ends. Host code never uses the array's name as a device address: on the
// first call on a device, with that device current, hipGetSymbolAddress
// gives that device's copy, and the kernels take that address as they take
// the caller's workspace. Every call on one GPU shares its copy, so calls
// that overlap on one GPU race; the array keeps the last call's contents,
// and nothing in it needs to start at zero (the schedule's counters are
// cleared by the first kernel of each call, every record is written before
// it is read).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "element_kernels.hpp"
#include "sparlab_gpu.h"

namespace sbk {
int capi_fail(int code, const std::string& message);
int capi_check_element_batch(const sparlab_material* material, const sparlab_rule* rule, int kinematics,
                             int want_tangent, const sparlab_element_batch* b, int stages, ElementArgs& a);
int capi_launch_element_batch(int want_tangent, const sparlab_element_batch* b, double* workspace, int stages,
                              void* stream, ElementArgs& a);
}  // namespace sbk

namespace {

constexpr long long kScratchDoubles = SPARLAB_NOALLOC_SCRATCH_BYTES / sizeof(double);
static_assert(SPARLAB_NOALLOC_SCRATCH_BYTES % 256 == 0, "the scratch size is a multiple of its alignment");

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
double* g_scratch[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int device_scratch(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "no current HIP device (or more than 64 devices)");
  std::lock_guard<std::mutex> hold(g_lock);
  if (g_scratch[dev] == nullptr) {
    void* ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || ptr == nullptr)
      return sbk::capi_fail(SPARLAB_E_DEVICE, std::string("the scratch array's device address is unavailable: ") +
                                                  hipGetErrorString(err));
    g_scratch[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch[dev];
  return SPARLAB_OK;
}

}  // namespace

extern "C" int sparlab_elements_evaluate_noalloc_device(const sparlab_material* material, const sparlab_rule* rule,
                                                        int kinematics, int want_tangent,
                                                        const sparlab_element_batch* batch, void* stream) {
  sbk::ElementArgs a{};
  const int st = sbk::capi_check_element_batch(material, rule, kinematics, want_tangent, batch, 3, a);
  if (st != SPARLAB_OK) return st;
  if (batch->num_elements == 0) return SPARLAB_OK;
  const long long need = sbk::workspace_doubles(batch->num_elements, a.rule.n);
  if (need > kScratchDoubles)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT,
                          "the batch needs " + std::to_string(need * 8) + " bytes of scratch, more than the " +
                              std::to_string(static_cast<long long>(SPARLAB_NOALLOC_SCRATCH_BYTES)) +
                              " the no-allocation entry point holds; split it into smaller batches");
  double* ws = nullptr;
  const int sc = device_scratch(&ws);
  if (sc != SPARLAB_OK) return sc;
  return sbk::capi_launch_element_batch(want_tangent, batch, ws, 3, stream, a);
}
