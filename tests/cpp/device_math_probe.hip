// Test probe of the device math helpers' rounding (recip_positive and
// rsqrt_positive of springback-precomp_amd/csrc/device_math.hpp), built into
// tests/cpp/_build/libdevice_math_probe.so for tests/test_device_math_gpu.py:
// for each argument x, the helpers' values beside the IEEE operations'.
#include <hip/hip_runtime.h>

#include "device_math.hpp"

namespace {

__global__ __launch_bounds__(256) void probe_kernel(const double* __restrict__ x, long long n,
                                                    double* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  out[4 * i + 0] = sbk::recip_positive(v);
  out[4 * i + 1] = 1.0 / v;
  out[4 * i + 2] = sbk::rsqrt_positive(v);
  out[4 * i + 3] = 1.0 / sqrt(v);
}

}  // namespace

// x: n doubles (device), out: 4 n doubles (device); returns a hipError_t.
extern "C" int device_math_probe(const double* x, long long n, double* out, void* stream) {
  if (n <= 0) return 0;
  const unsigned int blocks = static_cast<unsigned int>((n + 255) / 256);
  hipLaunchKernelGGL(probe_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), x, n, out);
  return static_cast<int>(hipGetLastError());
}
