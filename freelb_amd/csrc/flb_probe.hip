// Device copy-bandwidth probe: one float4 per thread, no loop, the access
// pattern that measures highest on this hardware. Used to calibrate the
// roofline figures and the HBM traffic counters.
#include <hip/hip_runtime.h>

#include "flb_internal.hpp"

namespace flb {
namespace {

__global__ __launch_bounds__(kMaxThreads) void k_copy_probe(float4* __restrict__ dst,
                                                            const float4* __restrict__ src,
                                                            size_t n4) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n4) dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy_probe(void* dst, const void* src, size_t n4, hipStream_t s) {
  const size_t nb = (n4 + kMaxThreads - 1) / kMaxThreads;
  hipLaunchKernelGGL(k_copy_probe, dim3(unsigned(nb)), dim3(kMaxThreads), 0, s,
                     static_cast<float4*>(dst), static_cast<const float4*>(src), n4);
  return hipGetLastError();
}

}  // namespace flb
