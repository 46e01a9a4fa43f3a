// Device bandwidth probes: copy, read-only and write-only kernels over float4
// elements, in the shapes a streaming kernel can take on this hardware:
//   - U independent 16-byte accesses in flight per lane (U = 1, 2, 4, 8);
//   - plain or non-temporal loads and stores;
//   - a one-shot grid (every element visited by its own lane once) or a
//     persistent grid of k workgroups per CU that strides over the buffer.
// The best of these is the achievable HBM rate against which the lattice
// steps are set, next to the vendor peak. The one-shot U = 1 copy also
// calibrates the HBM traffic counters (FETCH_SIZE / WRITE_SIZE).
#include <hip/hip_runtime.h>

#include "flb_internal.hpp"

namespace flb {
namespace {

using f4 = float __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ f4 ld4(const f4* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st4(f4* p, f4 v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

__global__ __launch_bounds__(kMaxThreads) void k_copy_probe(float4* __restrict__ dst,
                                                            const float4* __restrict__ src,
                                                            size_t n4) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n4) dst[i] = src[i];
}

// MODE 0 copy, 1 read, 2 write. Workgroup b covers the U * 256 elements
// starting at (b + k * gridDim.x) * U * 256, lane t the elements t + j * 256:
// each of the U accesses of a wavefront is one contiguous 1 KiB run. n4 is a
// multiple of U * 256 (the host rounds it down).
template <int MODE, int U, bool NT>
__global__ __launch_bounds__(kMaxThreads) void k_bw_probe(f4* __restrict__ dst, const f4* __restrict__ src,
                                                          size_t n4) {
  constexpr size_t tile = size_t(U) * kMaxThreads;
  const size_t ntile = n4 / tile;
  uint32_t acc = 0u;
  for (size_t b = blockIdx.x; b < ntile; b += gridDim.x) {
    const size_t base = b * tile + threadIdx.x;
    if constexpr (MODE == 2) {
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const float v = float(base + size_t(j) * kMaxThreads);
        st4<NT>(dst + base + size_t(j) * kMaxThreads, f4{v, v, v, v});
      }
    } else {
      f4 v[U];
#pragma unroll
      for (int j = 0; j < U; ++j) v[j] = ld4<NT>(src + base + size_t(j) * kMaxThreads);
      if constexpr (MODE == 0) {
#pragma unroll
        for (int j = 0; j < U; ++j) st4<NT>(dst + base + size_t(j) * kMaxThreads, v[j]);
      } else {
#pragma unroll
        for (int j = 0; j < U; ++j)
          acc ^= __float_as_uint(v[j].x) ^ __float_as_uint(v[j].y) ^ __float_as_uint(v[j].z) ^
                 __float_as_uint(v[j].w);
      }
    }
  }
  // read-only form: keep the loads live without storing what they read (the
  // source holds zeros, so the condition never holds)
  if (MODE == 1 && acc == 0x9E3779B9u) dst[threadIdx.x] = f4{1.f, 1.f, 1.f, 1.f};
}

template <int MODE, bool NT>
hipError_t launch_bw_u(int unroll, unsigned grid, f4* dst, const f4* src, size_t n4, hipStream_t s) {
  switch (unroll) {
    case 1: hipLaunchKernelGGL((k_bw_probe<MODE, 1, NT>), dim3(grid), dim3(kMaxThreads), 0, s, dst, src, n4); break;
    case 2: hipLaunchKernelGGL((k_bw_probe<MODE, 2, NT>), dim3(grid), dim3(kMaxThreads), 0, s, dst, src, n4); break;
    case 4: hipLaunchKernelGGL((k_bw_probe<MODE, 4, NT>), dim3(grid), dim3(kMaxThreads), 0, s, dst, src, n4); break;
    case 8: hipLaunchKernelGGL((k_bw_probe<MODE, 8, NT>), dim3(grid), dim3(kMaxThreads), 0, s, dst, src, n4); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_copy_probe(void* dst, const void* src, size_t n4, hipStream_t s) {
  const size_t nb = (n4 + kMaxThreads - 1) / kMaxThreads;
  hipLaunchKernelGGL(k_copy_probe, dim3(unsigned(nb)), dim3(kMaxThreads), 0, s,
                     static_cast<float4*>(dst), static_cast<const float4*>(src), n4);
  return hipGetLastError();
}

hipError_t launch_bw_probe(int mode, int unroll, bool nt, unsigned grid, void* dst, const void* src, size_t n4,
                           hipStream_t s) {
  f4* d = static_cast<f4*>(dst);
  const f4* r = static_cast<const f4*>(src);
  switch (mode * 2 + int(nt)) {
    case 0: return launch_bw_u<0, false>(unroll, grid, d, r, n4, s);
    case 1: return launch_bw_u<0, true>(unroll, grid, d, r, n4, s);
    case 2: return launch_bw_u<1, false>(unroll, grid, d, r, n4, s);
    case 3: return launch_bw_u<1, true>(unroll, grid, d, r, n4, s);
    case 4: return launch_bw_u<2, false>(unroll, grid, d, r, n4, s);
    case 5: return launch_bw_u<2, true>(unroll, grid, d, r, n4, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace flb
