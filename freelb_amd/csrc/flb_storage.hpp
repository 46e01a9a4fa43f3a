// Population storage formats in HBM.
//
// Each direction i owns one array (SoA) in the padded slot layout described
// in flb_internal.hpp. The shifted half formats store f_i - w_i (rho0 = 1):
// around equilibrium the stored value is a small deviation, so the 11-bit
// (FP16) or 8-bit (BF16) significand resolves it relative to its own size
// instead of relative to w_i. All arithmetic is done in the compute type R
// (float, or double for F64). w_i = w_{opp i}, so the shift of a value does
// not depend on which of the two AA slots it sits in.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "flb_cell.hpp"

namespace flb {

template <class T, int N>
using vec_t = T __attribute__((ext_vector_type(N)));

struct StoreF32 {
  using E = float;  // element in HBM
  using R = float;  // compute type
  static constexpr int kind = 0;
  template <int Q>
  __device__ __forceinline__ static R from(E e, int) {
    return e;
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int) {
    return v;
  }
};

struct StoreF64 {
  using E = double;
  using R = double;
  static constexpr int kind = 3;
  template <int Q>
  __device__ __forceinline__ static R from(E e, int) {
    return e;
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int) {
    return v;
  }
};

#ifndef FLB_F16_VARIANT
#define FLB_F16_VARIANT 0
#endif

struct StoreF16S {
  using E = uint16_t;
  using R = float;
  static constexpr int kind = 1;
#if FLB_F16_VARIANT == 0
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    return __half2float(__ushort_as_half(e)) + Lat<Q>::template w<float>(i);
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    return __half_as_ushort(__float2half_rn(v - Lat<Q>::template w<float>(i)));
  }
#elif FLB_F16_VARIANT == 1 || FLB_F16_VARIANT == 3
  // IEEE half of (f - w) * 2^14: the shifted values stay in the normal range
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    return __half2float(__ushort_as_half(e)) * 6.103515625e-05f + Lat<Q>::template w<float>(i);
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    return __half_as_ushort(__float2half_rn((v - Lat<Q>::template w<float>(i)) * 16384.0f));
  }
#elif FLB_F16_VARIANT == 2
  // 1 sign, 4 exponent, 11 fraction bits: |x| = (1 + m/2^11) 2^(E-18), E = 1..15,
  // m 2^-28 for E = 0; largest 0.2499, round to nearest even, saturating
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    const uint32_t m = uint32_t(e) & 0x7FFFu;
    float a = (m >> 11) == 0u ? float(m) * 3.7252902984619140625e-09f
                              : __uint_as_float((m + (109u << 11)) << 12);
    return ((e & 0x8000u) ? -a : a) + Lat<Q>::template w<float>(i);
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    const float x = v - Lat<Q>::template w<float>(i);
    const uint32_t u = __float_as_uint(x);
    const uint32_t sgn = (u >> 16) & 0x8000u;
    const uint32_t a = u & 0x7FFFFFFFu;
    uint32_t code;
    if (a < (110u << 23)) {  // below 2^-17
      code = uint32_t(__float2uint_rn(__uint_as_float(a) * 268435456.0f));
    } else {
      const uint32_t r = (a + 0x7FFu + ((a >> 12) & 1u)) >> 12;
      code = r - (109u << 11);
      if (code > 0x7FFFu) code = 0x7FFFu;
    }
    return E(sgn | code);
  }
#endif
};

struct StoreBF16S {
  using E = uint16_t;
  using R = float;
  static constexpr int kind = 2;
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    return __uint_as_float(uint32_t(e) << 16) + Lat<Q>::template w<float>(i);
  }
  // Round to nearest even on the upper 16 bits (values are finite here).
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    const uint32_t u = __float_as_uint(v - Lat<Q>::template w<float>(i));
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return E(r >> 16);
  }
};

// VEC consecutive elements from an address aligned to VEC*sizeof(E).
template <class E, int VEC>
__device__ __forceinline__ vec_t<E, VEC> load_vec(const E* p) {
  return *reinterpret_cast<const vec_t<E, VEC>*>(p);
}
template <class E, int VEC>
__device__ __forceinline__ void store_vec(E* p, vec_t<E, VEC> v) {
  *reinterpret_cast<vec_t<E, VEC>*>(p) = v;
}

}  // namespace flb
