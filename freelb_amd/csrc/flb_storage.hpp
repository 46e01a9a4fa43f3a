// Population storage formats in HBM.
//
// Each direction i owns one array (SoA) in the padded slot layout described
// in flb_internal.hpp. The shifted 16-bit formats store f_i - w_i (rho0 = 1):
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

struct StoreF16S {
  using E = uint16_t;
  using R = float;
  static constexpr int kind = 1;
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    return __half2float(__ushort_as_half(e)) + Lat<Q>::template w<float>(i);
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    return __half_as_ushort(__float2half_rn(v - Lat<Q>::template w<float>(i)));
  }
};

// 16-bit float of f - w with 1 sign, 4 exponent and 11 fraction bits:
// |x| = (1 + m/2^11) 2^(E-18) for E = 1..15 and m 2^-28 for E = 0, so the
// range is [2^-28, 0.2499] with 12 significant bits, one more than IEEE half
// (f - w stays far below 0.25 at lattice speeds). Round to nearest even;
// larger magnitudes saturate. The E = 0 codes are formed by an FP32 add of
// 2^-5, whose ulp is exactly 2^-28.
struct StoreF16C {
  using E = uint16_t;
  using R = float;
  static constexpr int kind = 4;
  static constexpr uint32_t kTwoM5 = 122u << 23;  // bits of 2^-5
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    const uint32_t m = uint32_t(e) & 0x7FFFu;
    const float sub = __uint_as_float(kTwoM5 + m) - 0.03125f;
    const float nor = __uint_as_float((m + (109u << 11)) << 12);
    const float a = m < 2048u ? sub : nor;
    return ((e & 0x8000u) ? -a : a) + Lat<Q>::template w<float>(i);
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    const uint32_t u = __float_as_uint(v - Lat<Q>::template w<float>(i));
    const uint32_t a = u & 0x7FFFFFFFu;
    const int cn = int((a + 0x7FFu + ((a >> 12) & 1u)) >> 12) - int(109u << 11);
    const uint32_t nor = uint32_t(cn < 0x7FFF ? cn : 0x7FFF);
    const uint32_t sub = __float_as_uint(__uint_as_float(a) + 0.03125f) - kTwoM5;
    return E(((u >> 16) & 0x8000u) | (a < (110u << 23) ? sub : nor));
  }
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
