// Population storage formats in HBM.
//
// Each direction i owns one array (SoA) in the padded slot layout described
// in flb_internal.hpp. The shifted 16-bit formats store f_i - w_i (rho0 = 1):
// around equilibrium the stored value is a small deviation, so the 11-bit
// (FP16) or 8-bit (BF16) significand resolves it relative to its own size
// instead of relative to w_i. All arithmetic is done in the compute type R
// (float, or double for F64). w_i = w_{opp i}, so the shift of a value does
// not depend on which of the two AA slots it sits in.
//
// from / to convert between a stored element and f_i. The shifted kinds
// (kShifted) also give dev / undev, the conversion between the element and
// the deviation g_i = f_i - w_i itself, for kernels that compute on g.
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
  static constexpr bool kShifted = false;
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
  static constexpr bool kShifted = false;
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
  static constexpr bool kShifted = true;
  __device__ __forceinline__ static R dev(E e) { return __half2float(__ushort_as_half(e)); }
  __device__ __forceinline__ static E undev(R d) { return __half_as_ushort(__float2half_rn(d)); }
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    return dev(e) + Lat<Q>::template w<float>(i);
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    return undev(v - Lat<Q>::template w<float>(i));
  }
};

// 16-bit float of f - w with 1 sign, 4 exponent and 11 fraction bits:
// |x| = (1 + m/2^11) 2^(E-18) for E = 1..15 and m 2^-28 for E = 0, so the
// range is [2^-28, 0.2499] with 12 significant bits, one more than IEEE half
// (f - w stays far below 0.25 at lattice speeds). Round to nearest even;
// larger magnitudes saturate. The E = 0 codes are formed by an FP32 add of
// 2^-5, whose ulp is exactly 2^-28.
//
// The code is the FP32 word of x 2^-109 with its 12 leading significant bits
// kept: an FP32 number with exponent field E and fraction m 2^12, i.e.
// (1 + m/2^11) 2^(E-127) for E >= 1 and the subnormal m 2^-137 for E = 0, is
// exactly the F16C value times 2^-109 (the kernels keep FP32 subnormals).
// - Decoding places the code's magnitude bits at bit 12 and its sign at bit
//   31 of an FP32 word; one fused multiply-add by 2^109 (exact) adds w.
// - Encoding scales f - w by 2^-109, which is exact here: below 2^-17 the
//   difference of f and w is itself exact (Sterbenz) and has no bit under
//   2^-31, above it the product stays normal. The integer add of 0x7FF plus
//   the lowest kept bit rounds the FP32 word to nearest even at bit 12, in
//   the normal and the subnormal range alike.
// - undev encodes a deviation computed as such (not as a difference): below
//   2^-17 its scaled value can round once at 2^-149 before the rounding at
//   bit 12, so a value within 2^-40 of a tie may round to the other
//   neighbour (an error of at most half a code step plus 2^-40).
struct StoreF16C {
  using E = uint16_t;
  using R = float;
  static constexpr int kind = 4;
  static constexpr bool kShifted = true;
  // the code as an FP32 number: the value times 2^-109
  __device__ __forceinline__ static float scaled(E e) {
    const uint32_t c = uint32_t(e);
    return __uint_as_float(((c & 0x8000u) << 16) | ((c & 0x7FFFu) << 12));
  }
  __device__ __forceinline__ static R dev(E e) { return scaled(e) * 0x1p109f; }
  __device__ __forceinline__ static E undev(R d) {
    const uint32_t u = __float_as_uint(d * 0x1p-109f);
    const uint32_t a = u & 0x7FFFFFFFu;
    const uint32_t r = (a + 0x7FFu + ((a >> 12) & 1u)) >> 12;
    return E(((u >> 16) & 0x8000u) | (r < 0x7FFFu ? r : 0x7FFFu));
  }
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    return __fmaf_rn(scaled(e), 0x1p109f, Lat<Q>::template w<float>(i));
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    return undev(v - Lat<Q>::template w<float>(i));
  }
};

struct StoreBF16S {
  using E = uint16_t;
  using R = float;
  static constexpr int kind = 2;
  static constexpr bool kShifted = true;
  __device__ __forceinline__ static R dev(E e) { return __uint_as_float(uint32_t(e) << 16); }
  // Round to nearest even on the upper 16 bits (values are finite here).
  __device__ __forceinline__ static E undev(R d) {
    const uint32_t u = __float_as_uint(d);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return E(r >> 16);
  }
  template <int Q>
  __device__ __forceinline__ static R from(E e, int i) {
    return dev(e) + Lat<Q>::template w<float>(i);
  }
  template <int Q>
  __device__ __forceinline__ static E to(R v, int i) {
    return undev(v - Lat<Q>::template w<float>(i));
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
