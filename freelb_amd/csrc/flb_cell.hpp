// Per-cell lattice-Boltzmann arithmetic for the gfx950 kernels.
//
// Every function here works on one cell's q populations held in registers.
// Expression order follows FreeLB's unrolled specialisations (the *.ur.h
// files, which the reference builds with -D_UNROLLFOR) term by term, and the
// kernels are compiled with floating-point contraction off, so an FP32 run
// rounds exactly as FreeLB's FP32 run does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace flb {

// Logical workgroup index: the hardware hands consecutive workgroups to the
// 8 XCDs in turn, so workgroup b runs on XCD b % 8 as its (b / 8)-th one.
// Mapping it to xcd * (nb / 8) + b / 8 gives every XCD (and its L2) a
// contiguous range of the grid's work.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned per = nb >> 3, rem = nb & 7u, xcd = b & 7u, slot = b >> 3;
  return xcd * per + (xcd < rem ? xcd : rem) + slot;
}

// ---------------------------------------------------------------------------
// Velocity sets in FreeLB's order (lbm/lattice_set.h:168-194 D2Q9,
// 259-296 D3Q19). Opposite pairs are (2k-1, 2k).
// ---------------------------------------------------------------------------
template <int Q>
struct Lat;

template <>
struct Lat<19> {
  static constexpr int D = 3;
  __host__ __device__ static constexpr int cx(int i) {
    constexpr int v[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
    return v[i];
  }
  __host__ __device__ static constexpr int cy(int i) {
    constexpr int v[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
    return v[i];
  }
  __host__ __device__ static constexpr int cz(int i) {
    constexpr int v[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
    return v[i];
  }
  // weight class: 0 -> 1/3, 1 -> 1/18, 2 -> 1/36
  __host__ __device__ static constexpr int wclass(int i) { return i == 0 ? 0 : (i <= 6 ? 1 : 2); }
  template <class R>
  __host__ __device__ static constexpr R w(int i) {
    return wclass(i) == 0 ? R(1) / R(3) : (wclass(i) == 1 ? R(1) / R(18) : R(1) / R(36));
  }
};

template <>
struct Lat<9> {
  static constexpr int D = 2;
  __host__ __device__ static constexpr int cx(int i) {
    constexpr int v[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
    return v[i];
  }
  __host__ __device__ static constexpr int cy(int i) {
    constexpr int v[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
    return v[i];
  }
  __host__ __device__ static constexpr int cz(int) { return 0; }
  __host__ __device__ static constexpr int wclass(int i) { return i == 0 ? 0 : (i <= 4 ? 1 : 2); }
  template <class R>
  __host__ __device__ static constexpr R w(int i) {
    return wclass(i) == 0 ? R(4) / R(9) : (wclass(i) == 1 ? R(1) / R(9) : R(1) / R(36));
  }
};

// Advection-diffusion sets D2Q5 (lattice_set.h:144-155) and D3Q7 (196-207):
// the rest direction and the axis pairs, weights 1/3, 1/6 and 1/4, 1/8. Both
// keep cs2 = 1/3 in the equilibrium, as FreeLB's do.
template <>
struct Lat<5> {
  static constexpr int D = 2;
  __host__ __device__ static constexpr int cx(int i) { return i == 1 ? 1 : (i == 2 ? -1 : 0); }
  __host__ __device__ static constexpr int cy(int i) { return i == 3 ? 1 : (i == 4 ? -1 : 0); }
  __host__ __device__ static constexpr int cz(int) { return 0; }
  __host__ __device__ static constexpr int wclass(int i) { return i == 0 ? 0 : 1; }
  template <class R>
  __host__ __device__ static constexpr R w(int i) {
    return i == 0 ? R(1) / R(3) : R(1) / R(6);
  }
};

template <>
struct Lat<7> {
  static constexpr int D = 3;
  __host__ __device__ static constexpr int cx(int i) { return i == 1 ? 1 : (i == 2 ? -1 : 0); }
  __host__ __device__ static constexpr int cy(int i) { return i == 3 ? 1 : (i == 4 ? -1 : 0); }
  __host__ __device__ static constexpr int cz(int i) { return i == 5 ? 1 : (i == 6 ? -1 : 0); }
  __host__ __device__ static constexpr int wclass(int i) { return i == 0 ? 0 : 1; }
  template <class R>
  __host__ __device__ static constexpr R w(int i) {
    return i == 0 ? R(1) / R(4) : R(1) / R(8);
  }
};

// Advection-diffusion (scalar) velocity sets.
template <int Q>
constexpr bool kScalarLat = Q == 5 || Q == 7;

__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

// moment::rhoImpl<D2Q5 / D3Q7> (moment.ur.h:47-82): the scalar (temperature,
// concentration) is the left-to-right sum of the populations.
template <class R, int Q>
__device__ __forceinline__ R rho_sum(const R (&c)[Q]) {
  R r = c[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) r = r + c[i];
  return r;
}

// Read-out of a scalar lattice for flb_moments: the sum and sum c f / sum.
template <class R>
__device__ __forceinline__ void rhou(const R (&c)[5], R& rho, R (&u)[3]) {
  rho = rho_sum(c);
  u[0] = (c[1] - c[2]) / rho;
  u[1] = (c[3] - c[4]) / rho;
  u[2] = R(0);
}
template <class R>
__device__ __forceinline__ void rhou(const R (&c)[7], R& rho, R (&u)[3]) {
  rho = rho_sum(c);
  u[0] = (c[1] - c[2]) / rho;
  u[1] = (c[3] - c[4]) / rho;
  u[2] = (c[5] - c[6]) / rho;
}

// equilibrium::SecondOrderImpl<D2Q5> (equilibrium.ur.h:43-63) and <D3Q7>
// (equilibrium.ur.h:96-119): one weight class per axis pair.
template <class R, int Q>
__device__ __forceinline__ void feq_axes(R (&e)[Q], R rho, const R (&u)[3]) {
  const R InvCs2 = R(3);
  const R InvCs2x = R(0.5) * InvCs2;
  constexpr int D = Lat<Q>::D;
  const R var0 = D == 3 ? R(1) - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * InvCs2x
                        : R(1) - (u[0] * u[0] + u[1] * u[1]) * InvCs2x;
  const R rhowk0 = Lat<Q>::template w<R>(0) * rho;
  const R rhowk1 = Lat<Q>::template w<R>(1) * rho;
  e[0] = rhowk0 * var0;
#pragma unroll
  for (int p = 0; p < D; ++p) {
    const R a = InvCs2 * u[p];
    const R v = var0 + a * a * R(0.5);
    e[2 * p + 1] = rhowk1 * (v + a);
    e[2 * p + 2] = rhowk1 * (v - a);
  }
}
template <class R>
__device__ __forceinline__ void feq(R (&e)[5], R rho, const R (&u)[3]) {
  feq_axes<R, 5>(e, rho, u);
}
template <class R>
__device__ __forceinline__ void feq(R (&e)[7], R rho, const R (&u)[3]) {
  feq_axes<R, 7>(e, rho, u);
}

// ---------------------------------------------------------------------------
// Moments: moment::rhoUImpl (moment.ur.h:461-472 D2Q9, 505-517 D3Q19).
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void rhou(const R (&c)[19], R& rho, R (&u)[3]) {
  const R r = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8] + c[9] + c[10] +
              c[11] + c[12] + c[13] + c[14] + c[15] + c[16] + c[17] + c[18];
  rho = r;
  u[0] = (c[1] - c[2] + c[7] - c[8] + c[9] - c[10] + c[13] - c[14] + c[15] - c[16]) / r;
  u[1] = (c[3] - c[4] + c[7] - c[8] + c[11] - c[12] - c[13] + c[14] + c[17] - c[18]) / r;
  u[2] = (c[5] - c[6] + c[9] - c[10] + c[11] - c[12] - c[15] + c[16] - c[17] + c[18]) / r;
}

template <class R>
__device__ __forceinline__ void rhou(const R (&c)[9], R& rho, R (&u)[3]) {
  const R r = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8];
  rho = r;
  u[0] = (c[1] - c[2] + c[5] - c[6] + c[7] - c[8]) / r;
  u[1] = (c[3] - c[4] + c[5] - c[6] - c[7] + c[8]) / r;
  u[2] = R(0);
}

// ---------------------------------------------------------------------------
// equilibrium::SecondOrderImpl (equilibrium.ur.h:65-94 D2Q9, 165-214 D3Q19).
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void feq(R (&e)[19], R rho, const R (&u)[3]) {
  const R InvCs2 = R(3);
  const R InvCs2x = R(0.5) * InvCs2;
  const R var0 = R(1) - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * InvCs2x;
  const R rhowk0 = Lat<19>::w<R>(0) * rho;
  const R rhowk1 = Lat<19>::w<R>(1) * rho;
  const R rhowk7 = Lat<19>::w<R>(7) * rho;
  const R cu[9] = {u[0],        u[1],        u[2],        u[0] + u[1], u[0] + u[2],
                   u[1] + u[2], u[0] - u[1], u[0] - u[2], u[1] - u[2]};
  e[0] = rhowk0 * var0;
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const R a = InvCs2 * cu[p];
    const R v = var0 + a * a * R(0.5);
    const R rw = p < 3 ? rhowk1 : rhowk7;
    e[2 * p + 1] = rw * (v + a);
    e[2 * p + 2] = rw * (v - a);
  }
}

template <class R>
__device__ __forceinline__ void feq(R (&e)[9], R rho, const R (&u)[3]) {
  const R InvCs2 = R(3);
  const R InvCs2x = R(0.5) * InvCs2;
  const R var0 = R(1) - (u[0] * u[0] + u[1] * u[1]) * InvCs2x;
  const R rhowk0 = Lat<9>::w<R>(0) * rho;
  const R rhowk1 = Lat<9>::w<R>(1) * rho;
  const R rhowk5 = Lat<9>::w<R>(5) * rho;
  const R cu[4] = {u[0], u[1], u[0] + u[1], u[0] - u[1]};
  e[0] = rhowk0 * var0;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const R a = InvCs2 * cu[p];
    const R v = var0 + a * a * R(0.5);
    const R rw = p < 2 ? rhowk1 : rhowk5;
    e[2 * p + 1] = rw * (v + a);
    e[2 * p + 2] = rw * (v - a);
  }
}

// c_i . u accumulated from zero over the axes (Vector.h:229-236).
template <int Q, class R>
__device__ __forceinline__ R cdotu(int i, const R (&u)[3]) {
  R r = R(0);
  r += u[0] * R(Lat<Q>::cx(i));
  r += u[1] * R(Lat<Q>::cy(i));
  if (Lat<Q>::D == 3) r += u[2] * R(Lat<Q>::cz(i));
  return r;
}

// Halfway moving-wall term 2*InvCs2*w_k*RHO*(VELOCITY.c_k)
// (bounce_back_boundary.h:246-250).
template <int Q, class R>
__device__ __forceinline__ R movingwall_term(int k, R rho_f, const R (&u_f)[3]) {
  return R(6) * Lat<Q>::template w<R>(k) * rho_f * cdotu<Q>(k, u_f);
}

// Halfway anti-bounce-back of a scalar lattice, bounceback::anti_simple
// (bounce_back_boundary.h:183-193): f_k = 2 RHO w_k - f*_{opp k}, RHO the
// cell's scalar field (the wall value on Dirichlet walls).
template <int Q, class R>
__device__ __forceinline__ R anti_term(int k, R rho_f, R prev) {
  return R(2) * rho_f * Lat<Q>::template w<R>(k) - prev;
}

// ---------------------------------------------------------------------------
// Collisions.
// ---------------------------------------------------------------------------

// collision::BGK (collision.h:39-56): f_i <- omega*feq_i + (1-omega)*f_i.
template <int Q, class R>
__device__ __forceinline__ void bgk_with(R (&f)[Q], R rho, const R (&u)[3], R omega, R _omega) {
  R e[Q];
  feq(e, rho, u);
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = omega * e[i] + _omega * f[i];
}

// BGK on the deviations g_i = f_i - w_i that the shifted storage kinds hold
// (D2Q9 / D3Q19). Same algebra as collision::BGK on f: with sum w_i = 1 and
// sum c_i w_i = 0,
//   rho - 1 = sum g_i,   rho u = sum c_i g_i,
//   feq_i - w_i = w_i (rho (1 + a_i + a_i^2 / 2 - 3 u.u / 2) - 1),  a_i = 3 c_i.u,
// and g_i <- omega (feq_i - w_i) + (1 - omega) g_i. With d = rho - 1,
// A = d - rho 3 u.u / 2 and t_i = rho a_i / 2, the bracket of direction i is
// A + t_i a_i + 2 t_i and that of opp(i) (a = -a_i) A + t_i a_i - 2 t_i: no
// sum near 1 is formed, so the density keeps the resolution of its deviation.
// The rounding is not FreeLB's (the FP32 and FP64 kinds keep that order).

// c_i . v with the zero and unit components of c_i folded away.
template <int Q, class R>
__device__ __forceinline__ R cdot_folded(int i, const R (&v)[3]) {
  R r = R(0);
  bool any = false;
#pragma unroll
  for (int d = 0; d < Lat<Q>::D; ++d) {
    const int c = d == 0 ? Lat<Q>::cx(i) : (d == 1 ? Lat<Q>::cy(i) : Lat<Q>::cz(i));
    if (c > 0) {
      r = any ? r + v[d] : v[d];
      any = true;
    } else if (c < 0) {
      r = any ? r - v[d] : -v[d];
      any = true;
    }
  }
  return r;
}

template <int Q, class R>
__device__ __forceinline__ void rhou_dev(const R (&g)[Q], R& d, R& rho, R (&u)[3]) {
  R s = g[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) s += g[i];
  d = s;
  rho = R(1) + s;
  const R inv = R(1) / rho;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k >= Lat<Q>::D) {
      u[k] = R(0);
      continue;
    }
    R j = R(0);
    bool any = false;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const int c = k == 0 ? Lat<Q>::cx(i) : (k == 1 ? Lat<Q>::cy(i) : Lat<Q>::cz(i));
      if (c > 0) {
        j = any ? j + g[i] : g[i];
        any = true;
      } else if (c < 0) {
        j = any ? j - g[i] : -g[i];
        any = true;
      }
    }
    u[k] = j * inv;
  }
}

template <int Q, class R>
__device__ __forceinline__ void bgk_dev(R (&g)[Q], R d, R rho, const R (&u)[3], R omega, R _omega) {
  static_assert(Q == 9 || Q == 19, "flow lattices");
  R U[3];
  R uu = u[0] * u[0] + u[1] * u[1];
  if (Lat<Q>::D == 3) uu += u[2] * u[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) U[k] = R(3) * u[k];
  const R A = d - rho * (R(1.5) * uu);
  const R rh = R(0.5) * rho;
  g[0] = omega * Lat<Q>::template w<R>(0) * A + _omega * g[0];
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    const R ai = cdot_folded<Q>(i, U);
    const R t = rh * ai;
    const R q = A + t * ai;
    const R ow = omega * Lat<Q>::template w<R>(i);
    g[i] = ow * (q + R(2) * t) + _omega * g[i];
    g[i + 1] = ow * (q - R(2) * t) + _omega * g[i + 1];
  }
}

// moment::forcerhoU (forcerhoUImpl<D3Q19> moment.ur.h:639-663, <D2Q9>
// 564-574): u = (sum c f + F/2) / rho, each component divided on its own.
template <class R>
__device__ __forceinline__ void force_rhou(const R (&c)[19], const R (&F)[3], R& rho, R (&u)[3]) {
  rho = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8] + c[9] + c[10] + c[11] + c[12] +
        c[13] + c[14] + c[15] + c[16] + c[17] + c[18];
  u[0] = (c[1] - c[2] + c[7] - c[8] + c[9] - c[10] + c[13] - c[14] + c[15] - c[16] + F[0] * R(0.5)) / rho;
  u[1] = (c[3] - c[4] + c[7] - c[8] + c[11] - c[12] - c[13] + c[14] + c[17] - c[18] + F[1] * R(0.5)) / rho;
  u[2] = (c[5] - c[6] + c[9] - c[10] + c[11] - c[12] - c[15] + c[16] - c[17] + c[18] + F[2] * R(0.5)) / rho;
}
template <class R>
__device__ __forceinline__ void force_rhou(const R (&c)[9], const R (&F)[3], R& rho, R (&u)[3]) {
  rho = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8];
  u[0] = (c[1] - c[2] + c[5] - c[6] + c[7] - c[8] + F[0] * R(0.5)) / rho;
  u[1] = (c[3] - c[4] + c[5] - c[6] - c[7] + c[8] + F[1] * R(0.5)) / rho;
  u[2] = R(0);
}

// c_i . u of the first direction of opposite pair p (directions 2p+1, 2p+2),
// summed the way the unrolled specialisations write it (first nonzero
// component first): u0, u1, u2, u0+u1, u0+u2, u1+u2, u0-u1, u0-u2, u1-u2
// (D3Q19), u0, u1, u0+u1, u0-u1 (D2Q9).
template <int Q, class R>
__device__ __forceinline__ R pair_cu(int p, const R (&u)[3]) {
  return cdot_folded<Q>(2 * p + 1, u);
}

// force::ForcePopImpl (the D3Q19 / D2Q9 forms of force.ur.h; generic form
// force.h:29-36): Fi = w_i InvCs2 sum_a F_a (c_ia - u_a + InvCs2 (c_i.u) c_ia).
// The unrolled form leaves out the zero components of c_i: a term is
// -u_a where c_ia = 0, else (+-1 - u_a) +- InvCs2 (c.u) of the pair, with
// the pair's c.u formed for its first direction (first component +). The
// terms are generated from c_i in that association; the rest direction is
// w_0 InvCs2 (-(F.u)).
template <int Q, class R>
__device__ __forceinline__ void force_pop(const R (&u)[3], const R (&F)[3], R (&Fi)[Q]) {
  constexpr int D = Lat<Q>::D;
  const R I2 = R(3);
  R T1[3], _T1[3], nu[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    T1[a] = R(1) - u[a];
    _T1[a] = R(-1) - u[a];
    nu[a] = -u[a];
  }
  R fu = F[0] * u[0] + F[1] * u[1];
  if constexpr (D == 3) fu = fu + F[2] * u[2];
  Fi[0] = Lat<Q>::template w<R>(0) * I2 * (-fu);
#pragma unroll
  for (int p = 0; p < (Q - 1) / 2; ++p) {
    const R cp = I2 * pair_cu<Q>(p, u);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = 2 * p + 1 + h;  // h = 1: the opposite direction, whose c.u is -cp
      R acc = R(0);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const int c = a == 0 ? Lat<Q>::cx(k) : (a == 1 ? Lat<Q>::cy(k) : Lat<Q>::cz(k));
        const R t = c == 0 ? nu[a] : ((c > 0) == (h == 0) ? (c > 0 ? T1[a] : _T1[a]) + cp
                                                             : (c > 0 ? T1[a] : _T1[a]) - cp);
        acc = a == 0 ? F[0] * t : acc + F[a] * t;
      }
      Fi[k] = Lat<Q>::template w<R>(k) * I2 * acc;
    }
  }
}

// collision::BGKForce (collision.h:68-87): omega feq + (1 - omega) f +
// (1 - omega/2) Fi, fOmega formed as FreeLB's BlockLattice does
// (block_lattice.hh:33-34).
template <int Q, class R>
__device__ __forceinline__ void bgk_force(R (&f)[Q], const R (&F)[3], R omega, R _omega) {
  R rho, u[3];
  force_rhou(f, F, rho, u);
  R Fi[Q];
  force_pop(u, F, Fi);
  R e[Q];
  feq(e, rho, u);
  const R fomega = R(1) - R(0.5) * omega;
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = omega * e[i] + _omega * f[i] + fomega * Fi[i];
}

// ---------------------------------------------------------------------------
// Scalar force along the last axis (y in 2-D, z in 3-D): force::ScalarForce
// (force.h:109-133) as the buoyancy of a coupled temperature / concentration
// lattice sets it.
// ---------------------------------------------------------------------------


// forcerhoUImpl's scalar overload (moment.h:423-435, unrolled at
// moment.ur.h:564-574, 652-663): rho and the sums c f, then F/2 added to the
// force axis, then each component divided by rho.
template <int Q, class R>
__device__ __forceinline__ void scalar_force_rhou(const R (&c)[Q], R F, R& rho, R (&u)[3]) {
  constexpr int D = Lat<Q>::D;
  if constexpr (Q == 19) {
    rho = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8] + c[9] + c[10] + c[11] + c[12] +
          c[13] + c[14] + c[15] + c[16] + c[17] + c[18];
    u[0] = c[1] - c[2] + c[7] - c[8] + c[9] - c[10] + c[13] - c[14] + c[15] - c[16];
    u[1] = c[3] - c[4] + c[7] - c[8] + c[11] - c[12] - c[13] + c[14] + c[17] - c[18];
    u[2] = c[5] - c[6] + c[9] - c[10] + c[11] - c[12] - c[15] + c[16] - c[17] + c[18];
  } else {
    rho = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8];
    u[0] = c[1] - c[2] + c[5] - c[6] + c[7] - c[8];
    u[1] = c[3] - c[4] + c[5] - c[6] - c[7] + c[8];
    u[2] = R(0);
  }
  u[D - 1] += F * R(0.5);
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] /= rho;
}

// ScalarForcePopImpl<D2Q9, 1> / <D3Q19, 2> (force.ur.h:332-353, 584-616):
// Fi = w_i InvCs2 F (c_id - u_d + InvCs2 (c_i.u) c_id), written per direction
// as (+-1 - u_d) +- InvCs2 (c.u) of the pair, or -u_d where c_id = 0.
template <int Q, class R>
__device__ __forceinline__ void scalar_force_pop(const R (&u)[3], R F, R (&Fi)[Q]) {
  constexpr int D = Lat<Q>::D;
  const R I2 = R(3);
  const R ud = u[D - 1];
  const R T1_ud = R(1) - ud, _T1_ud = R(-1) - ud, _ud = -ud;
  Fi[0] = Lat<Q>::template w<R>(0) * I2 * F * (_ud);
#pragma unroll
  for (int p = 0; p < (Q - 1) / 2; ++p) {
    const int i = 2 * p + 1;
    const int s = D == 3 ? Lat<Q>::cz(i) : Lat<Q>::cy(i);  // c_id of the pair's first direction
    if (s == 0) {
      Fi[i] = Lat<Q>::template w<R>(i) * I2 * F * (_ud);
      Fi[i + 1] = Lat<Q>::template w<R>(i + 1) * I2 * F * (_ud);
    } else {
      const R cu = I2 * pair_cu<Q>(p, u);
      Fi[i] = Lat<Q>::template w<R>(i) * I2 * F * (s > 0 ? T1_ud + cu : _T1_ud - cu);
      Fi[i + 1] = Lat<Q>::template w<R>(i + 1) * I2 * F * (s > 0 ? _T1_ud + cu : T1_ud - cu);
    }
  }
}

// collision::BGKForce with forcerhoU<CELL, ScalarForce> and ScalarForce
// (collision.h:68-87): f <- omega feq + (1 - omega) f + (1 - omega/2) Fi.
template <int Q, class R>
__device__ __forceinline__ void bgk_scalar_force(R (&f)[Q], R F, R omega, R _omega, R& rho, R (&u)[3]) {
  scalar_force_rhou<Q>(f, F, rho, u);
  R Fi[Q];
  scalar_force_pop<Q>(u, F, Fi);
  R e[Q];
  feq(e, rho, u);
  const R fomega = R(1) - R(0.5) * omega;
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = omega * e[i] + _omega * f[i] + fomega * Fi[i];
}

// collision::BounceBack (collision.ur.h:44-93).
template <int Q, class R>
__device__ __forceinline__ void bounceback(R (&f)[Q]) {
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    const R t = f[i];
    f[i] = f[i + 1];
    f[i + 1] = t;
  }
}

// collision::BounceBackMovingWall (collision.ur.h:97-187): uc_i = (c_i.u) w_i rhox,
// where c_i.u sums only the non-zero components (u0, u0+u1, u0-u1, ...).
template <int Q, class R>
__device__ __forceinline__ void bounceback_moving(R (&f)[Q], R rho_f, const R (&u)[3]) {
  const R rhox = R(6) * rho_f;
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    const int c[3] = {Lat<Q>::cx(i), Lat<Q>::cy(i), Lat<Q>::cz(i)};
    R cu = R(0);
    bool first = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (c[a] == 0) continue;
      if (first) {
        cu = c[a] > 0 ? u[a] : -u[a];
        first = false;
      } else {
        cu = c[a] > 0 ? cu + u[a] : cu - u[a];
      }
    }
    const R uc = cu * Lat<Q>::template w<R>(i) * rhox;
    const R t = f[i];
    f[i] = f[i + 1] + uc;
    f[i + 1] = t - uc;
  }
}

template <class T>
__device__ __forceinline__ T dsqrt(T v);
template <>
__device__ __forceinline__ float dsqrt<float>(float v) {
  return sqrtf(v);
}
template <>
__device__ __forceinline__ double dsqrt<double>(double v) {
  return sqrt(v);
}

// The Smagorinsky omega for |Pi_neq| = norm (collisionLES.h:36-59); the
// coefficient is formed in double because std::sqrt(2) is a double, and with
// the usual RHO field of 1 its double division is by 1 and is skipped (the
// same bits).
template <class R>
__device__ __forceinline__ R smag_from_norm(R norm, R omega, R cs, R rho_f) {
  const double cd = double(cs * cs * R(3) * R(3) * R(2)) * 1.4142135623730951;
  const R coeff = rho_f == R(1) ? R(cd) : R(cd / double(rho_f));
  const R rt_mol = R(1) / omega;
  const R rt_tur = R(0.5) * (dsqrt(rt_mol * rt_mol + coeff * norm) - rt_mol);
  return R(1) / (rt_mol + rt_tur);
}

// Smagorinsky effective omega (collisionLES.h:36-59): |Pi_neq| from
// Pi_ab_neq (moment.ur.h:708-718, 750-763) and NormSquare (util.h:418-434);
// the coefficient is formed in double because std::sqrt(2) is a double.
// DEV: c holds the deviations f - w and d = rho - 1; sum c_a^2 w = cs2 and
// sum c_a c_b w = 0 (a != b), so the diagonal terms subtract rho u_a^2 + d cs2.
template <bool DEV = false, class R>
__device__ __forceinline__ R smag_omega(const R (&c)[19], R rho, const R (&u)[3], R omega, R cs,
                                        R rho_f, R d = R(0)) {
  const R cs2 = R(1) / R(3);
  auto diag = [&](R s, R ua) { return DEV ? s - (rho * (ua * ua) + d * cs2) : s - rho * (ua * ua + cs2); };
  const R t0 = diag(c[1] + c[2] + c[7] + c[8] + c[9] + c[10] + c[13] + c[14] + c[15] + c[16], u[0]);
  const R t1 = c[7] + c[8] - c[13] - c[14] - rho * (u[0] * u[1]);
  const R t2 = c[9] + c[10] - c[15] - c[16] - rho * (u[0] * u[2]);
  const R t3 = diag(c[3] + c[4] + c[7] + c[8] + c[11] + c[12] + c[13] + c[14] + c[17] + c[18], u[1]);
  const R t4 = c[11] + c[12] - c[17] - c[18] - rho * (u[1] * u[2]);
  const R t5 = diag(c[5] + c[6] + c[9] + c[10] + c[11] + c[12] + c[15] + c[16] + c[17] + c[18], u[2]);
  const R ns = t0 * t0 + t1 * t1 * R(2) + t2 * t2 * R(2) + t3 * t3 + t4 * t4 * R(2) + t5 * t5;
  return smag_from_norm(dsqrt(ns), omega, cs, rho_f);
}

template <class R>
__device__ __forceinline__ R smag_omega(const R (&c)[9], R rho, const R (&u)[3], R omega, R cs,
                                        R rho_f) {
  const R cs2 = R(1) / R(3);
  const R t0 = c[1] + c[2] + c[5] + c[6] + c[7] + c[8] - rho * (u[0] * u[0] + cs2);
  const R t1 = c[5] + c[6] - c[7] - c[8] - rho * (u[0] * u[1]);
  const R t2 = c[3] + c[4] + c[5] + c[6] + c[7] + c[8] - rho * (u[1] * u[1] + cs2);
  const R ns = t0 * t0 + t1 * t1 * R(2) + t2 * t2;
  const R norm = dsqrt(ns);
  const R coeff =
      R(double(cs * cs * R(3) * R(3) * R(2)) * 1.4142135623730951 / double(rho_f));
  const R rt_mol = R(1) / omega;
  const R rt_tur = R(0.5) * (dsqrt(rt_mol * rt_mol + coeff * norm) - rt_mol);
  return R(1) / (rt_mol + rt_tur);
}

// ---------------------------------------------------------------------------
// MRT. D2Q9 as collision::MRT_Feq_RhoU<D2Q9> (collisionMRT.ur.h:45-129) with
// the data of collisionMRT.h:55-90. D3Q19 on the orthogonal basis of
// d'Humieres et al. (2002): m = M(f - feq), f <- f - M^T diag(s/|M_j|^2) m.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int m19(int j, int i) {
  // rows: rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz, pxz, mx, my, mz
  const int x = Lat<19>::cx(i), y = Lat<19>::cy(i), z = Lat<19>::cz(i);
  const int c2 = x * x + y * y + z * z;
  switch (j) {
    case 0: return 1;
    case 1: return 19 * c2 - 30;
    case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
    case 3: return x;
    case 4: return (5 * c2 - 9) * x;
    case 5: return y;
    case 6: return (5 * c2 - 9) * y;
    case 7: return z;
    case 8: return (5 * c2 - 9) * z;
    case 9: return 3 * x * x - c2;
    case 10: return (3 * c2 - 5) * (3 * x * x - c2);
    case 11: return y * y - z * z;
    case 12: return (3 * c2 - 5) * (y * y - z * z);
    case 13: return x * y;
    case 14: return y * z;
    case 15: return x * z;
    case 16: return (y * y - z * z) * x;
    case 17: return (z * z - x * x) * y;
    default: return (x * x - y * y) * z;
  }
}

__host__ __device__ constexpr int m19_norm(int j) {
  int n = 0;
  for (int i = 0; i < 19; ++i) n += m19(j, i) * m19(j, i);
  return n;
}

// Default D3Q19 rates; a negative entry marks the shear rates (take omega).
__host__ __device__ constexpr double m19_default_rate(int j) {
  constexpr double s[19] = {0.0, 1.19, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 1.2, -1.0,
                            1.4, -1.0, 1.4, -1.0, -1.0, -1.0, 1.98, 1.98, 1.98};
  return s[j];
}

// Calls f(std::integral_constant<int, I>) for I = 0 .. N-1, each with a
// compile-time index, so that matrix entries taken from constexpr functions
// fold into constants (a loop the compiler leaves rolled would evaluate them
// at run time).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// The relaxation factor of row j, s_j / |M_j|^2. For a fixed rate it is
// formed once per block on the host (the rate over the norm in double,
// rounded to R: sn[j]); a shear row (bit j of `shear`) takes the cell's
// omega times the rounded reciprocal of its norm.
template <class R>
__device__ __forceinline__ R mrt_factor(int j, const R* sn, uint32_t shear, R om) {
  return ((shear >> j) & 1u) ? om * R(1.0 / double(m19_norm(j))) : sn[j];
}

// Rows odd in c (j, q, m): M_j(opp i) = -M_j(i); the others are even,
// M_j(opp i) = M_j(i) (every row is one or the other).
__host__ __device__ constexpr bool m19_odd(int j) {
  for (int i = 1; i < 19; i += 2)
    if (m19(j, i + 1) != -m19(j, i)) return false;
  return m19(j, 0) == 0;
}
__host__ __device__ constexpr bool m19_rows_have_parity() {
  for (int j = 0; j < 19; ++j)
    for (int i = 1; i < 19; i += 2)
      if (m19(j, i + 1) != (m19_odd(j) ? -m19(j, i) : m19(j, i))) return false;
  return true;
}
static_assert(m19_rows_have_parity(), "d'Humieres rows are even or odd in c");

// The equilibrium the MRT and LES of a cell relax towards, in the form the
// populations are held in: feq_i (FreeLB's evaluation order), or on the
// deviations feq_i - w_i = w_i (A + t_i a_i +- 2 t_i) as in bgk_dev.
template <bool DEV, class R>
__device__ __forceinline__ void mrt_eq(R (&e)[19], R d, R rho, const R (&u)[3]) {
  if constexpr (!DEV) {
    feq(e, rho, u);
  } else {
    R U[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) U[k] = R(3) * u[k];
    const R A = d - rho * (R(1.5) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const R rh = R(0.5) * rho;
    e[0] = Lat<19>::w<R>(0) * A;
#pragma unroll
    for (int i = 1; i < 19; i += 2) {
      const R ai = cdot_folded<19>(i, U);
      const R t = rh * ai;
      const R q = A + t * ai;
      e[i] = Lat<19>::w<R>(i) * (q + R(2) * t);
      e[i + 1] = Lat<19>::w<R>(i) * (q - R(2) * t);
    }
  }
}

// D3Q19 MRT on the pairs of opposite directions: with n = f - feq,
// s_p = n_i + n_opp(i) and d_p = n_i - n_opp(i) (i = 2p + 1), an even row is
// M_j0 n_0 + sum_p M_ji s_p and an odd row sum_p M_ji d_p; back, with
// k_j = (M n)_j s_j / |M_j|^2, E_p = sum over even rows of M_ji k_j and O_p
// over odd rows, f_i -= E_p + O_p and f_opp(i) -= E_p - O_p. Every sum runs
// over the nonzero entries in ascending index from zero. DEV: f holds the
// deviations f - w (d = rho - 1), relaxed towards feq - w.
template <bool DEV, class R>
__device__ __forceinline__ void mrt(R (&f)[19], R d, R rho, const R (&u)[3], R om, const R* sn, uint32_t shear) {
  R e[19];
  mrt_eq<DEV>(e, d, rho, u);
  const R n0 = f[0] - e[0];
  R sp[9], dp[9];
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const R a = f[2 * p + 1] - e[2 * p + 1];
    const R b = f[2 * p + 2] - e[2 * p + 2];
    sp[p] = a + b;
    dp[p] = a - b;
  }
  R k[19];
  static_for<19>([&](auto J) {
    constexpr int j = decltype(J)::value;
    R dm = R(0);
    if constexpr (!m19_odd(j) && m19(j, 0) != 0) dm += R(m19(j, 0)) * n0;
    static_for<9>([&](auto P) {
      constexpr int p = decltype(P)::value;
      constexpr int m = m19(j, 2 * p + 1);
      if constexpr (m != 0) dm += R(m) * (m19_odd(j) ? dp[p] : sp[p]);
    });
    k[j] = dm * mrt_factor<R>(j, sn, shear, om);
  });
  R d0 = R(0);
  static_for<19>([&](auto J) {
    constexpr int j = decltype(J)::value;
    if constexpr (!m19_odd(j) && m19(j, 0) != 0) d0 += R(m19(j, 0)) * k[j];
  });
  f[0] -= d0;
  static_for<9>([&](auto P) {
    constexpr int p = decltype(P)::value;
    R E = R(0), O = R(0);
    static_for<19>([&](auto J) {
      constexpr int j = decltype(J)::value;
      constexpr int m = m19(j, 2 * p + 1);
      if constexpr (m != 0) {
        if constexpr (m19_odd(j)) O += R(m) * k[j];
        else E += R(m) * k[j];
      }
    });
    f[2 * p + 1] -= E + O;
    f[2 * p + 2] -= E - O;
  });
}

// D3Q19 MRT, with the Smagorinsky shear rates when `les`, on the deviations
// g = f - w of the shifted storage kinds (d = rho - 1), from the moments
// themselves: m = M g through the pair sums and differences, less the
// equilibrium moments in closed form. For the second-order feq, M feq - M w
// is
//   e: -11 d + 19 rho u.u      eps: 3 d - 5.5 rho u.u    q_a: -2/3 rho u_a
//   3pxx: rho (3 ux^2 - u.u)   3pixx: -3pxx / 2          pww: rho (uy^2 - uz^2)
//   piww: -pww / 2             p_ab: rho u_a u_b          m_a: 0
// (M w is 1, -11 and 3 in rows rho, e, eps, 0 elsewhere). The conserved rows
// (rho, j) have zero non-equilibrium part and are not relaxed. The LES
// stress is read from the same moments: with T = sum c^2 n = dm_e / 19,
// Pi_xx = (dm_3pxx + T) / 3, Pi_yy, Pi_zz = (T - Pi_xx +- dm_pww) / 2 and
// Pi_xy, Pi_yz, Pi_xz the p rows, the same |Pi_neq| as smag_omega in exact
// arithmetic. No feq and no f - feq are formed.
__host__ __device__ constexpr bool m19_conserved(int j) { return j == 0 || j == 3 || j == 5 || j == 7; }

// The density and velocity come from the same pair sums and differences:
// rho - 1 = g_0 + sum_p s_p, rho u_a = sum_p c_ai d_p.
template <class R>
__device__ __forceinline__ void mrt_dev(R (&g)[19], R om, bool les, R cs, R rho_f, const R* sn, uint32_t shear) {
  R sp[9], dp[9];
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    sp[p] = g[2 * p + 1] + g[2 * p + 2];
    dp[p] = g[2 * p + 1] - g[2 * p + 2];
  }
  R d = g[0];
#pragma unroll
  for (int p = 0; p < 9; ++p) d += sp[p];
  const R rho = R(1) + d;
  const R inv = R(1) / rho;
  R u[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    R j = R(0);
    bool any = false;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
      const int c = k == 0 ? Lat<19>::cx(2 * p + 1) : (k == 1 ? Lat<19>::cy(2 * p + 1) : Lat<19>::cz(2 * p + 1));
      if (c > 0) {
        j = any ? j + dp[p] : dp[p];
        any = true;
      } else if (c < 0) {
        j = any ? j - dp[p] : -dp[p];
        any = true;
      }
    }
    u[k] = j * inv;
  }
  R dm[19];
  static_for<19>([&](auto J) {
    constexpr int j = decltype(J)::value;
    if constexpr (!m19_conserved(j)) {
      R m = R(0);
      if constexpr (!m19_odd(j) && m19(j, 0) != 0) m += R(m19(j, 0)) * g[0];
      static_for<9>([&](auto P) {
        constexpr int p = decltype(P)::value;
        constexpr int c = m19(j, 2 * p + 1);
        if constexpr (c != 0) m += R(c) * (m19_odd(j) ? dp[p] : sp[p]);
      });
      dm[j] = m;
    }
  });
  const R rx = rho * u[0], ry = rho * u[1], rz = rho * u[2];
  const R xx = rx * u[0], yy = ry * u[1], zz = rz * u[2];
  const R uu = xx + yy + zz;
  const R p3xx = R(3) * xx - uu, pww = yy - zz;
  dm[1] -= R(19) * uu - R(11) * d;
  dm[2] -= R(3) * d - R(5.5) * uu;
  dm[4] += R(2.0 / 3.0) * rx;
  dm[6] += R(2.0 / 3.0) * ry;
  dm[8] += R(2.0 / 3.0) * rz;
  dm[9] -= p3xx;
  dm[10] += R(0.5) * p3xx;
  dm[11] -= pww;
  dm[12] += R(0.5) * pww;
  dm[13] -= rx * u[1];
  dm[14] -= ry * u[2];
  dm[15] -= rx * u[2];
  R eo = om;
  if (les) {
    const R T = dm[1] * R(1.0 / 19.0);
    const R pxx = (dm[9] + T) * R(1.0 / 3.0);
    const R pyy = R(0.5) * (T - pxx + dm[11]);
    const R pzz = R(0.5) * (T - pxx - dm[11]);
    const R ns = pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (dm[13] * dm[13] + dm[14] * dm[14] + dm[15] * dm[15]);
    eo = smag_from_norm(dsqrt(ns), om, cs, rho_f);
  }
  R k[19];
  static_for<19>([&](auto J) {
    constexpr int j = decltype(J)::value;
    if constexpr (!m19_conserved(j)) k[j] = dm[j] * mrt_factor<R>(j, sn, shear, eo);
  });
  R d0 = R(0);
  static_for<19>([&](auto J) {
    constexpr int j = decltype(J)::value;
    if constexpr (!m19_conserved(j) && !m19_odd(j) && m19(j, 0) != 0) d0 += R(m19(j, 0)) * k[j];
  });
  g[0] -= d0;
  static_for<9>([&](auto P) {
    constexpr int p = decltype(P)::value;
    R E = R(0), O = R(0);
    static_for<19>([&](auto J) {
      constexpr int j = decltype(J)::value;
      constexpr int c = m19(j, 2 * p + 1);
      if constexpr (c != 0 && !m19_conserved(j)) {
        if constexpr (m19_odd(j)) O += R(c) * k[j];
        else E += R(c) * k[j];
      }
    });
    g[2 * p + 1] -= E + O;
    g[2 * p + 2] -= E - O;
  });
}

__host__ __device__ constexpr int m9(int i, int j) {
  constexpr int M[9][9] = {{1, 1, 1, 1, 1, 1, 1, 1, 1},      {-4, -1, -1, -1, -1, 2, 2, 2, 2},
                           {4, -2, -2, -2, -2, 1, 1, 1, 1},  {0, 1, -1, 0, 0, 1, -1, 1, -1},
                           {0, -2, 2, 0, 0, 1, -1, 1, -1},   {0, 0, 0, 1, -1, 1, -1, -1, 1},
                           {0, 0, 0, -2, 2, 1, -1, -1, 1},   {0, 1, 1, -1, -1, 0, 0, 0, 0},
                           {0, 0, 0, 0, 0, 1, 1, -1, -1}};
  return M[i][j];
}
__host__ __device__ constexpr int invm9_n(int i, int j) {
  constexpr int N[9][9] = {{1, -1, 1, 0, 0, 0, 0, 0, 0},     {1, -1, -1, 1, -1, 0, 0, 1, 0},
                           {1, -1, -1, -1, 1, 0, 0, 1, 0},   {1, -1, -1, 0, 0, 1, -1, -1, 0},
                           {1, -1, -1, 0, 0, -1, 1, -1, 0},  {1, 1, 1, 1, 1, 1, 1, 0, 1},
                           {1, 1, 1, -1, -1, -1, -1, 0, 1},  {1, 1, 1, 1, 1, -1, -1, 0, -1},
                           {1, 1, 1, -1, -1, 1, 1, 0, -1}};
  return N[i][j];
}
__host__ __device__ constexpr int invm9_d(int i, int j) {
  constexpr int Dn[9][9] = {{9, 9, 9, 1, 1, 1, 1, 1, 1},     {9, 36, 18, 6, 6, 1, 1, 4, 1},
                            {9, 36, 18, 6, 6, 1, 1, 4, 1},   {9, 36, 18, 1, 1, 6, 6, 4, 1},
                            {9, 36, 18, 1, 1, 6, 6, 4, 1},   {9, 18, 36, 6, 12, 6, 12, 1, 4},
                            {9, 18, 36, 6, 12, 6, 12, 1, 4}, {9, 18, 36, 6, 12, 6, 12, 1, 4},
                            {9, 18, 36, 6, 12, 6, 12, 1, 4}};
  return Dn[i][j];
}

template <class R>
__device__ __forceinline__ void mrt(R (&f)[9], R rho, const R (&u)[3], R om, const double* rates) {
  R rt[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const double sr = rates[j];
    rt[j] = (sr >= 0.0) ? R(sr) : om;
  }
  R m[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    R s = R(0);
#pragma unroll
    for (int j = 0; j < 9; ++j) s += R(m9(i, j)) * f[j];
    m[i] = s;
  }
  R dm[9];
  const R ux2 = u[0] * u[0];
  const R uy2 = u[1] * u[1];
  const R rhoux = rho * u[0];
  const R rhouy = rho * u[1];
  const R cse0 = R(3) * rho * (ux2 + uy2);
  dm[0] = m[0] - rho;
  dm[1] = m[1] + R(2) * rho - cse0;
  dm[2] = m[2] - R(9) * rho * ux2 * uy2 + cse0 - rho;
  dm[3] = m[3] - rhoux;
  dm[4] = m[4] - rhoux * (R(3) * uy2 - R(1));
  dm[5] = m[5] - rhouy;
  dm[6] = m[6] - rhouy * (R(3) * ux2 - R(1));
  dm[7] = m[7] - rho * (ux2 - uy2);
  dm[8] = m[8] - rho * u[0] * u[1];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    R s = R(0);
#pragma unroll
    for (int j = 0; j < 9; ++j) s += ((R(invm9_n(i, j)) / R(invm9_d(i, j))) * rt[j]) * dm[j];
    f[i] -= s;
  }
}

}  // namespace flb
