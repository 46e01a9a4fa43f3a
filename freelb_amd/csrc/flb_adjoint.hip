// Differentiable D3Q19 lattice update: forward BGK with a body force and a
// Brinkman drag, the hand-written discrete adjoint of that step, and the
// checkpoint / recompute driver, behind the C ABI (include/flb_abi.h).
//
// Model (one time step, every fluid cell x):
//   rho = sum_i f_i,  j = sum_i c_i f_i,  D = 1 + alpha*s(x)/2
//   u   = (j + G/2) / (rho*D)                 (FreeLB's forcerhoU, lbm/moment.h:401-438,
//                                             with the implicit drag folded in)
//   F   = G - alpha*s(x)*rho*u                 (body force density G, Brinkman drag)
//   F_i = w_i [3 (c_i - u) + 9 (c_i . u) c_i] . F        (ForcePopImpl, lbm/force.h:29-36)
//   f*_i = omega feq_i(rho, u) + (1 - omega) f_i + (1 - omega/2) F_i
//                                             (BGKForce, lbm/collision.h:61-90)
//   f_i(x + c_i, t+1) = f*_i(x), or f_{opp i}(x, t+1) = f*_i(x) when x + c_i is
//   solid or outside a non-periodic axis (halfway bounce-back,
//   boundary/bounce_back_boundary.h:169-181).
// Objective: J = (1/T) sum_{t=1..T} sum_{x on plane x = outlet} sum_i c_ix f_i(x, t),
// summed in FP64 inside the forward step that writes state t (obj_fold).
//
// Adjoint. With lambda^t the adjoint of f^t and mu^t = P^T lambda^{t+1} the
// adjoint of the post-collision populations (mu_i(x) = lambda_i(x + c_i), or
// lambda_{opp i}(x) across a wall link), one backward step is per cell
//   lambda_j = dJ/df_j + (1 - omega) mu_j + A_rho + A_u . (c_j/(rho D) - u/rho)
//   dJ/ds   += A_u . (-alpha u / (2 D)) - (1 - omega/2) alpha rho (B . u)
// with the weighted moments of mu, M = sum w_i mu_i, V = sum w_i mu_i c_i,
// T = sum w_i mu_i c_i c_i^T, and
//   B     = 3 V - 3 M u + 9 T u          (= sum mu_i dF_i/dF = sum mu_i dfeq_i/du / rho)
//   A_rho = omega [M (1 - 1.5 u.u) + 3 V.u + 4.5 u.T.u] - (1 - omega/2) alpha s (B . u)
//   A_u   = omega rho B + (1 - omega/2) [-3 M F + 9 T F - alpha s rho B].
// No reference implementation exists (FreeLB has no adjoint); the derivation
// is pinned by FP64 finite differences and by autodiff of an FP64 CPU model.
//
// Layout: one dense array per direction, [19][N], N = nx*ny*nz,
// id = x + nx*(y + ny*z) (FreeLB's index order); states are kept whole for
// checkpoints, so no AA parity is involved. The adjoint step reads a forward
// state only through its moment record, [4][N]: rho - 1 and u, written by the
// forward step that reads that state (or by k_mom). The arrays hold g_i = f_i - w_i
// and all moments are formed from g (rho = 1 + sum g, j = sum c g): the
// velocities of a slow channel flow (u ~ 1e-5) then keep full precision in
// FP32 instead of drowning in the rounding of f ~ w.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/flb_abi.h"
#include "flb_cell.hpp"
#include "flb_host.hpp"

namespace flb {
namespace adj {

constexpr int Q = 19;
constexpr int kThreads = 256;

template <class R>
struct AdjArgs {
  int nx, ny, nz;
  int px, py, pz;
  int64_t N;
  const uint8_t* solid;  // nonzero: not updated, links into it bounce back
  const uint32_t* link;  // per cell: bit 0 solid, bit i the link along c_i is a wall
  const R* s;            // Brinkman solid fraction per cell
  R omega;
  R G[3];
  R alpha;
  int outlet;            // x of the objective plane
  R invT;                // 1 / T
  R om1, fom, Gh[3];     // 1 - omega, 1 - omega / 2, G / 2
  uint32_t sy, sz, sn;   // bytes between y rows, z planes and population arrays
  uint32_t mx, shx, px1; // div_magic multiplier, shift and pass mask of nx
  uint32_t my, shy, py1; // ... of ny
  int solids;            // nonzero: some cell is solid (else links are wall links only at faces)
};

// Neighbour id of (x, y, z) along c_i, or -1 for a wall link.
template <class R>
__device__ __forceinline__ int64_t neighbour(const AdjArgs<R>& a, int x, int y, int z, int i) {
  int xn = x + Lat<Q>::cx(i), yn = y + Lat<Q>::cy(i), zn = z + Lat<Q>::cz(i);
  if (xn < 0 || xn >= a.nx) {
    if (!a.px) return -1;
    xn = xn < 0 ? xn + a.nx : xn - a.nx;
  }
  if (yn < 0 || yn >= a.ny) {
    if (!a.py) return -1;
    yn = yn < 0 ? yn + a.ny : yn - a.ny;
  }
  if (zn < 0 || zn >= a.nz) {
    if (!a.pz) return -1;
    zn = zn < 0 ? zn + a.nz : zn - a.nz;
  }
  const int64_t id = int64_t(xn) + int64_t(a.nx) * (yn + int64_t(a.ny) * zn);
  return a.solid[id] ? -1 : id;
}

// r += sgn * v, where the first term of a sum replaces r (no 0 + v is formed,
// so sums over the nonzero components of c_i cost one add per term). `any`
// is known at compile time once the loops are unrolled.
template <class R>
__device__ __forceinline__ void acc(R& r, bool& any, R v, int sgn) {
  if (sgn == 0) return;
  const R t = sgn > 0 ? v : -v;
  r = any ? (sgn > 0 ? r + v : r - v) : t;
  any = true;
}

// Sum and difference of every opposite pair (i = 2p + 1, opp(i) = i + 1).
template <class R>
__device__ __forceinline__ void pair_split(const R (&f)[Q], R (&sp)[Q / 2], R (&dp)[Q / 2]) {
#pragma unroll
  for (int p = 0; p < Q / 2; ++p) {
    sp[p] = f[2 * p + 1] + f[2 * p + 2];
    dp[p] = f[2 * p + 1] - f[2 * p + 2];
  }
}

// 1 / (rho D) and the force F = G - alpha s rho u of a cell. The forward step
// and the adjoint step (from the forward step's moment record) form them by
// the same expressions, so both see the same bits.
template <class R>
__device__ __forceinline__ R inv_rho_d(R rho, R s, const AdjArgs<R>& a) {
  return R(1) / (rho * (R(1) + (a.alpha * s) * R(0.5)));
}
template <class R>
__device__ __forceinline__ void drag_force(R rho, const R (&u)[3], R s, const AdjArgs<R>& a, R (&F)[3]) {
  const R asr = (a.alpha * s) * rho;
#pragma unroll
  for (int d = 0; d < 3; ++d) F[d] = a.G[d] - asr * u[d];
}

// rho - 1, rho, u, F and 1 / (rho D) of a cell from its shifted populations g
// (the forward model above), the sums taken over opposite pairs.
template <class R>
__device__ __forceinline__ void force_state(const R (&g)[Q], R s, const AdjArgs<R>& a, R& drho, R& rho,
                                            R (&u)[3], R (&F)[3], R& irD) {
  R sp[Q / 2], dp[Q / 2];
  pair_split(g, sp, dp);
  drho = g[0];
#pragma unroll
  for (int p = 0; p < Q / 2; ++p) drho += sp[p];
  R j[3] = {R(0), R(0), R(0)};
  bool any[3] = {false, false, false};
#pragma unroll
  for (int p = 0; p < Q / 2; ++p) {
    acc(j[0], any[0], dp[p], Lat<Q>::cx(2 * p + 1));
    acc(j[1], any[1], dp[p], Lat<Q>::cy(2 * p + 1));
    acc(j[2], any[2], dp[p], Lat<Q>::cz(2 * p + 1));
  }
  rho = R(1) + drho;
  irD = inv_rho_d(rho, s, a);
#pragma unroll
  for (int d = 0; d < 3; ++d) u[d] = (j[d] + a.Gh[d]) * irD;
  drag_force(rho, u, s, a, F);
}

// Post-collision values v_i = omega feq_i + (1 - omega) g_i + (1 - omega/2) F_i
// (shifted by w_i). With P = omega rho u, H = (1 - omega/2) F and, per
// direction, a = c.P, b = c.H, cu = c.u:
//   v_i - (1 - omega) g_i = w_i [A0 + 9 cu (a/2 + b) + 3 (a + b)],
//   A0 = omega (rho - 1) - 1.5 omega rho u.u - 3 (1 - omega/2) u.F;
// the bracket's first two terms are even in c and the last odd, so each
// opposite pair shares them.
// sink(i, v_i) takes each value as soon as it is formed (the forward step
// stores it there, spreading its stores over the arithmetic).
template <class R, class Sink>
__device__ __forceinline__ void post_collision(const R (&g)[Q], R drho, R rho, const R (&u)[3],
                                               const R (&F)[3], const AdjArgs<R>& a, R (&v)[Q], Sink&& sink) {
  const R om = a.omega, om1 = a.om1, fom = a.fom;
  const R orho = om * rho;
  const R usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const R uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
  const R A0 = om * drho - R(1.5) * orho * usq - R(3) * fom * uF;
  const R P[3] = {orho * u[0], orho * u[1], orho * u[2]};
  const R H[3] = {fom * F[0], fom * F[1], fom * F[2]};
  v[0] = om1 * g[0] + Lat<Q>::template w<R>(0) * A0;
  sink(0, v[0]);
#pragma unroll
  for (int p = 0; p < Q / 2; ++p) {
    const int i = 2 * p + 1;
    const R cu = cdot_folded<Q>(i, u), ca = cdot_folded<Q>(i, P), cb = cdot_folded<Q>(i, H);
    const R w = Lat<Q>::template w<R>(i);
    const R ws = w * (A0 + R(9) * cu * (R(0.5) * ca + cb));
    const R wa = (R(3) * w) * (ca + cb);
    v[i] = om1 * g[i] + (ws + wa);
    v[i + 1] = om1 * g[i + 1] + (ws - wa);
    sink(i, v[i]);
    sink(i + 1, v[i + 1]);
  }
}

// feq_i(rho, u) - w_i, formed without cancellation.
template <class R>
__device__ __forceinline__ R geq(int i, R drho, R rho, const R (&u)[3], R usq);

template <class R>
__device__ __forceinline__ R cdot(int i, const R (&v)[3]) {
  return R(Lat<Q>::cx(i)) * v[0] + R(Lat<Q>::cy(i)) * v[1] + R(Lat<Q>::cz(i)) * v[2];
}

template <class R>
__device__ __forceinline__ R geq(int i, R drho, R rho, const R (&u)[3], R usq) {
  const R cu = cdot(i, u);
  return Lat<Q>::template w<R>(i) * (drho + rho * (R(3) * cu + R(4.5) * cu * cu - R(1.5) * usq));
}

// Boxes hold at most 2^30 cells (checked at creation), so the index splits
// with 32-bit divisions (64-bit ones cost several times the instructions).
__device__ __forceinline__ void cell_xyz(int64_t id, int nx, int ny, int& x, int& y, int& z) {
  const unsigned i = unsigned(id), yz = i / unsigned(nx);
  x = int(i - yz * unsigned(nx));
  const unsigned zz = yz / unsigned(ny);
  y = int(yz - zz * unsigned(ny));
  z = int(zz);
}

// ---- link words: bit 0 the cell is solid, bit i its neighbour along c_i is a
// wall (solid, or beyond a non-periodic face). Rebuilt whenever the solid mask
// changes; the steps then test one word per cell instead of loading the 18
// neighbours' solid flags.
template <class R>
__global__ __launch_bounds__(kThreads) void k_links(AdjArgs<R> a, uint32_t* __restrict__ link) {
  const int64_t id = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (id >= a.N) return;
  int x, y, z;
  cell_xyz(id, a.nx, a.ny, x, y, z);
  uint32_t w = a.solid[id] ? 1u : 0u;
#pragma unroll
  for (int i = 1; i < Q; ++i)
    if (neighbour(a, x, y, z, i) < 0) w |= 1u << i;
  link[id] = w;
}

// q = n / d for n < 2^31 by one high multiply. For d >= 2, with
// l = ceil(log2 d), m = ceil(2^(31+l) / d) fits 32 bits and n m / 2^(31+l)
// exceeds n / d by less than n / (d 2^31) < 1 / d, so its floor is the
// quotient (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994); sh = l - 1 and pass = 0. d == 1 has m = 0,
// sh = 0 and pass = ~0, so that n passes through. No branch: the batched
// steps read m, sh and pass from a table, and a branch on them would hold
// back the other reads of the table behind it.
__device__ __forceinline__ uint32_t div_magic(uint32_t n, uint32_t m, uint32_t sh, uint32_t pass) {
  return (__umulhi(n, m) >> sh) + (n & pass);
}

// A cell's x and the byte offsets, in one [N] population array, of the cell
// (o) and of its neighbours. The neighbour along c_i sits at
//   X(c_ix) + Y(c_iy) + Z(c_iz) + c_iy sy + c_iz sz,
// X(0) = o and X(+-1) = oxp, oxm (o +- one element, wrapped across the x
// faces); Y(+-1) = yp, ym and Z(+-1) = zp, zm are the corrections of a wrap
// across a y or z face (0 inside). Wraps are taken on every axis, periodic or
// not, so every offset names a cell of the box; a wall link's offset is then
// in bounds and never used. The sums run modulo 2^32: every true offset lies
// in [0, N size) and N size <= 2^32 (checked at creation).
struct Hops {
  uint32_t x, y, z, o, oxp, oxm, yp, ym, zp, zm;
};

template <class R>
__device__ __forceinline__ Hops hops(const AdjArgs<R>& a, uint32_t id) {
  constexpr uint32_t e = sizeof(R);
  const uint32_t yz = div_magic(id, a.mx, a.shx, a.px1), z = div_magic(yz, a.my, a.shy, a.py1);
  const uint32_t x = id - yz * uint32_t(a.nx), y = yz - z * uint32_t(a.ny);
  Hops h;
  h.x = x;
  h.y = y;
  h.z = z;
  h.o = id * e;
  h.oxp = h.o + (x == uint32_t(a.nx - 1) ? e - a.sy : e);
  h.oxm = h.o - (x == 0u ? e - a.sy : e);
  h.yp = y == uint32_t(a.ny - 1) ? 0u - a.sz : 0u;
  h.ym = y == 0u ? a.sz : 0u;
  h.zp = z == uint32_t(a.nz - 1) ? 0u - a.sn : 0u;
  h.zm = z == 0u ? a.sn : 0u;
  return h;
}

// Byte offset of the neighbour along c_i (i > 0), or of the cell itself (i == 0).
template <class R>
__device__ __forceinline__ uint32_t hop(const Hops& h, const AdjArgs<R>& a, int i) {
  const int cx = Lat<Q>::cx(i), cy = Lat<Q>::cy(i), cz = Lat<Q>::cz(i);
  uint32_t o = cx > 0 ? h.oxp : (cx < 0 ? h.oxm : h.o);
  if (cy != 0) o += cy > 0 ? h.yp : h.ym;
  if (cz != 0) o += cz > 0 ? h.zp : h.zm;
  return o + uint32_t(cy) * a.sy + uint32_t(cz) * a.sz;
}

// Directions whose component along `axis` equals sgn (bit i for c_i).
constexpr uint32_t face_mask(int axis, int sgn) {
  uint32_t m = 0;
  for (int i = 1; i < Q; ++i) {
    const int c = axis == 0 ? Lat<Q>::cx(i) : (axis == 1 ? Lat<Q>::cy(i) : Lat<Q>::cz(i));
    if (c == sgn) m |= 1u << i;
  }
  return m;
}

// The link word of a cell of a box without solid cells, from its place: the
// links that leave the box through a non-periodic face (what k_links stores
// for such a box), with no load.
template <class R>
__device__ __forceinline__ uint32_t face_links(const AdjArgs<R>& a, const Hops& h) {
  uint32_t lk = 0;
  if (!a.px) {
    if (h.x == uint32_t(a.nx - 1)) lk |= face_mask(0, 1);
    if (h.x == 0u) lk |= face_mask(0, -1);
  }
  if (!a.py) {
    if (h.y == uint32_t(a.ny - 1)) lk |= face_mask(1, 1);
    if (h.y == 0u) lk |= face_mask(1, -1);
  }
  if (!a.pz) {
    if (h.z == uint32_t(a.nz - 1)) lk |= face_mask(2, 1);
    if (h.z == 0u) lk |= face_mask(2, -1);
  }
  return lk;
}

// Population i at byte offset `off` of a [19][N] state: the direction's
// array base, wave-uniform, plus a zero-extended 32-bit lane offset, which the
// global load and store instructions take as is (scalar base, vector offset),
// so an access costs no 64-bit address arithmetic. The empty asm keeps the
// base an opaque scalar value: without it the compiler folds i N into the
// lane offset and forms a 64-bit address per lane and access.
template <class T>
using gptr = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<const char> sbase(const T* f) {
  gptr<const char> p = (gptr<const char>)f;
  asm("" : "+s"(p));
  return p;
}
template <class R>
__device__ __forceinline__ gptr<const char> dir_base(const R* f, int64_t N, int i) {
  return sbase(f + int64_t(i) * N);
}
// Element of a per-cell [N] array of R at the cell's byte offset.
template <class R>
__device__ __forceinline__ R ld_cell(const R* f, uint32_t off) {
  return *(gptr<const R>)(sbase(f) + off);
}
template <class R>
__device__ __forceinline__ gptr<R> cell_ref(R* f, uint32_t off) {
  return (gptr<R>)(gptr<char>)(sbase(f) + off);
}
template <class R>
__device__ __forceinline__ R ld_pop(const R* __restrict__ f, int64_t N, int i, uint32_t off) {
  return *(gptr<const R>)(dir_base(f, N, i) + off);
}
template <class R>
__device__ __forceinline__ void st_pop(R* __restrict__ f, int64_t N, int i, uint32_t off, R v) {
  *(gptr<R>)(gptr<char>)(dir_base(f, N, i) + off) = v;
}

// Uses the loaded values at this point, so that their loads stay ahead of
// the solid-cell test that follows (the compiler would sink them below it,
// behind the link word's latency). Waits for them, issues nothing.
template <class R, int M>
__device__ __forceinline__ void pin(const R (&v)[M]) {
#pragma unroll
  for (int i = 0; i < M; ++i) asm volatile("" ::"v"(v[i]));
}

// ---- forward step: state src -> dst (push, halfway bounce-back) -------------
// OBJ: also the cell's share of the objective of state dst, the FP64 sum, in
// direction order, of c_jx times every value it pushes into population j of a
// cell on the outlet plane (its own, across a wall link, or the neighbour's);
// every population of the plane's cells is pushed by exactly one cell.
// Wavefronts without a wall link store without testing links, and those with
// no cell on or beside the outlet plane skip the objective.
// MOM: also the moment record of the source state, [4][N]: rho - 1 and u
// (all the adjoint step needs of it).
template <class R, bool OBJ, bool MOM>
__device__ __forceinline__ double fwd_cell(const AdjArgs<R>& a, const R* __restrict__ src,
                                           R* __restrict__ dst, R* __restrict__ mom, int64_t id) {
  double c = 0.0;
  if (id >= a.N) return c;
  // the populations are read together with the link word (every offset is in
  // bounds, a solid cell's too), not after it
  const Hops h = hops(a, uint32_t(id));
  R g[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) g[i] = ld_pop(src, a.N, i, h.o);
  const R s = ld_cell(a.s, h.o);
  uint32_t lk;
  if (a.solids) {
    lk = ((gptr<const uint32_t>)a.link)[id];
    pin(g);
    asm volatile("" : "+v"(lk));  // tests of lk come after every load above
    if (lk & 1u) return c;
  } else {
    lk = face_links(a, h);
  }
  R drho, rho, u[3], F[3], irD;
  force_state(g, s, a, drho, rho, u, F, irD);
  if constexpr (MOM) {
    st_pop(mom, a.N, 0, h.o, drho);
#pragma unroll
    for (int d = 0; d < 3; ++d) st_pop(mom, a.N, d + 1, h.o, u[d]);
  }
  R v[Q];
  if (__ballot(lk != 0u) == 0) {
    post_collision(g, drho, rho, u, F, a, v, [&](int i, R x) { st_pop(dst, a.N, i, hop(h, a, i), x); });
  } else {
    post_collision(g, drho, rho, u, F, a, v, [&](int i, R x) {
      if (i != 0 && ((lk >> i) & 1u))
        st_pop(dst, a.N, opp(i), h.o, x);
      else
        st_pop(dst, a.N, i, hop(h, a, i), x);
    });
  }
  if constexpr (OBJ) {
    const uint32_t x = h.x, last = uint32_t(a.nx - 1), out = uint32_t(a.outlet);
    const bool on0 = x == out, onp = (x == last ? 0u : x + 1u) == out, onm = (x == 0u ? last : x - 1u) == out;
    if (__ballot(on0 || onp || onm) != 0) {
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        const int cx = Lat<Q>::cx(i);
        if (cx == 0) continue;
        const bool wall = (lk >> i) & 1u;
        // across a wall the value lands in opp(i) of this cell (c_x = -cx)
        const bool hit = wall ? on0 : (cx > 0 ? onp : onm);
        const bool up = wall ? cx < 0 : cx > 0;
        const double dv = double(v[i]);
        if (hit) c += up ? dv : -dv;
      }
    }
  }
  return c;
}

// Fixed-order sum of v over the workgroup (every thread calls it; thread 0
// gets the result): a shuffle tree within each wavefront, then the
// wavefronts in order.
__device__ __forceinline__ double wg_sum(double v) {
  __shared__ double ws[kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = ws[0];
#pragma unroll
  for (int w = 1; w < kThreads / 64; ++w) s += ws[w];
  __syncthreads();
  return s;
}

// Every thread of the workgroup holding chunk `chunk` (256 cells of the box)
// calls this with its cell's share; the chunk's fixed-order sum goes to
// partial[chunk]. Nothing is shared between workgroups (no atomics): k_obj_sum
// adds the chunks of every step after the run, in chunk order, so the sums do
// not depend on the order in which workgroups run, and a box stepped in a
// batch (chunks aligned to 256 cells there too) gives the same bits as
// stepped alone.
__device__ __forceinline__ void obj_fold(double c, double* __restrict__ partial, unsigned chunk) {
  const double s = wg_sum(c);
  if (threadIdx.x == 0) partial[chunk] = s;
}

// Workgroup t (t = 1 .. T): jt[t] = sum of the nchunks partial sums of step t
// (partial + t * nchunks), each thread over a fixed stride, then wg_sum.
__device__ __forceinline__ void obj_sum(const double* __restrict__ partial, unsigned nchunks, int t,
                                        double* __restrict__ jt) {
  const double* p = partial + int64_t(t) * nchunks;
  double v = 0.0;
  for (unsigned k = threadIdx.x; k < nchunks; k += kThreads) v += p[k];
  const double s = wg_sum(v);
  if (threadIdx.x == 0) jt[t] = s;
}

__global__ __launch_bounds__(kThreads) void k_obj_sum(const double* __restrict__ partial, unsigned nchunks,
                                                      double* __restrict__ jt) {
  obj_sum(partial, nchunks, int(blockIdx.x) + 1, jt);
}

// partial: the step's chunk sums (OBJ)
template <class R, bool OBJ, bool MOM>
__global__ __launch_bounds__(kThreads) void k_fwd(AdjArgs<R> a, const R* __restrict__ src, R* __restrict__ dst,
                                                  R* __restrict__ mom, double* __restrict__ partial) {
  const unsigned wg = xcd_block(blockIdx.x, gridDim.x);
  const double c = fwd_cell<R, OBJ, MOM>(a, src, dst, mom, int64_t(wg) * kThreads + threadIdx.x);
  if constexpr (OBJ) obj_fold(c, partial, wg);
}

// ---- the moment record alone (the last state of a recomputed segment) ------------
template <class R>
__device__ __forceinline__ void mom_cell(const AdjArgs<R>& a, const R* __restrict__ src, R* __restrict__ mom,
                                         int64_t id) {
  if (id >= a.N) return;
  const uint32_t lk = ((gptr<const uint32_t>)a.link)[id];
  if (lk & 1u) return;
  const uint32_t o = uint32_t(id) * uint32_t(sizeof(R));
  R g[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) g[i] = ld_pop(src, a.N, i, o);
  R drho, rho, u[3], F[3], irD;
  force_state(g, ld_cell(a.s, o), a, drho, rho, u, F, irD);
  st_pop(mom, a.N, 0, o, drho);
#pragma unroll
  for (int d = 0; d < 3; ++d) st_pop(mom, a.N, d + 1, o, u[d]);
}

template <class R>
__global__ __launch_bounds__(kThreads) void k_mom(AdjArgs<R> a, const R* __restrict__ src, R* __restrict__ mom) {
  mom_cell(a, src, mom, int64_t(blockIdx.x) * kThreads + threadIdx.x);
}

// ---- equilibrium initial state ----------------------------------------------
template <class R>
__global__ __launch_bounds__(kThreads) void k_init(AdjArgs<R> a, const R* __restrict__ rhou,
                                                   R* __restrict__ f) {
  const int64_t id = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (id >= a.N) return;
  const R u[3] = {rhou[a.N + id], rhou[2 * a.N + id], rhou[3 * a.N + id]};
  const R rho = rhou[id], usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const bool sol = a.solid[id] != 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) f[int64_t(i) * a.N + id] = sol ? R(0) : geq(i, rho - R(1), rho, u, usq);
}

// ---- adjoint seed: lambda^T = dJ/df^T -----------------------------------------
template <class R>
__device__ __forceinline__ void seed_cell(const AdjArgs<R>& a, R* __restrict__ lam, int64_t id) {
  if (id >= a.N) return;
  const bool on = !a.solid[id] && int(id % a.nx) == a.outlet;
#pragma unroll
  for (int i = 0; i < Q; ++i) lam[int64_t(i) * a.N + id] = on ? R(Lat<Q>::cx(i)) * a.invT : R(0);
}

template <class R>
__global__ __launch_bounds__(kThreads) void k_seed(AdjArgs<R> a, R* __restrict__ lam) {
  seed_cell(a, lam, int64_t(blockIdx.x) * kThreads + threadIdx.x);
}

// ---- one backward step: lambda^{t+1}, f^t -> lambda^t, grad += ds ----------------
// The weighted moments of mu are summed per weight class over opposite pairs
// (M from the pair sums, V from the differences, T from the sums), and
// lambda^t is written pairwise: beside their own (1 - omega) mu terms,
// lambda_i and lambda_opp(i) share A_rho - A_u.u / rho and differ in the sign
// of c_i.A_u / (rho D).
// f^t enters only through rho and u, read from its moment record mom ([4][N]:
// rho - 1, u), 16 bytes per cell in FP32 instead of the 76 of the populations.
template <class R>
__device__ __forceinline__ void bwd_cell(const AdjArgs<R>& a, const R* __restrict__ mom,
                                         const R* __restrict__ lam1, R* __restrict__ lam0,
                                         R* __restrict__ grad, int with_obj, int64_t id) {
  if (id >= a.N) return;
  // mu = P^T lambda^{t+1}: every lane first reads lambda_i(x + c_i), issued
  // together with the moments and the link word (every offset is in bounds); a
  // wall link then takes lambda_opp(i)(x) instead
  const Hops h = hops(a, uint32_t(id));
  R m[4], mu[Q];
#pragma unroll
  for (int k = 0; k < 4; ++k) m[k] = ld_pop(mom, a.N, k, h.o);
  const R s = ld_cell(a.s, h.o);
  uint32_t lk;
  if (a.solids) {
#pragma unroll
    for (int i = 0; i < Q; ++i) mu[i] = ld_pop(lam1, a.N, i, hop(h, a, i));
    lk = ((gptr<const uint32_t>)a.link)[id];
    pin(m);
    pin(mu);
    asm volatile("" : "+v"(lk));  // tests of lk come after every load above
    if (lk & 1u) return;
  } else {
    // wall links known from the place, not from memory: the bounce-back
    // loads below go out in the same round as these
    lk = face_links(a, h);
#pragma unroll
    for (int i = 0; i < Q; ++i) mu[i] = ld_pop(lam1, a.N, i, hop(h, a, i));
  }
  if (__ballot(lk != 0u) != 0) {
#pragma unroll
    for (int i = 1; i < Q; ++i)
      if ((lk >> i) & 1u) mu[i] = ld_pop(lam1, a.N, opp(i), h.o);
  }
  const R rho = R(1) + m[0];
  const R u[3] = {m[1], m[2], m[3]};
  const R irD = inv_rho_d(rho, s, a);
  R F[3];
  drag_force(rho, u, s, a, F);
  R sm[Q / 2], dm[Q / 2];
  pair_split(mu, sm, dm);
  // per weight class k (1: axes, 2: diagonals): sums of mu, c mu and c c mu
  R Ms[3] = {mu[0], R(0), R(0)}, Vs[3][3] = {}, Ts[3][6] = {};
  bool mo[3] = {true, false, false}, vo[3][3] = {}, to[3][6] = {};
#pragma unroll
  for (int p = 0; p < Q / 2; ++p) {
    const int i = 2 * p + 1, k = Lat<Q>::wclass(i);
    const int c[3] = {Lat<Q>::cx(i), Lat<Q>::cy(i), Lat<Q>::cz(i)};
    acc(Ms[k], mo[k], sm[p], 1);
#pragma unroll
    for (int d = 0; d < 3; ++d) acc(Vs[k][d], vo[k][d], dm[p], c[d]);
    acc(Ts[k][0], to[k][0], sm[p], c[0] * c[0]);
    acc(Ts[k][1], to[k][1], sm[p], c[1] * c[1]);
    acc(Ts[k][2], to[k][2], sm[p], c[2] * c[2]);
    acc(Ts[k][3], to[k][3], sm[p], c[0] * c[1]);
    acc(Ts[k][4], to[k][4], sm[p], c[0] * c[2]);
    acc(Ts[k][5], to[k][5], sm[p], c[1] * c[2]);
  }
  const R w0 = Lat<Q>::template w<R>(0), w1 = Lat<Q>::template w<R>(1), w2 = Lat<Q>::template w<R>(7);
  // weighted sum of the two classes' partial sums (a class may hold none)
  auto wsum = [&](R s1, bool o1, R s2, bool o2) -> R {
    return o1 ? (o2 ? w1 * s1 + w2 * s2 : w1 * s1) : (o2 ? w2 * s2 : R(0));
  };
  const R M = w0 * Ms[0] + w1 * Ms[1] + w2 * Ms[2];
  R V[3], T[6];
#pragma unroll
  for (int d = 0; d < 3; ++d) V[d] = wsum(Vs[1][d], vo[1][d], Vs[2][d], vo[2][d]);
#pragma unroll
  for (int e = 0; e < 6; ++e) T[e] = wsum(Ts[1][e], to[1][e], Ts[2][e], to[2][e]);
  const R Txx = T[0], Tyy = T[1], Tzz = T[2], Txy = T[3], Txz = T[4], Tyz = T[5];
  const R Tu[3] = {Txx * u[0] + Txy * u[1] + Txz * u[2], Txy * u[0] + Tyy * u[1] + Tyz * u[2],
                   Txz * u[0] + Tyz * u[1] + Tzz * u[2]};
  const R TF[3] = {Txx * F[0] + Txy * F[1] + Txz * F[2], Txy * F[0] + Tyy * F[1] + Tyz * F[2],
                   Txz * F[0] + Tyz * F[1] + Tzz * F[2]};
  R B[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) B[d] = R(3) * (V[d] - M * u[d]) + R(9) * Tu[d];
  const R om = a.omega, fom = a.fom;
  const R uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const R Vu = V[0] * u[0] + V[1] * u[1] + V[2] * u[2];
  const R uTu = u[0] * Tu[0] + u[1] * Tu[1] + u[2] * Tu[2];
  const R Bu = B[0] * u[0] + B[1] * u[1] + B[2] * u[2];
  const R as = a.alpha * s, asr = as * rho;
  const R Arho = om * (M * (R(1) - R(1.5) * uu) + R(3) * Vu + R(4.5) * uTu) - fom * as * Bu;
  R Au[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) Au[d] = (om * rho - fom * asr) * B[d] + fom * (R(9) * TF[d] - R(3) * M * F[d]);
  const R Auu = Au[0] * u[0] + Au[1] * u[1] + Au[2] * u[2];
  // 1 / D = rho / (rho D) and 1 / rho = D / (rho D)
  const R D = R(1) + as * R(0.5);
  const gptr<R> gr = cell_ref(grad, h.o);
  *gr += -(R(0.5) * a.alpha * rho * irD) * Auu - fom * a.alpha * rho * Bu;
  const R base = Arho - Auu * (D * irD);
  // c_j.A_u / (rho D), with the objective's c_jx / T folded into the x part
  R Ar[3] = {Au[0] * irD, Au[1] * irD, Au[2] * irD};
  if (with_obj && h.x == uint32_t(a.outlet)) Ar[0] += a.invT;
  const R om1 = a.om1;
  // the cell's offset as a value of this block, so that the stores take it as
  // their 32-bit lane offset (a value carried in from an earlier block reaches
  // them widened to 64 bits, and each store then forms a 64-bit address)
  uint32_t o = h.o;
  asm volatile("" : "+v"(o));
  st_pop(lam0, a.N, 0, o, om1 * mu[0] + base);
#pragma unroll
  for (int p = 0; p < Q / 2; ++p) {
    const int i = 2 * p + 1;
    const R ca = cdot_folded<Q>(i, Ar);
    st_pop(lam0, a.N, i, o, om1 * mu[i] + (base + ca));
    st_pop(lam0, a.N, i + 1, o, om1 * mu[i + 1] + (base - ca));
  }
}

template <class R>
__global__ __launch_bounds__(kThreads) void k_bwd(AdjArgs<R> a, const R* __restrict__ mom,
                                                  const R* __restrict__ lam1, R* __restrict__ lam0,
                                                  R* __restrict__ grad, int with_obj) {
  bwd_cell(a, mom, lam1, lam0, grad, with_obj, int64_t(xcd_block(blockIdx.x, gridDim.x)) * kThreads + threadIdx.x);
}

// ---- many boxes in one launch ------------------------------------------------------
// Entry e: the box's arguments, its state buffers by slot number (the same
// slot schedule for every box, since all share T and the checkpoint
// interval), its gradient and objective arrays, and its first global cell
// (a multiple of 64, so a wavefront never spans two boxes).
template <class R>
struct Entry {
  AdjArgs<R> a;
  R* const* slots;
  R* grad;
  double* jt;
  double* partial;  // objective partial sums, [T + 1][chunks of 256 cells]
  int64_t cell0;    // first cell in the batch numbering (a multiple of 256)
};

// Buffer `k` of a box's slot table, read through the constant address space:
// the table entry is the same for every lane, so this is a scalar load and
// the pointer lands in scalar registers, as the steps need for their bases.
template <class R>
__device__ __forceinline__ R* slot(const Entry<R>& E, int k) {
  typedef R* P;
  return ((const __attribute__((address_space(4))) P*)E.slots)[k];
}

// The box of logical workgroup g: the table holds, after its n entries, one
// int per workgroup of the grid (every workgroup covers 256 cells of one box).
// A scalar load, in place of a search over the entries' first cells.
template <class R>
__device__ __forceinline__ const Entry<R>& entry_of(const Entry<R>* tab, int n, unsigned g) {
  return tab[((const __attribute__((address_space(4))) int*)(tab + n))[g]];
}

// One workgroup covers 256 cells of one box (boxes start on multiples of
// 256 in the batch numbering).
// mom: the slot of the moment record (MOM)
template <class R, bool OBJ, bool MOM>
__global__ __launch_bounds__(kThreads) void k_fwd_b(const Entry<R>* __restrict__ tab, int n, int64_t total,
                                                    int src, int dst, int mom, int t) {
  const unsigned g = xcd_block(blockIdx.x, gridDim.x);
  const int64_t k0 = int64_t(g) * kThreads;
  if (k0 >= total) return;
  const Entry<R>& E = entry_of(tab, n, g);
  const int64_t id0 = k0 - E.cell0;
  const double c =
      fwd_cell<R, OBJ, MOM>(E.a, slot(E, src), slot(E, dst), MOM ? slot(E, mom) : nullptr, id0 + threadIdx.x);
  if constexpr (OBJ) {
    const unsigned nchunks = unsigned((E.a.N + kThreads - 1) / kThreads);
    obj_fold(c, E.partial + int64_t(t) * nchunks, unsigned(id0 / kThreads));
  }
}

// grid (T, boxes)
template <class R>
__global__ __launch_bounds__(kThreads) void k_obj_sum_b(const Entry<R>* __restrict__ tab) {
  const Entry<R>& E = tab[blockIdx.y];
  obj_sum(E.partial, unsigned((E.a.N + kThreads - 1) / kThreads), int(blockIdx.x) + 1, E.jt);
}

template <class R>
__global__ __launch_bounds__(kThreads) void k_seed_b(const Entry<R>* __restrict__ tab, int n, int64_t total,
                                                     int lam) {
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= total) return;
  const Entry<R>& E = entry_of(tab, n, blockIdx.x);
  seed_cell(E.a, E.slots[lam], k - E.cell0);
}

template <class R>
__global__ __launch_bounds__(kThreads) void k_mom_b(const Entry<R>* __restrict__ tab, int n, int64_t total,
                                                    int src, int mom) {
  const int64_t k0 = int64_t(blockIdx.x) * kThreads;
  if (k0 >= total) return;
  const Entry<R>& E = entry_of(tab, n, blockIdx.x);
  mom_cell(E.a, slot(E, src), slot(E, mom), k0 - E.cell0 + threadIdx.x);
}

template <class R>
__global__ __launch_bounds__(kThreads) void k_bwd_b(const Entry<R>* __restrict__ tab, int n, int64_t total,
                                                    int mom, int lam1, int lam0, int with_obj) {
  const unsigned g = xcd_block(blockIdx.x, gridDim.x);
  const int64_t k0 = int64_t(g) * kThreads;
  if (k0 >= total) return;
  const Entry<R>& E = entry_of(tab, n, g);
  bwd_cell(E.a, slot(E, mom), slot(E, lam1), slot(E, lam0), E.grad, with_obj, k0 - E.cell0 + threadIdx.x);
}

}  // namespace adj
}  // namespace flb

using namespace flbh;
namespace A = flb::adj;

struct flb_adjoint_s {
  int device = 0;
  hipStream_t stream = nullptr;
  flb_adjoint_desc desc{};
  int64_t N = 0;
  size_t real = 4;
  std::vector<uint8_t> h_solid;
  bool any_solid = false;
  uint8_t* d_solid = nullptr;
  uint32_t* d_link = nullptr;  // link words (k_links) of d_solid
  void* d_s = nullptr;
  void* d_f0 = nullptr;
  void* d_work[2] = {nullptr, nullptr};
  void* d_lam[2] = {nullptr, nullptr};
  void* d_grad = nullptr;
  double* d_jt = nullptr;
  int jt_cap = 0;
  double* d_partial = nullptr;  // objective partial sums, [T + 1][chunks of 256 cells]
  int partial_cap = 0;          // in steps
  std::vector<void*> ckpt;  // ckpt[k] = state at step k*K (ckpt[0] aliases d_f0)
  // moment records (rho - 1, u; [4][N]) of the states the adjoint steps read:
  // those of one recomputed segment (K > 0), or of every state (K == 0)
  std::vector<void*> mtape;
  void* d_rec[2] = {nullptr, nullptr};  // recompute ping-pong (K > 0)
  int T = 0;
  int K = -1;
  bool recorded = false;
  void* last = nullptr;  // buffer holding the final forward state
  double J = 0.0;
  float fwd_ms = 0.0f, bwd_ms = 0.0f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

size_t state_bytes(const flb_adjoint_s* h) { return size_t(A::Q) * size_t(h->N) * h->real; }
size_t mom_bytes(const flb_adjoint_s* h) { return size_t(4) * size_t(h->N) * h->real; }

// Multiplier, shift and pass mask of A::div_magic for 1 <= d < 2^31.
void magic_of(uint32_t d, uint32_t& m, uint32_t& sh, uint32_t& pass) {
  pass = d <= 1 ? ~0u : 0u;
  if (d <= 1) {
    m = 0;
    sh = 0;
    return;
  }
  uint32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  m = uint32_t(((uint64_t(1) << (31 + l)) + d - 1) / d);
  sh = l - 1;
}

template <class R>
A::AdjArgs<R> adj_args(const flb_adjoint_s* h) {
  A::AdjArgs<R> a{};
  a.nx = h->desc.nx;
  a.ny = h->desc.ny;
  a.nz = h->desc.nz;
  a.px = h->desc.periodic[0] != 0;
  a.py = h->desc.periodic[1] != 0;
  a.pz = h->desc.periodic[2] != 0;
  a.N = h->N;
  a.solid = h->d_solid;
  a.link = h->d_link;
  a.s = static_cast<const R*>(h->d_s);
  a.omega = R(h->desc.omega);
  for (int d = 0; d < 3; ++d) a.G[d] = R(h->desc.force[d]);
  a.alpha = R(h->desc.brinkman_alpha);
  a.outlet = h->desc.outlet_x;
  a.invT = h->T > 0 ? R(1) / R(h->T) : R(0);
  a.om1 = R(1) - a.omega;
  a.fom = R(1) - a.omega * R(0.5);
  for (int d = 0; d < 3; ++d) a.Gh[d] = a.G[d] * R(0.5);
  // byte strides modulo 2^32 (a population array holds at most 2^32 bytes)
  const uint64_t e = sizeof(R);
  a.sy = uint32_t(e * uint64_t(a.nx));
  a.sz = uint32_t(e * uint64_t(a.nx) * uint64_t(a.ny));
  a.sn = uint32_t(e * uint64_t(h->N));
  magic_of(uint32_t(a.nx), a.mx, a.shx, a.px1);
  magic_of(uint32_t(a.ny), a.my, a.shy, a.py1);
  a.solids = h->any_solid ? 1 : 0;
  return a;
}

unsigned grid_of(const flb_adjoint_s* h) { return unsigned((h->N + A::kThreads - 1) / A::kThreads); }

// Link words of the current solid mask (on the object's stream).
hipError_t build_links(const flb_adjoint_s* h) {
  hipLaunchKernelGGL(A::k_links<float>, dim3(grid_of(h)), dim3(A::kThreads), 0, h->stream, adj_args<float>(h),
                     h->d_link);
  return hipGetLastError();
}

// One forward step src -> dst; OBJ: the chunk sums of its objective (step t);
// MOM: the moment record of src into mom.
template <class R, bool OBJ, bool MOM>
hipError_t fwd_step(const flb_adjoint_s* h, const void* src, void* dst, void* mom, int t) {
  hipLaunchKernelGGL((A::k_fwd<R, OBJ, MOM>), dim3(grid_of(h)), dim3(A::kThreads), 0, h->stream, adj_args<R>(h),
                     static_cast<const R*>(src), static_cast<R*>(dst), static_cast<R*>(mom),
                     h->d_partial + int64_t(t) * grid_of(h));
  return hipGetLastError();
}

template <class R>
hipError_t mom_step(const flb_adjoint_s* h, const void* src, void* mom) {
  hipLaunchKernelGGL(A::k_mom<R>, dim3(grid_of(h)), dim3(A::kThreads), 0, h->stream, adj_args<R>(h),
                     static_cast<const R*>(src), static_cast<R*>(mom));
  return hipGetLastError();
}

int alloc_state(flb_adjoint_s* h, void** p) {
  if (*p) return FLB_OK;
  if (hipMalloc(p, state_bytes(h)) != hipSuccess)
    return fail(FLB_ERR_NOMEM, "adjoint state allocation of %zu bytes failed", state_bytes(h));
  // solid cells are never written by the step kernels: keep them zero
  HIP_TRY(hipMemsetAsync(*p, 0, state_bytes(h), h->stream));
  return FLB_OK;
}

int alloc_mom(flb_adjoint_s* h, void** p) {
  if (*p) return FLB_OK;
  if (hipMalloc(p, mom_bytes(h)) != hipSuccess)
    return fail(FLB_ERR_NOMEM, "moment record allocation of %zu bytes failed", mom_bytes(h));
  return FLB_OK;
}

// Keeps the first n moment records of v (allocating what is missing) and frees the rest.
int size_moms(flb_adjoint_s* h, std::vector<void*>& v, int n) {
  for (size_t k = size_t(n); k < v.size(); ++k)
    if (v[k]) (void)hipFree(v[k]);
  v.resize(size_t(n), nullptr);
  int rc = FLB_OK;
  for (void*& p : v)
    if ((rc = alloc_mom(h, &p))) return rc;
  return FLB_OK;
}

void free_vec(std::vector<void*>& v, const void* keep) {
  for (void* p : v)
    if (p && p != keep) (void)hipFree(p);
  v.clear();
}

// Allocates the buffers a forward run of T steps with checkpoint interval K
// needs (kept across runs of the same shape) and records T and K.
int prepare_forward(flb_adjoint_s* h, int T, int K) {
  h->T = T;
  h->K = K;
  h->recorded = false;
  int rc = FLB_OK;
  for (void*& w : h->d_work)
    if ((rc = alloc_state(h, &w))) return rc;
  const int nck = K > 0 ? (T + K - 1) / K : 0;
  if (K > 0) {
    // keep existing checkpoint buffers when the count matches
    if (int(h->ckpt.size()) != nck) {
      free_vec(h->ckpt, h->d_f0);
      h->ckpt.assign(size_t(nck), nullptr);
      h->ckpt[0] = h->d_f0;
    }
    for (int k = 1; k < nck; ++k)
      if ((rc = alloc_state(h, &h->ckpt[size_t(k)]))) return rc;
    for (void*& r : h->d_rec)
      if ((rc = alloc_state(h, &r))) return rc;
    if ((rc = size_moms(h, h->mtape, std::min(K, T)))) return rc;
  } else if (K == 0) {
    if ((rc = size_moms(h, h->mtape, T))) return rc;
  }
  if (h->jt_cap < T + 1) {
    if (h->d_jt) (void)hipFree(h->d_jt);
    h->d_jt = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&h->d_jt), size_t(T + 1) * sizeof(double)) != hipSuccess)
      return fail(FLB_ERR_NOMEM, "objective buffer allocation failed");
    h->jt_cap = T + 1;
  }
  if (h->partial_cap < T + 1) {
    if (h->d_partial) (void)hipFree(h->d_partial);
    h->d_partial = nullptr;
    h->partial_cap = 0;
    if (hipMalloc(reinterpret_cast<void**>(&h->d_partial), size_t(T + 1) * grid_of(h) * sizeof(double)) !=
        hipSuccess)
      return fail(FLB_ERR_NOMEM, "objective partial-sum buffer allocation failed");
    h->partial_cap = T + 1;
  }
  return FLB_OK;
}

// Records the result of a forward run: the objective from the per-step plane
// sums jt[1..T] and the buffer holding the final state.
void finish_forward(flb_adjoint_s* h, const double* jt, void* last) {
  double s = 0.0;
  for (int t = 1; t <= h->T; ++t) s += jt[t];
  h->J = h->T > 0 ? s / double(h->T) : 0.0;
  h->last = last;
  h->recorded = h->K >= 0;
}

// Buffer receiving state t+1 of a forward run whose state t sits in src.
void* forward_target(const flb_adjoint_s* h, int t, const void* src) {
  const int K = h->K, T = h->T;
  if (K > 0 && (t + 1) % K == 0 && t + 1 < T) return h->ckpt[size_t((t + 1) / K)];
  return (src == h->d_work[0]) ? h->d_work[1] : h->d_work[0];
}

template <class R>
int forward_impl(flb_adjoint_s* h, int T, int K, double* J) {
  int rc = prepare_forward(h, T, K);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(h->e0, h->stream));
  const void* src = h->d_f0;
  for (int t = 0; t < T; ++t) {
    void* dst = forward_target(h, t, src);
    if (K == 0)
      HIP_TRY((fwd_step<R, true, true>(h, src, dst, h->mtape[size_t(t)], t + 1)));
    else
      HIP_TRY((fwd_step<R, true, false>(h, src, dst, nullptr, t + 1)));
    src = dst;
  }
  if (T > 0) {
    hipLaunchKernelGGL(A::k_obj_sum, dim3(unsigned(T)), dim3(A::kThreads), 0, h->stream, h->d_partial, grid_of(h),
                       h->d_jt);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(h->e1, h->stream));
  std::vector<double> jt(size_t(T + 1), 0.0);
  if (T > 0)
    HIP_TRY(hipMemcpyAsync(jt.data() + 1, h->d_jt + 1, size_t(T) * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipEventElapsedTime(&h->fwd_ms, h->e0, h->e1));
  finish_forward(h, jt.data(), const_cast<void*>(src));
  if (J) *J = h->J;
  return FLB_OK;
}

// One adjoint step; mom: the moment record of state t.
template <class R>
hipError_t bwd_step(const flb_adjoint_s* h, const void* mom, const void* lam1, void* lam0, int with_obj) {
  hipLaunchKernelGGL(A::k_bwd<R>, dim3(grid_of(h)), dim3(A::kThreads), 0, h->stream, adj_args<R>(h),
                     static_cast<const R*>(mom), static_cast<const R*>(lam1), static_cast<R*>(lam0),
                     static_cast<R*>(h->d_grad), with_obj);
  return hipGetLastError();
}

int prepare_backward(flb_adjoint_s* h) {
  int rc = FLB_OK;
  for (void*& l : h->d_lam)
    if ((rc = alloc_state(h, &l))) return rc;
  if (!h->d_grad) {
    if (hipMalloc(&h->d_grad, size_t(h->N) * h->real) != hipSuccess)
      return fail(FLB_ERR_NOMEM, "gradient allocation failed");
  }
  return FLB_OK;
}

template <class R>
int backward_impl(flb_adjoint_s* h, double* grad) {
  const int T = h->T, K = h->K;
  int rc = prepare_backward(h);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(h->e0, h->stream));
  HIP_TRY(hipMemsetAsync(h->d_grad, 0, size_t(h->N) * h->real, h->stream));
  hipLaunchKernelGGL(A::k_seed<R>, dim3(grid_of(h)), dim3(A::kThreads), 0, h->stream, adj_args<R>(h),
                     static_cast<R*>(h->d_lam[0]));
  HIP_TRY(hipGetLastError());
  int cur = 0;
  if (K == 0) {
    for (int t = T - 1; t >= 0; --t) {
      HIP_TRY(bwd_step<R>(h, h->mtape[size_t(t)], h->d_lam[cur], h->d_lam[cur ^ 1], t >= 1));
      cur ^= 1;
    }
  } else {
    const int nseg = (T + K - 1) / K;
    for (int k = nseg - 1; k >= 0; --k) {
      const int t0 = k * K, t1 = std::min((k + 1) * K, T);
      // moment records of states t0 .. t1-1: the recompute steps from the
      // checkpoint write those of their sources, k_mom that of the last state
      const void* src = h->ckpt[size_t(k)];
      for (int t = t0; t < t1 - 1; ++t) {
        void* dst = h->d_rec[(t - t0) & 1];
        HIP_TRY((fwd_step<R, false, true>(h, src, dst, h->mtape[size_t(t - t0)], 0)));
        src = dst;
      }
      HIP_TRY(mom_step<R>(h, src, h->mtape[size_t(t1 - 1 - t0)]));
      for (int t = t1 - 1; t >= t0; --t) {
        HIP_TRY(bwd_step<R>(h, h->mtape[size_t(t - t0)], h->d_lam[cur], h->d_lam[cur ^ 1], t >= 1));
        cur ^= 1;
      }
    }
  }
  HIP_TRY(hipEventRecord(h->e1, h->stream));
  std::vector<R> g(size_t(h->N));
  HIP_TRY(hipMemcpyAsync(g.data(), h->d_grad, size_t(h->N) * sizeof(R), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipEventElapsedTime(&h->bwd_ms, h->e0, h->e1));
  for (int64_t i = 0; i < h->N; ++i) grad[i] = double(g[size_t(i)]);
  return FLB_OK;
}

#define CHECK_ADJ(h) \
  if (!(h)) return fail(FLB_ERR_ARG, "null adjoint handle")

}  // namespace

// ---- many boxes stepped together -------------------------------------------
// Every member keeps its own fields and state buffers; the batch steps all of
// them with one launch per time step over the concatenated cells (each box
// starts at a multiple of 64 cells). All members share T and the checkpoint
// interval, so one slot schedule (which buffer holds which state) serves them
// all: slot 0 = f^0, 1-2 = forward ping-pong, 3-4 = adjoint ping-pong,
// 5-6 = recompute ping-pong, 7.. = checkpoints, then the moment records.
struct flb_adjoint_batch_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<flb_adjoint_s*> members;
  std::vector<int64_t> cell0;
  int64_t total = 0;
  int64_t cells = 0;  // sum of member cell counts
  size_t real = 4;
  void* d_tab = nullptr;
  size_t tab_cap = 0;
  void* d_slots = nullptr;
  size_t slots_cap = 0;
  int nslots = 0;
  int T = 0, K = -1;
  bool recorded = false;
  float fwd_ms = 0.0f, bwd_ms = 0.0f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

enum { kSlotF0 = 0, kSlotW0 = 1, kSlotW1 = 2, kSlotL0 = 3, kSlotL1 = 4, kSlotR0 = 5, kSlotR1 = 6, kSlotCk = 7 };

int tape_base(int T, int K) { return kSlotCk + (K > 0 ? (T + K - 1) / K : 0); }

// the slot receiving state t+1 when state t sits in slot src (forward_target
// in slot numbers)
int forward_slot(int T, int K, int t, int src) {
  if (K > 0 && (t + 1) % K == 0 && t + 1 < T) return kSlotCk + (t + 1) / K;
  return src == kSlotW0 ? kSlotW1 : kSlotW0;
}

unsigned batch_grid(const flb_adjoint_batch_s* b) {
  return unsigned((b->total + A::kThreads - 1) / A::kThreads);
}

// Uploads the entry table and every member's slot pointers.
template <class R>
int upload_table(flb_adjoint_batch_s* b) {
  const int T = b->T, K = b->K;
  const size_t n = b->members.size();
  const int tb = tape_base(T, K);
  // moment records of this schedule (prepare_forward sized every member's so)
  const int ntape = K > 0 ? std::min(K, T) : (K == 0 ? T : 0);
  const int nslots = tb + ntape;
  size_t bytes_slots = n * size_t(nslots) * sizeof(R*);
  // the entries, then the box of every workgroup (A::entry_of)
  const size_t nwg = size_t(batch_grid(b));
  size_t bytes_tab = n * sizeof(A::Entry<R>) + nwg * sizeof(int);
  if (bytes_slots > b->slots_cap) {
    if (b->d_slots) (void)hipFree(b->d_slots);
    b->d_slots = nullptr;
    b->slots_cap = 0;
    if (hipMalloc(&b->d_slots, bytes_slots) != hipSuccess) return fail(FLB_ERR_NOMEM, "slot table allocation failed");
    b->slots_cap = bytes_slots;
  }
  if (bytes_tab > b->tab_cap) {
    if (b->d_tab) (void)hipFree(b->d_tab);
    b->d_tab = nullptr;
    b->tab_cap = 0;
    if (hipMalloc(&b->d_tab, bytes_tab) != hipSuccess) return fail(FLB_ERR_NOMEM, "entry table allocation failed");
    b->tab_cap = bytes_tab;
  }
  std::vector<R*> slots(n * size_t(nslots), nullptr);
  std::vector<A::Entry<R>> tab(n);
  R** dslots = static_cast<R**>(b->d_slots);
  for (size_t m = 0; m < n; ++m) {
    flb_adjoint_s* h = b->members[m];
    R** sl = slots.data() + m * size_t(nslots);
    sl[kSlotF0] = static_cast<R*>(h->d_f0);
    sl[kSlotW0] = static_cast<R*>(h->d_work[0]);
    sl[kSlotW1] = static_cast<R*>(h->d_work[1]);
    sl[kSlotL0] = static_cast<R*>(h->d_lam[0]);
    sl[kSlotL1] = static_cast<R*>(h->d_lam[1]);
    sl[kSlotR0] = static_cast<R*>(h->d_rec[0]);
    sl[kSlotR1] = static_cast<R*>(h->d_rec[1]);
    for (int k = 0; k < tb - kSlotCk; ++k) sl[kSlotCk + k] = static_cast<R*>(h->ckpt[size_t(k)]);
    for (int k = 0; k < ntape; ++k) sl[tb + k] = static_cast<R*>(h->mtape[size_t(k)]);
    tab[m].a = adj_args<R>(h);
    tab[m].slots = dslots + m * size_t(nslots);
    tab[m].grad = static_cast<R*>(h->d_grad);
    tab[m].jt = h->d_jt;
    tab[m].partial = h->d_partial;
    tab[m].cell0 = b->cell0[m];
  }
  HIP_TRY(hipMemcpyAsync(b->d_slots, slots.data(), bytes_slots, hipMemcpyHostToDevice, b->stream));
  std::vector<int> wgbox(nwg, 0);
  for (size_t m = 0; m < n; ++m) {
    const int64_t w0 = b->cell0[m] / A::kThreads, w1 = (b->cell0[m] + b->members[m]->N + A::kThreads - 1) / A::kThreads;
    for (int64_t w = w0; w < w1; ++w) wgbox[size_t(w)] = int(m);
  }
  std::vector<uint8_t> blob(bytes_tab);
  std::memcpy(blob.data(), tab.data(), n * sizeof(A::Entry<R>));
  std::memcpy(blob.data() + n * sizeof(A::Entry<R>), wgbox.data(), nwg * sizeof(int));
  HIP_TRY(hipMemcpyAsync(b->d_tab, blob.data(), bytes_tab, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->nslots = nslots;
  return FLB_OK;
}

// One forward step of every box, slot src -> dst; OBJ: the objectives into jt[t];
// MOM: the moment records of src into slot mom.
template <class R, bool OBJ, bool MOM>
hipError_t fwd_batch(const flb_adjoint_batch_s* b, int src, int dst, int mom, int t) {
  hipLaunchKernelGGL((A::k_fwd_b<R, OBJ, MOM>), dim3(batch_grid(b)), dim3(A::kThreads), 0, b->stream,
                     static_cast<const A::Entry<R>*>(b->d_tab), int(b->members.size()), b->total, src, dst, mom, t);
  return hipGetLastError();
}

template <class R>
hipError_t mom_batch(const flb_adjoint_batch_s* b, int src, int mom) {
  hipLaunchKernelGGL(A::k_mom_b<R>, dim3(batch_grid(b)), dim3(A::kThreads), 0, b->stream,
                     static_cast<const A::Entry<R>*>(b->d_tab), int(b->members.size()), b->total, src, mom);
  return hipGetLastError();
}

template <class R>
hipError_t bwd_batch(const flb_adjoint_batch_s* b, int mom, int lam1, int lam0, int with_obj) {
  hipLaunchKernelGGL(A::k_bwd_b<R>, dim3(batch_grid(b)), dim3(A::kThreads), 0, b->stream,
                     static_cast<const A::Entry<R>*>(b->d_tab), int(b->members.size()), b->total, mom, lam1,
                     lam0, with_obj);
  return hipGetLastError();
}

template <class R>
int batch_forward_impl(flb_adjoint_batch_s* b, int T, int K, double* J) {
  b->recorded = false;
  int rc = FLB_OK;
  for (flb_adjoint_s* h : b->members) {
    if ((rc = prepare_forward(h, T, K))) return rc;
    if (K >= 0 && (rc = prepare_backward(h))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  b->T = T;
  b->K = K;
  if ((rc = upload_table<R>(b))) return rc;
  const unsigned nm = unsigned(b->members.size());
  HIP_TRY(hipEventRecord(b->e0, b->stream));
  int src = kSlotF0;
  const int tb = tape_base(T, K);
  for (int t = 0; t < T; ++t) {
    const int dst = forward_slot(T, K, t, src);
    if (K == 0)
      HIP_TRY((fwd_batch<R, true, true>(b, src, dst, tb + t, t + 1)));
    else
      HIP_TRY((fwd_batch<R, true, false>(b, src, dst, 0, t + 1)));
    src = dst;
  }
  if (T > 0) {
    hipLaunchKernelGGL(A::k_obj_sum_b<R>, dim3(unsigned(T), nm), dim3(A::kThreads), 0, b->stream,
                       static_cast<const A::Entry<R>*>(b->d_tab));
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(b->e1, b->stream));
  std::vector<double> jt(size_t(nm) * size_t(T + 1), 0.0);
  for (unsigned m = 0; m < nm; ++m)
    HIP_TRY(hipMemcpyAsync(jt.data() + size_t(m) * size_t(T + 1) + 1, b->members[m]->d_jt + 1,
                           size_t(T) * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipEventElapsedTime(&b->fwd_ms, b->e0, b->e1));
  for (unsigned m = 0; m < nm; ++m) {
    flb_adjoint_s* h = b->members[m];
    // the final state's buffer, by the same slot numbering as upload_table
    void* buf = nullptr;
    if (src == kSlotF0) buf = h->d_f0;
    else if (src == kSlotW0) buf = h->d_work[0];
    else if (src == kSlotW1) buf = h->d_work[1];
    else buf = h->ckpt[size_t(src - kSlotCk)];
    finish_forward(h, jt.data() + size_t(m) * size_t(T + 1), buf);
    h->fwd_ms = b->fwd_ms;
    if (J) J[m] = h->J;
  }
  b->recorded = K >= 0;
  return FLB_OK;
}

template <class R>
int batch_backward_impl(flb_adjoint_batch_s* b, double* grad) {
  const int T = b->T, K = b->K;
  const unsigned nm = unsigned(b->members.size());
  int rc = FLB_OK;
  for (flb_adjoint_s* h : b->members) {
    if (!h->recorded || h->T != T || h->K != K)
      return fail(FLB_ERR_STATE, "a member was run on its own since the batch forward run");
    if ((rc = prepare_backward(h))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  if ((rc = upload_table<R>(b))) return rc;
  const auto* tab = static_cast<const A::Entry<R>*>(b->d_tab);
  HIP_TRY(hipEventRecord(b->e0, b->stream));
  for (flb_adjoint_s* h : b->members)
    HIP_TRY(hipMemsetAsync(h->d_grad, 0, size_t(h->N) * h->real, b->stream));
  hipLaunchKernelGGL(A::k_seed_b<R>, dim3(batch_grid(b)), dim3(A::kThreads), 0, b->stream, tab, int(nm), b->total,
                     int(kSlotL0));
  HIP_TRY(hipGetLastError());
  const int tb = tape_base(T, K);
  int cur = 0;
  if (K == 0) {
    for (int t = T - 1; t >= 0; --t) {
      HIP_TRY(bwd_batch<R>(b, tb + t, kSlotL0 + cur, kSlotL0 + (cur ^ 1), t >= 1));
      cur ^= 1;
    }
  } else {
    const int nseg = (T + K - 1) / K;
    for (int k = nseg - 1; k >= 0; --k) {
      const int t0 = k * K, t1 = std::min((k + 1) * K, T);
      // as backward_impl: moment records of states t0 .. t1-1, then the adjoint steps
      int src = kSlotCk + k;
      for (int t = t0; t < t1 - 1; ++t) {
        const int dst = kSlotR0 + ((t - t0) & 1);
        HIP_TRY((fwd_batch<R, false, true>(b, src, dst, tb + (t - t0), 0)));
        src = dst;
      }
      HIP_TRY(mom_batch<R>(b, src, tb + (t1 - 1 - t0)));
      for (int t = t1 - 1; t >= t0; --t) {
        HIP_TRY(bwd_batch<R>(b, tb + (t - t0), kSlotL0 + cur, kSlotL0 + (cur ^ 1), t >= 1));
        cur ^= 1;
      }
    }
  }
  HIP_TRY(hipEventRecord(b->e1, b->stream));
  std::vector<R> g(size_t(b->cells));
  size_t off = 0;
  for (flb_adjoint_s* h : b->members) {
    HIP_TRY(hipMemcpyAsync(g.data() + off, h->d_grad, size_t(h->N) * sizeof(R), hipMemcpyDeviceToHost, b->stream));
    off += size_t(h->N);
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipEventElapsedTime(&b->bwd_ms, b->e0, b->e1));
  for (flb_adjoint_s* h : b->members) h->bwd_ms = b->bwd_ms;
  for (size_t i = 0; i < g.size(); ++i) grad[i] = double(g[i]);
  return FLB_OK;
}

}  // namespace

extern "C" {

int flb_adjoint_create(const flb_adjoint_desc* desc, flb_adjoint* out) {
  if (!desc || !out) return fail(FLB_ERR_ARG, "null descriptor or output");
  *out = nullptr;
  if (desc->storage != FLB_STORE_F32 && desc->storage != FLB_STORE_F64)
    return fail(FLB_ERR_UNSUPPORTED, "the adjoint runs in FP32 or FP64 storage, got kind %d", desc->storage);
  if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1)
    return fail(FLB_ERR_ARG, "extents must be positive (%d,%d,%d)", desc->nx, desc->ny, desc->nz);
  if (!(desc->omega > 0.0 && desc->omega < 2.0)) return fail(FLB_ERR_ARG, "omega must lie in (0, 2)");
  if (!(desc->brinkman_alpha >= 0.0)) return fail(FLB_ERR_ARG, "Brinkman coefficient must be >= 0");
  if (desc->outlet_x < 0 || desc->outlet_x >= desc->nx) return fail(FLB_ERR_ARG, "outlet plane outside the box");
  // The steps address a population array with 32-bit byte offsets.
  const int64_t real = desc->storage == FLB_STORE_F64 ? 8 : 4;
  const double cells = double(desc->nx) * double(desc->ny) * double(desc->nz);
  const int64_t N = cells * double(real) > 4294967296.0 ? int64_t(-1) : int64_t(desc->nx) * desc->ny * desc->nz;
  if (N < 0)
    return fail(FLB_ERR_ARG, "box too large: %.0f cells, at most 2^32 bytes per population array (%lld cells)",
                cells, static_cast<long long>((int64_t(1) << 32) / real));
  flb_adjoint_s* h = new flb_adjoint_s();
  h->desc = *desc;
  h->N = N;
  h->real = desc->storage == FLB_STORE_F64 ? 8 : 4;
  h->stream = static_cast<hipStream_t>(desc->stream);
  int rc = FLB_OK;
  if (hipGetDevice(&h->device) != hipSuccess) rc = fail(FLB_ERR_HIP, "no HIP device");
  if (rc == FLB_OK && (hipEventCreate(&h->e0) != hipSuccess || hipEventCreate(&h->e1) != hipSuccess))
    rc = fail(FLB_ERR_HIP, "event creation failed");
  if (rc == FLB_OK && (hipMalloc(reinterpret_cast<void**>(&h->d_solid), size_t(N)) != hipSuccess ||
                       hipMalloc(reinterpret_cast<void**>(&h->d_link), size_t(N) * sizeof(uint32_t)) != hipSuccess ||
                       hipMalloc(&h->d_s, size_t(N) * h->real) != hipSuccess))
    rc = fail(FLB_ERR_NOMEM, "adjoint field allocation failed");
  if (rc == FLB_OK) rc = alloc_state(h, &h->d_f0);
  if (rc == FLB_OK && (hipMemsetAsync(h->d_solid, 0, size_t(N), h->stream) != hipSuccess ||
                       hipMemsetAsync(h->d_s, 0, size_t(N) * h->real, h->stream) != hipSuccess ||
                       build_links(h) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess))
    rc = fail(FLB_ERR_HIP, "initial clear failed");
  if (rc != FLB_OK) {
    flb_adjoint_destroy(h);
    return rc;
  }
  *out = h;
  return FLB_OK;
}

int flb_adjoint_destroy(flb_adjoint h) {
  if (!h) return FLB_OK;
  DeviceGuard g(h->device);
  (void)hipStreamSynchronize(h->stream);
  free_vec(h->ckpt, h->d_f0);
  free_vec(h->mtape, nullptr);
  void* ptrs[] = {h->d_solid, h->d_link, h->d_s, h->d_f0, h->d_work[0], h->d_work[1], h->d_lam[0], h->d_lam[1],
                  h->d_rec[0], h->d_rec[1], h->d_grad, h->d_jt, h->d_partial};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  delete h;
  return FLB_OK;
}

int flb_adjoint_set_solid(flb_adjoint h, const uint8_t* solid) {
  CHECK_ADJ(h);
  if (!solid) return fail(FLB_ERR_ARG, "null solid mask");
  DeviceGuard g(h->device);
  std::vector<uint8_t> m(size_t(h->N));
  for (int64_t i = 0; i < h->N; ++i) m[size_t(i)] = solid[i] ? 1 : 0;
  h->h_solid = m;
  h->any_solid = std::find(m.begin(), m.end(), uint8_t(1)) != m.end();
  HIP_TRY(hipMemcpyAsync(h->d_solid, m.data(), size_t(h->N), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(build_links(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->recorded = false;
  return FLB_OK;
}

int flb_adjoint_set_fraction(flb_adjoint h, const double* s) {
  CHECK_ADJ(h);
  if (!s) return fail(FLB_ERR_ARG, "null solid fraction");
  DeviceGuard g(h->device);
  if (h->real == 8) {
    HIP_TRY(hipMemcpyAsync(h->d_s, s, size_t(h->N) * 8, hipMemcpyHostToDevice, h->stream));
  } else {
    std::vector<float> t(size_t(h->N));
    for (int64_t i = 0; i < h->N; ++i) t[size_t(i)] = float(s[i]);
    HIP_TRY(hipMemcpyAsync(h->d_s, t.data(), size_t(h->N) * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->recorded = false;
  return FLB_OK;
}

int flb_adjoint_set_fraction_at(flb_adjoint h, int64_t cell, double s) {
  CHECK_ADJ(h);
  if (cell < 0 || cell >= h->N) return fail(FLB_ERR_ARG, "cell %lld outside the box", (long long)cell);
  DeviceGuard g(h->device);
  char* dst = static_cast<char*>(h->d_s) + size_t(cell) * h->real;
  if (h->real == 8) {
    HIP_TRY(hipMemcpyAsync(dst, &s, 8, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  } else {
    const float v = float(s);
    HIP_TRY(hipMemcpyAsync(dst, &v, 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  h->recorded = false;
  return FLB_OK;
}

int flb_adjoint_init_equilibrium(flb_adjoint h, const double* rhou) {
  CHECK_ADJ(h);
  if (!rhou) return fail(FLB_ERR_ARG, "null rhou");
  DeviceGuard g(h->device);
  const size_t count = size_t(4) * size_t(h->N);
  void* stage = nullptr;
  if (hipMalloc(&stage, count * h->real) != hipSuccess) return fail(FLB_ERR_NOMEM, "staging allocation failed");
  hipError_t e;
  if (h->real == 8) {
    e = hipMemcpyAsync(stage, rhou, count * 8, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(A::k_init<double>, dim3(grid_of(h)), dim3(A::kThreads), 0, h->stream,
                         adj_args<double>(h), static_cast<const double*>(stage), static_cast<double*>(h->d_f0));
      e = hipGetLastError();
    }
  } else {
    std::vector<float> t(count);
    for (size_t i = 0; i < count; ++i) t[i] = float(rhou[i]);
    e = hipMemcpyAsync(stage, t.data(), count * 4, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(A::k_init<float>, dim3(grid_of(h)), dim3(A::kThreads), 0, h->stream,
                         adj_args<float>(h), static_cast<const float*>(stage), static_cast<float*>(h->d_f0));
      e = hipGetLastError();
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(stage);
  if (e != hipSuccess) return fail(FLB_ERR_HIP, "equilibrium initialisation failed: %s", hipGetErrorString(e));
  h->recorded = false;
  return FLB_OK;
}

int flb_adjoint_forward(flb_adjoint h, int steps, int checkpoint_every, double* J) {
  CHECK_ADJ(h);
  if (steps < 1) return fail(FLB_ERR_ARG, "forward needs at least one step");
  if (checkpoint_every < -1) return fail(FLB_ERR_ARG, "checkpoint_every must be -1, 0 or positive");
  DeviceGuard g(h->device);
  Range range("flb_adjoint_forward");
  return h->real == 8 ? forward_impl<double>(h, steps, checkpoint_every, J)
                      : forward_impl<float>(h, steps, checkpoint_every, J);
}

int flb_adjoint_backward(flb_adjoint h, double* grad) {
  CHECK_ADJ(h);
  if (!grad) return fail(FLB_ERR_ARG, "null gradient output");
  if (!h->recorded) return fail(FLB_ERR_STATE, "no recorded forward run (forward with checkpoint_every >= 0 first)");
  DeviceGuard g(h->device);
  Range range("flb_adjoint_backward");
  return h->real == 8 ? backward_impl<double>(h, grad) : backward_impl<float>(h, grad);
}

int flb_adjoint_state(flb_adjoint h, double* pops) {
  CHECK_ADJ(h);
  if (!pops) return fail(FLB_ERR_ARG, "null output");
  if (!h->last) return fail(FLB_ERR_STATE, "no forward run yet");
  DeviceGuard g(h->device);
  const size_t n = size_t(A::Q) * size_t(h->N);
  if (h->real == 8) {
    HIP_TRY(hipMemcpyAsync(pops, h->last, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  } else {
    std::vector<float> t(n);
    HIP_TRY(hipMemcpyAsync(t.data(), h->last, n * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; ++i) pops[i] = double(t[i]);
  }
  // stored g = f - w; solid cells hold no populations
  for (int i = 0; i < A::Q; ++i) {
    const double w = flb::Lat<A::Q>::template w<double>(i);
    double* p = pops + size_t(i) * size_t(h->N);
    for (int64_t c = 0; c < h->N; ++c)
      p[c] = (!h->h_solid.empty() && h->h_solid[size_t(c)]) ? 0.0 : p[c] + w;
  }
  return FLB_OK;
}

int flb_adjoint_timing(flb_adjoint h, double* forward_ms, double* backward_ms) {
  CHECK_ADJ(h);
  if (forward_ms) *forward_ms = h->fwd_ms;
  if (backward_ms) *backward_ms = h->bwd_ms;
  return FLB_OK;
}

int flb_adjoint_batch_create(const flb_adjoint* members, int n, void* stream, flb_adjoint_batch* out) {
  if (!members || !out) return fail(FLB_ERR_ARG, "null member list or output");
  *out = nullptr;
  if (n < 1) return fail(FLB_ERR_ARG, "a batch needs at least one member");
  for (int m = 0; m < n; ++m) {
    if (!members[m]) return fail(FLB_ERR_ARG, "member %d is null", m);
    for (int k = 0; k < m; ++k)
      if (members[k] == members[m]) return fail(FLB_ERR_ARG, "member %d appears twice", m);
    if (members[m]->device != members[0]->device) return fail(FLB_ERR_ARG, "members live on different devices");
    if (members[m]->real != members[0]->real) return fail(FLB_ERR_ARG, "members mix storage kinds");
  }
  flb_adjoint_batch_s* b = new flb_adjoint_batch_s();
  b->device = members[0]->device;
  b->stream = static_cast<hipStream_t>(stream);
  b->real = members[0]->real;
  b->members.assign(members, members + n);
  for (int m = 0; m < n; ++m) {
    b->cell0.push_back(b->total);
    b->total += (members[m]->N + A::kThreads - 1) / A::kThreads * A::kThreads;
    b->cells += members[m]->N;
  }
  DeviceGuard g(b->device);
  if (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) {
    flb_adjoint_batch_destroy(b);
    return fail(FLB_ERR_HIP, "event creation failed");
  }
  *out = b;
  return FLB_OK;
}

int flb_adjoint_batch_destroy(flb_adjoint_batch b) {
  if (!b) return FLB_OK;
  DeviceGuard g(b->device);
  (void)hipStreamSynchronize(b->stream);
  if (b->d_tab) (void)hipFree(b->d_tab);
  if (b->d_slots) (void)hipFree(b->d_slots);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  delete b;
  return FLB_OK;
}

int flb_adjoint_batch_forward(flb_adjoint_batch b, int steps, int checkpoint_every, double* J) {
  if (!b) return fail(FLB_ERR_ARG, "null batch handle");
  if (steps < 1) return fail(FLB_ERR_ARG, "forward needs at least one step");
  if (checkpoint_every < -1) return fail(FLB_ERR_ARG, "checkpoint_every must be -1, 0 or positive");
  DeviceGuard g(b->device);
  Range range("flb_adjoint_batch_forward");
  return b->real == 8 ? batch_forward_impl<double>(b, steps, checkpoint_every, J)
                      : batch_forward_impl<float>(b, steps, checkpoint_every, J);
}

int flb_adjoint_batch_backward(flb_adjoint_batch b, double* grad) {
  if (!b) return fail(FLB_ERR_ARG, "null batch handle");
  if (!grad) return fail(FLB_ERR_ARG, "null gradient output");
  if (!b->recorded) return fail(FLB_ERR_STATE, "no recorded batch forward run (checkpoint_every >= 0 first)");
  DeviceGuard g(b->device);
  Range range("flb_adjoint_batch_backward");
  return b->real == 8 ? batch_backward_impl<double>(b, grad) : batch_backward_impl<float>(b, grad);
}

int flb_adjoint_batch_timing(flb_adjoint_batch b, double* forward_ms, double* backward_ms) {
  if (!b) return fail(FLB_ERR_ARG, "null batch handle");
  if (forward_ms) *forward_ms = b->fwd_ms;
  if (backward_ms) *backward_ms = b->bwd_ms;
  return FLB_OK;
}

}  // extern "C"
