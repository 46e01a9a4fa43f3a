// Lattice kernels for gfx950: the fused AA-pattern collide-and-stream step
// and its read-out companions (templates; instantiated per storage kind in
// flb_inst_*.hip).
//
// AA (esoteric) in-place streaming, one array per direction (SoA):
//   even step  read  f_i <- A[i][x]              write A[opp i][x]   <- f*_i
//   odd step   read  f_i <- A[opp i][x - c_i]    write A[i][x + c_i] <- f*_i
// After an odd step the arrays hold the natural layout (A[i][x] = post-stream
// f_i(x)); after an even step A[opp i][x] holds the post-collision f*_i(x),
// i.e. the streamed state is implicit. Each cell update reads and writes its
// q populations exactly once: 2*q*sizeof(element) bytes per cell.
//
// Halfway bounce-back (FreeLB's bounceback::normal / movingwall, applied after
// Stream from getPrevious) needs no extra pass in this pattern. At an odd step
// a link whose upstream neighbour x - c_i is a wall reads the cell's own
// post-collision f*_{opp i}, which the even step left in A[i][x]; a link whose
// downstream neighbour x + c_i is a wall additionally writes f*_i into
// A[opp i][x], where the next even step reads f_{opp i}. The ordinary write of
// such a link lands in the wall cell's slot, which no halfway cell reads, and
// A[opp i][x] has no other writer, because its ordinary writer would be the
// (not updated) wall cell.
//
// Vector width. Each thread updates VEC consecutive cells of one row, so the
// even step moves whole VEC-element vectors per direction (rows start
// 256-byte aligned in the slot layout). At the odd step the directions with
// c_x = +-1 are one element off the vector grid: a thread loads its aligned
// vector and takes the missing element from the neighbouring lane with a
// cross-lane shuffle (only the first / last lane of a wavefront load it from
// memory), and writes are assembled the same way, so every direction still
// moves aligned vectors. A wavefront never spans two rows.
//
// Workgroups are remapped so that the workgroups the hardware places on one
// XCD (consecutive IDs go round-robin over the 8 XCDs) cover one contiguous
// slab of the block: the y/z neighbours an odd step reads then mostly sit in
// that XCD's own L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "flb_cell.hpp"
#include "flb_internal.hpp"
#include "flb_storage.hpp"

namespace flb {
namespace kern {

enum : int {
  DYN_NONE = 0,
  DYN_BGK = 1,
  DYN_BGK_FIELD = 2,
  DYN_BB = 3,
  DYN_BB_MW = 4,
  DYN_MRT = 5,
  DYN_SMAG_BGK = 6,
  DYN_SMAG_MRT = 7,
  DYN_BGK_FORCE = 8,
  DYN_BGK_WRITE = 9,
  DYN_BGK_SCALAR_FORCE = 10,
  DYN_SCALAR_BGK = 11,
  DYN_SCALAR_BGK_FIELD = 12
};
// HWBC_ANTI (scalar lattices) shares the second link word with HWBC_MOVING
// (flow lattices): a block has one or the other.
enum : int { HWBC_NORMAL = 0, HWBC_MOVING = 1, HWBC_ANTI = 2 };

template <class S>
__device__ __forceinline__ typename S::E* arr(const KArgs& a, int i) {
  return static_cast<typename S::E*>(a.pops) + int64_t(i) * a.dpop;
}

// Row / plane strides of the population slots (POP) or of the per-cell
// arrays -- flags and fields (!POP).
template <bool POP>
__device__ __forceinline__ int ystr(const KArgs& a) {
  return POP ? a.ypop : a.pitch;
}
template <bool POP>
__device__ __forceinline__ int64_t zstr(const KArgs& a) {
  return POP ? a.zpop : a.plane;
}

// Slot offset of the y/z part of c_i for a cell in row (y, z), wrapping
// periodic axes (uniform over a wavefront, which covers one row). 64-bit:
// the wrap across a periodic z axis spans the whole allocation, which holds
// more than 2^31 populations in blocks of 113 M cells and up.
template <int Q, bool PER, bool POP>
__device__ __forceinline__ int64_t yz_off(const KArgs& a, int i, int y, int z) {
  const int cy = Lat<Q>::cy(i), cz = Lat<Q>::cz(i);
  int64_t o = int64_t(cy * ystr<POP>(a)) + cz * zstr<POP>(a);
  if (PER) {
    if (a.py && cy != 0) {
      const int yn = y + cy;
      if (yn < 0) o += int64_t(a.ny) * ystr<POP>(a);
      else if (yn >= a.ny) o -= int64_t(a.ny) * ystr<POP>(a);
    }
    if (Lat<Q>::D == 3 && a.pz && cz != 0) {
      const int zn = z + cz;
      if (zn < 0) o += a.nz * zstr<POP>(a);
      else if (zn >= a.nz) o -= a.nz * zstr<POP>(a);
    }
  }
  return o;
}

// Full slot offset of the neighbour x + c_i.
template <int Q, bool PER, bool POP>
__device__ __forceinline__ int64_t nb_off(const KArgs& a, int i, int x, int y, int z) {
  const int cx = Lat<Q>::cx(i);
  int64_t o = cx + yz_off<Q, PER, POP>(a, i, y, z);
  if (PER && a.px && cx != 0) {
    const int xn = x + cx;
    if (xn < 0) o += a.nx;
    else if (xn >= a.nx) o -= a.nx;
  }
  return o;
}

// First task of the table matching the flag (TaskSelector, utils/tmp.h:283-309).
__device__ __forceinline__ int match_task(const KArgs& a, uint8_t fl, float& cs) {
  for (int t = 0; t < a.ntasks; ++t) {
    if (fl & a.tmask[t]) {
      cs = a.tcs[t];
      return a.tdyn[t];
    }
  }
  return DYN_NONE;
}

// Halfway wall links of a cell (bit k: neighbour x + c_k is a wall), split by
// kind; managers apply in order so a later match overrides an earlier one
// (BlockFixedBoundary::addtoBd, basic_boundary.h:198-212).
template <int Q, bool PER>
__device__ __forceinline__ void wall_links(const KArgs& a, int x, int y, int z, int slot,
                                           uint8_t fl, unsigned& ln, unsigned& lm) {
  ln = 0u;
  lm = 0u;
  for (int k = 0; k < a.nbc; ++k) {
    if (!(fl & a.bcmask[k])) continue;
    const uint8_t vm = a.bcvoid[k];
    unsigned v = 0u;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if (a.flags[slot + nb_off<Q, PER, false>(a, i, x, y, z)] & vm) v |= 1u << i;
    }
    if (a.bckind[k] != HWBC_NORMAL) {
      lm |= v;
      ln &= ~v;
    } else {
      ln |= v;
      lm &= ~v;
    }
  }
}

template <int Q, class R>
__device__ __forceinline__ void load_fields(const KArgs& a, int slot, R& rho_f, R (&u_f)[3]) {
  rho_f = a.rho_f ? static_cast<const R*>(a.rho_f)[slot] : R(1);
  const R* u = static_cast<const R*>(a.u_f);
#pragma unroll
  for (int d = 0; d < 3; ++d) u_f[d] = (u && d < Lat<Q>::D) ? u[d * a.fstride + slot] : R(0);
}

// moment::rhoU<CELL, true> (moment.h:359-377): the moments go to the RHO /
// VELOCITY fields of the cell.
template <int Q, class R>
__device__ __forceinline__ void store_fields(const KArgs& a, int slot, R rho, const R (&u)[3]) {
  if (a.rho_f) static_cast<R*>(const_cast<void*>(a.rho_f))[slot] = rho;
  R* uf = static_cast<R*>(const_cast<void*>(a.u_f));
  if (uf) {
#pragma unroll
    for (int d = 0; d < Lat<Q>::D; ++d) uf[d * a.fstride + slot] = u[d];
  }
}

// f_i + w_i and back: the shifted kinds' deviations as populations.
template <int Q, class R>
__device__ __forceinline__ void add_w(R (&f)[Q]) {
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] += Lat<Q>::template w<R>(i);
}
template <int Q, class R>
__device__ __forceinline__ void sub_w(R (&f)[Q]) {
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] -= Lat<Q>::template w<R>(i);
}

// The dynamics of the FAM_FULL family beyond BGK and bounce-back (forced BGK,
// BGK writing its moments, Smagorinsky BGK, MRT and MRT+LES). DEV (D3Q19
// only): f holds the deviations f - w; BGK, the Smagorinsky omega and MRT
// then run on them, the forced dynamics on f = g + w.
template <int Q, class R, bool DEV>
__device__ __forceinline__ void collide_full(const KArgs& a, int slot, int dyn, float cs, R (&f)[Q],
                                             R om, R _om) {
  static_assert(!DEV || Q == 19, "deviation form of the full family: D3Q19");
  R rho, u[3], rho_f, u_f[3];
  if (dyn == DYN_BGK_FORCE || dyn == DYN_BGK_SCALAR_FORCE) {
    if constexpr (DEV) add_w<Q>(f);
    if (dyn == DYN_BGK_FORCE) {
      const R F[3] = {R(a.force[0]), R(a.force[1]), R(a.force[2])};
      bgk_force<Q>(f, F, om, _om);
    } else {
      // BGKForce<forcerhoU<CELL, ScalarForce, true>, SecondOrder, ScalarForce>
      // with SCALARFORCE = 0 + the force::Buoyancy terms (force.h:160-181)
      // of the coupled scalar blocks, in coupling order
      R F = R(0);
      const uint8_t fl = a.flags[slot];
#pragma unroll
      for (int k = 0; k < kMaxCouplings; ++k) {
        if (k < a.ncpl && (fl & a.cpl_mask[k]))
          F += R(a.cpl_gbeta[k]) * (static_cast<const R*>(a.cpl_rho[k])[slot] - R(a.cpl_rho0[k]));
      }
      bgk_scalar_force<Q>(f, F, om, _om, rho, u);
      store_fields<Q>(a, slot, rho, u);
    }
    if constexpr (DEV) sub_w<Q>(f);
    return;
  }
  if constexpr (DEV) {
    // MRT and MRT + LES from the moments of the deviations (mrt_dev)
    if (dyn == DYN_MRT || dyn == DYN_SMAG_MRT) {
      const R* sn;
      if constexpr (sizeof(R) == 8) sn = a.mrt_sn_d;
      else sn = a.mrt_sn_f;
      rho_f = R(1);
      if (dyn == DYN_SMAG_MRT) load_fields<Q>(a, slot, rho_f, u_f);
      mrt_dev(f, om, dyn == DYN_SMAG_MRT, R(cs), rho_f, sn, a.mrt_shear);
      return;
    }
  }
  R d = R(0);
  if constexpr (DEV) rhou_dev<Q>(f, d, rho, u);
  else rhou(f, rho, u);
  if (dyn == DYN_BGK_WRITE) {
    store_fields<Q>(a, slot, rho, u);
    if constexpr (DEV) bgk_dev<Q>(f, d, rho, u, om, _om);
    else bgk_with<Q>(f, rho, u, om, _om);
    return;
  }
  R eo = om;
  if (dyn == DYN_SMAG_BGK || dyn == DYN_SMAG_MRT) {
    load_fields<Q>(a, slot, rho_f, u_f);
    if constexpr (Q == 19) eo = smag_omega<DEV>(f, rho, u, om, R(cs), rho_f, d);
    else eo = smag_omega(f, rho, u, om, R(cs), rho_f);
  }
  if (dyn == DYN_SMAG_BGK) {
    if constexpr (DEV) bgk_dev<Q>(f, d, rho, u, eo, R(1) - eo);
    else bgk_with<Q>(f, rho, u, eo, R(1) - eo);
  } else {
    if constexpr (Q == 19) {
      const R* sn;
      if constexpr (sizeof(R) == 8) sn = a.mrt_sn_d;
      else sn = a.mrt_sn_f;
      mrt<DEV>(f, d, rho, u, eo, sn, a.mrt_shear);
    } else {
      mrt(f, rho, u, eo, a.rates);
    }
  }
}

// One cell's collision. The RHO / VELOCITY fields of the cell (slot) are read
// here, by the dynamics that use them, so they hold no registers elsewhere.
// DEV: f holds the deviations g_i = f_i - w_i (shifted storage kinds on the
// flow lattices). BGK then runs on g (bgk_dev), bounce-back is the same on
// g as on f (w_i = w_opp(i)); collide_full takes the rest.
template <int Q, class R, int FAM, bool DEV = false>
__device__ __forceinline__ void collide(const KArgs& a, int slot, int dyn, float cs, R (&f)[Q],
                                        R om, R _om) {
  if constexpr (FAM == FAM_SCALAR) {
    // Advection-diffusion (collision::BGK, collision.h:39-56): the scalar is
    // advected by the VELOCITY field (moment::useFieldU, moment.h:56-69);
    // DYN_SCALAR_BGK sums it and writes it to the field
    // (MomentaTuple<rho<CELL, true>, useFieldU>, moment.h:159-182, 667-679),
    // DYN_SCALAR_BGK_FIELD takes it from the field (useFieldrhoU, moment.h:71-90).
    R rho, u[3] = {R(0), R(0), R(0)};
    const R* ue = static_cast<const R*>(a.u_ext ? a.u_ext : a.u_f);
    if (ue) {
#pragma unroll
      for (int d = 0; d < Lat<Q>::D; ++d) u[d] = ue[d * a.fstride + slot];
    }
    if (dyn == DYN_SCALAR_BGK) {
      rho = rho_sum(f);
      if (a.rho_f) static_cast<R*>(const_cast<void*>(a.rho_f))[slot] = rho;
    } else if (dyn == DYN_SCALAR_BGK_FIELD) {
      rho = a.rho_f ? static_cast<const R*>(a.rho_f)[slot] : R(1);
    } else {
      return;
    }
    bgk_with<Q>(f, rho, u, om, _om);
    return;
  } else {
  R rho, u[3], rho_f, u_f[3];
  switch (dyn) {
    case DYN_BGK:
      if constexpr (DEV) {
        R d;
        rhou_dev<Q>(f, d, rho, u);
        bgk_dev<Q>(f, d, rho, u, om, _om);
      } else {
        rhou(f, rho, u);
        bgk_with<Q>(f, rho, u, om, _om);
      }
      break;
    case DYN_BGK_FIELD:
      load_fields<Q>(a, slot, rho_f, u_f);
      if constexpr (DEV) bgk_dev<Q>(f, rho_f - R(1), rho_f, u_f, om, _om);
      else bgk_with<Q>(f, rho_f, u_f, om, _om);
      break;
    case DYN_BB:
      bounceback<Q>(f);
      break;
    case DYN_BB_MW:
      load_fields<Q>(a, slot, rho_f, u_f);
      bounceback_moving<Q>(f, rho_f, u_f);
      break;
    default:
      if constexpr (FAM == FAM_FULL) {
        // D2Q9's full family (FreeLB's closed-form MRT) runs on f
        constexpr bool dev19 = DEV && Q == 19;
        if constexpr (DEV && !dev19) add_w<Q>(f);
        collide_full<Q, R, dev19>(a, slot, dyn, cs, f, om, _om);
        if constexpr (DEV && !dev19) sub_w<Q>(f);
      }
      break;
  }
  }
}

// Whole-wavefront lane shifts (DPP wave_shr:1 / wave_shl:1): lane l receives
// lane l-1's (shr) or lane l+1's (shl) value; the lane with no source gets 0.
__device__ __forceinline__ float wave_shr(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float wave_shl(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ double wave_shr(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(b)), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(uint64_t(b) >> 32)), 0x138, 0xf, 0xf, false);
  return __longlong_as_double(int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo)));
}
__device__ __forceinline__ double wave_shl(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(b)), 0x130, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(uint64_t(b) >> 32)), 0x130, 0xf, 0xf, false);
  return __longlong_as_double(int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo)));
}
__device__ __forceinline__ unsigned wave_shr(unsigned v) {
  return unsigned(__builtin_amdgcn_update_dpp(0, int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ unsigned wave_shl(unsigned v) {
  return unsigned(__builtin_amdgcn_update_dpp(0, int(v), 0x130, 0xf, 0xf, false));
}

template <int VEC>
__device__ __forceinline__ void load_flags(const uint8_t* p, uint8_t (&fl)[VEC]) {
  if constexpr (VEC == 1) {
    fl[0] = p[0];
  } else if constexpr (VEC == 2) {
    const uint16_t w = *reinterpret_cast<const uint16_t*>(p);
    fl[0] = uint8_t(w);
    fl[1] = uint8_t(w >> 8);
  } else {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(p);
#pragma unroll
    for (int j = 0; j < VEC; ++j) fl[j] = uint8_t(w >> (8 * j));
  }
}

// The step kernels compute on the deviations g_i = f_i - w_i when the
// storage holds them (the shifted kinds) on the flow lattices: decoding skips
// the add of w_i and encoding its subtraction, and collide<..., true> runs
// BGK on g. Every step kernel and launch family takes the same path, so a
// block's result does not depend on which kernel stepped it.
template <int Q, class S>
constexpr bool kDev = S::kShifted && (Q == 9 || Q == 19);
template <int Q, class S>
__device__ __forceinline__ typename S::R pop_dec(typename S::E e, int i) {
  if constexpr (kDev<Q, S>) return S::dev(e);
  else return S::template from<Q>(e, i);
}
template <int Q, class S>
__device__ __forceinline__ typename S::E pop_enc(typename S::R v, int i) {
  if constexpr (kDev<Q, S>) return S::undev(v);
  else return S::template to<Q>(v, i);
}

// VEC elements of one direction at p (aligned), converted to the compute type.
template <int Q, class S, int VEC>
__device__ __forceinline__ void load_row(const typename S::E* p, int i, typename S::R (&v)[VEC]) {
  if constexpr (VEC == 1) {
    v[0] = pop_dec<Q, S>(p[0], i);
  } else {
    const auto w = load_vec<typename S::E, VEC>(p);
#pragma unroll
    for (int j = 0; j < VEC; ++j) v[j] = pop_dec<Q, S>(w[j], i);
  }
}

// Store the elements of v whose bit is set in `mask` (full mask: one vector store).
template <int Q, class S, int VEC>
__device__ __forceinline__ void store_row(typename S::E* p, int i, const typename S::R (&v)[VEC],
                                          unsigned mask) {
  using E = typename S::E;
  constexpr unsigned full = (1u << VEC) - 1u;
  if constexpr (VEC == 1) {
    if (mask & 1u) p[0] = pop_enc<Q, S>(v[0], i);
  } else {
    if (mask == full) {
      vec_t<E, VEC> w;
#pragma unroll
      for (int j = 0; j < VEC; ++j) w[j] = pop_enc<Q, S>(v[j], i);
      store_vec<E, VEC>(p, w);
    } else if (mask) {
#pragma unroll
      for (int j = 0; j < VEC; ++j)
        if ((mask >> j) & 1u) p[j] = pop_enc<Q, S>(v[j], i);
    }
  }
}

// Population loads and stores of the step kernels. Each population is read
// and written once per step; at the even step of a block with wide rows no
// other wavefront touches a cell's lines, so non-temporal accesses
// (FLB_NT_EVEN, PAR 0) keep them from displacing lines in the caches. Blocks
// with compact rows share lines between neighbouring rows and wavefronts, and
// batches mix blocks: their even step (PAR 2) keeps the lines cached. The odd
// step re-reads neighbour lines from L2 and keeps its loads cached
// (FLB_NT_ODD_STORE marks only its stores).
#ifndef FLB_NT_EVEN
#define FLB_NT_EVEN 1
#endif
// In blocks with wide rows, the odd step's directions with c_x = 0 read and
// write whole aligned rows that no other wavefront touches: non-temporal
// there too (FLB_NT_ODD_ALIGNED). Its x-shifted directions share their edge
// lines with the neighbouring wavefront and stay cached.
#ifndef FLB_NT_ODD_ALIGNED
#define FLB_NT_ODD_ALIGNED 1
#endif
#ifndef FLB_NT_ODD_STORE
#define FLB_NT_ODD_STORE 0
#endif
template <int PAR, class E>
__device__ __forceinline__ E ld_pop(const E* p) {
  if constexpr (PAR == 0 && FLB_NT_EVEN) return __builtin_nontemporal_load(p);
  else return *p;
}
template <int PAR, class E>
__device__ __forceinline__ void st_pop(E* p, E v) {
  if constexpr ((PAR == 0 && FLB_NT_EVEN) || (PAR == 1 && FLB_NT_ODD_STORE)) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Buffer-addressed population access: a descriptor in scalar registers whose
// base is a wavefront-uniform element e0 (index from the allocation start),
// a 32-bit byte offset per lane and a wavefront-uniform scalar offset per
// direction, so an address costs no 64-bit pointer arithmetic. The range runs
// from e0 to the end of the allocation, capped at 4 GiB - 1 bytes; the
// offsets the kernels form stay below 2 GiB (KArgs::buf, checked on the host
// for the block's strides), so allocations of any size take this form.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pop_rsrc(const KArgs& a, size_t elem, int64_t e0) {
  const uint64_t off = uint64_t(e0) * elem;
  const uint64_t rest = a.pop_total - off;
  const uint32_t nrec = rest < 0xFFFFFFFFull ? uint32_t(rest) : 0xFFFFFFFFu;
  // word 3 of a raw (unformatted) buffer descriptor on the gfx9 family
  return __builtin_amdgcn_make_buffer_rsrc(static_cast<char*>(a.pops) - kFrontPad * elem + off, 0, int(nrec),
                                           0x00020000);
}

template <int Q, class S, int PAR, int FAM, bool PER, bool BUF = false>
__device__ __forceinline__ void cell_update_at(const KArgs& a, int c0, int y, int z);

// One population at lane byte offset vo plus wavefront-uniform scalar byte
// offset so, and its store.
template <int PAR, class E>
__device__ __forceinline__ E bld_pop_s(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  constexpr int aux = (PAR == 0 && FLB_NT_EVEN) ? 2 : 0;
  if constexpr (sizeof(E) == 8) {
    const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, aux);
    return __longlong_as_double(int64_t((uint64_t(w[1]) << 32) | uint64_t(w[0])));
  } else if constexpr (sizeof(E) == 4) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, vo, so, aux));
  } else {
    return E(__builtin_amdgcn_raw_buffer_load_b16(r, vo, so, aux));
  }
}
template <int PAR, class E>
__device__ __forceinline__ void bst_pop_s(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, E v) {
  constexpr int aux = ((PAR == 0 && FLB_NT_EVEN) || (PAR == 1 && FLB_NT_ODD_STORE)) ? 2 : 0;
  if constexpr (sizeof(E) == 8) {
    typedef unsigned int u32x2 __attribute__((__vector_size__(8)));
    const uint64_t b = uint64_t(__double_as_longlong(v));
    const u32x2 w = {uint32_t(b), uint32_t(b >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(w, r, vo, so, aux);
  } else if constexpr (sizeof(E) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, vo, so, aux);
  } else {
    __builtin_amdgcn_raw_buffer_store_b16(uint16_t(v), r, vo, so, aux);
  }
}

// Element offset of the neighbour x + c_k from x (no periodic wrap).
template <int Q>
__device__ __forceinline__ int32_t nb_rel(int k, int32_t yp, int32_t zp) {
  return Lat<Q>::cx(k) + Lat<Q>::cy(k) * yp + Lat<Q>::cz(k) * zp;
}

// Halfway wall terms of the odd step (moving walls; anti-bounce-back on the
// scalar lattices), each applied as one branch around the whole block so that
// wavefronts without such a cell skip it. In: the population f_i that came
// back across moving-wall link opp(i) picks up the term before the
// collision. Out: the population f_i leaving across moving-wall link i picks
// it up before its store.
template <int Q, class R>
__device__ __forceinline__ void wall_terms_in(const KArgs& a, int s0, unsigned lm, R (&f)[Q]) {
  if (!lm) return;
  R rho_f, u_f[3];
  load_fields<Q>(a, s0, rho_f, u_f);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if ((lm >> opp(i)) & 1u) {
      if constexpr (kScalarLat<Q>) f[i] = anti_term<Q>(i, rho_f, f[i]);
      else f[i] = f[i] + movingwall_term<Q>(i, rho_f, u_f);
    }
  }
}
template <int Q, class R>
__device__ __forceinline__ void wall_terms_out(const KArgs& a, int s0, unsigned lm, R (&f)[Q]) {
  if (!lm) return;
  R rho_f, u_f[3];
  load_fields<Q>(a, s0, rho_f, u_f);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if ((lm >> i) & 1u) {
      if constexpr (kScalarLat<Q>) f[i] = anti_term<Q>(opp(i), rho_f, f[i]);
      else f[i] = f[i] + movingwall_term<Q>(opp(i), rho_f, u_f);
    }
  }
}

// Per-lane byte delta of an odd-step access: of the two candidate slots rn
// (the neighbour's) and rw (the cell's own, taken when bit k of the link
// word is set), the smaller is in the scalar offset; the lane adds the
// difference when it takes the larger one. The link bit becomes a sign mask
// (0 or all ones) ANDed with the scalar difference: no compare and select.
// lkx is the link word XOR the wavefront-uniform bits of the directions
// whose wall candidate is the smaller one, so bit k set = larger candidate.
template <uint32_t el>
__device__ __forceinline__ uint32_t lane_delta(unsigned lkx, int k, int32_t rn, int32_t rw, bool linked) {
  if (!linked) return 0u;
  const uint32_t d = uint32_t(rw > rn ? rw - rn : rn - rw) * el;
  const uint32_t mask = uint32_t(int32_t(lkx << (31 - k)) >> 31);
  return mask & d;
}

// The odd step through the buffer descriptor (no periodic axis). The
// descriptor starts B = zpop + ypop + 1 (2D: ypop + 1) elements before the
// wavefront's first row, B bounding the most negative neighbour offset (rows
// y, z >= 1 keep it inside the allocation). Every access is the lane's base
// prel + c0 (its row relative to the first) plus a per-direction part: a
// wavefront-uniform scalar offset m + B, where m is the smaller of the two
// candidate slots (the neighbour's, or the cell's own across a wall link),
// and a per-lane delta (candidate - m >= 0) chosen by the lane's link bit, so
// every term is non-negative. Same selection and
// arithmetic as the general path, bit for bit. NTA: the c_x = 0 directions
// (whole aligned rows of wide blocks) load and store non-temporally.
template <int Q, class S, int FAM, bool NTA>
__device__ __forceinline__ void odd_buf(const KArgs& a, __amdgpu_buffer_rsrc_t rs, uint32_t prel, int c0, int s0,
                                        int dyn, float cs, unsigned ln, unsigned lm) {
  using R = typename S::R;
  using E = typename S::E;
  constexpr uint32_t el = sizeof(E);
  const int32_t dp = int32_t(a.dpop), yp = a.ypop, zp = int32_t(a.zpop);
  const int32_t B = (Lat<Q>::D == 3 ? zp : 0) + yp + 1;
  const uint32_t vob = (prel + uint32_t(c0)) * el;
  const unsigned lk = ln | lm;
  unsigned flip = 0;  // bit k: the wall candidate of direction k is the smaller slot
#pragma unroll
  for (int k = 1; k < Q; ++k)
    if (int32_t(opp(k) * dp) < int32_t(k * dp + nb_rel<Q>(k, yp, zp))) flip |= 1u << k;
  const unsigned lkx = lk ^ __builtin_amdgcn_readfirstlane(flip);
  R f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    // upstream neighbour x - c_i a wall: own f*_{opp i}, left in A[i][x]
    const int o = opp(i);
    const int32_t rn = o * dp + nb_rel<Q>(o, yp, zp), rw = i * dp;
    const int32_t m = rn < rw ? rn : rw;
    const uint32_t v = vob + lane_delta<el>(lkx, o, rn, rw, i != 0);
    if (NTA && Lat<Q>::cx(i) == 0)
      f[i] = pop_dec<Q, S>(bld_pop_s<0, E>(rs, v, uint32_t(m + B) * el), i);
    else
      f[i] = pop_dec<Q, S>(bld_pop_s<1, E>(rs, v, uint32_t(m + B) * el), i);
  }
  wall_terms_in<Q>(a, s0, lm, f);
  const R om = R(a.omega);
  collide<Q, R, FAM, kDev<Q, S>>(a, s0, dyn, cs, f, om, R(1) - om);
  wall_terms_out<Q>(a, s0, lm, f);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    // downstream neighbour x + c_i a wall: f_{opp i}(x, t+1) = f*_i (+ wall term)
    const int32_t rn = i * dp + nb_rel<Q>(i, yp, zp), rw = opp(i) * dp;
    const int32_t m = rn < rw ? rn : rw;
    const uint32_t w = vob + lane_delta<el>(lkx, i, rn, rw, i != 0);
    if (NTA && Lat<Q>::cx(i) == 0)
      bst_pop_s<0, E>(rs, w, uint32_t(m + B) * el, pop_enc<Q, S>(f[i], i));
    else
      bst_pop_s<1, E>(rs, w, uint32_t(m + B) * el, pop_enc<Q, S>(f[i], i));
  }
}

// Wavefront gw of the block (row gw / wpr, x segment gw % wpr) updates its
// up to 64 cells; lane is the lane id. With vec == -3 (linear form) wavefronts run over the
// block's updated cells in x-fastest order instead.
template <int Q, class S, int PAR, int FAM, bool PER, bool BUF = false>
__device__ __forceinline__ void cell_update(const KArgs& a, int gw, int lane) {
  int c0, y, z;
  if (a.vec == -3) {
    // linear form: wavefront gw takes updated cells 64 gw .. 64 gw + 63 in
    // x-fastest order, across row ends (short rows fill whole wavefronts)
    const unsigned c = unsigned(gw) * unsigned(kWave) + unsigned(lane);
    const unsigned exy = unsigned(a.ex) * unsigned(a.ey);
    if (c >= exy * unsigned(a.ez)) return;
    // c / exy and r / ex through the FP64 reciprocal (exact for c < 2^31 after
    // one correction), far cheaper than the integer division sequence
    unsigned zz = unsigned(double(c) * (1.0 / double(exy)));
    int r = int(c - zz * exy);
    if (r < 0) {
      --zz;
      r += int(exy);
    } else if (r >= int(exy)) {
      ++zz;
      r -= int(exy);
    }
    int yy = int(double(r) * (1.0 / double(a.ex)));
    c0 = r - yy * a.ex;
    if (c0 < 0) {
      --yy;
      c0 += a.ex;
    } else if (c0 >= a.ex) {
      ++yy;
      c0 -= a.ex;
    }
    y = a.ys + yy;
    z = a.zs + int(zz);
  } else {
    const int row = gw / a.wpr;
    if (row >= a.ey * a.ez) return;  // the whole wavefront leaves together
    const int wv = gw - row * a.wpr;
    y = a.ys + row % a.ey;
    z = a.zs + row / a.ey;
    c0 = wv * kWave + lane;
    if (c0 >= a.ex) return;
  }
  cell_update_at<Q, S, PAR, FAM, PER, BUF>(a, c0, y, z);
}

// The update of the cell at x = xs + c0 in row (y, z).
template <int Q, class S, int PAR, int FAM, bool PER, bool BUF>
__device__ __forceinline__ void cell_update_at(const KArgs& a, int c0, int y, int z) {
  using R = typename S::R;
  using E = typename S::E;
  constexpr bool EVEN = PAR != 1;  // PAR 0: even step with non-temporal accesses, 2: without
  const int x = a.xs + c0;
  const int s0 = c0 + a.pitch * y + int(a.plane) * z;
  E* const P = static_cast<E*>(a.pops) + int64_t(a.zpop) * z + int64_t(a.ypop) * y;
  // population at P + rel (rel: direction, x and neighbour offsets)
  auto load = [&](int64_t rel) -> E { return ld_pop<PAR>(P + rel); };
  auto store = [&](int64_t rel, E v) { st_pop<PAR>(P + rel, v); };
  // buffer form: the descriptor starts B elements before the row of the
  // wavefront's first active lane (the smallest row of its cells; every lane
  // of a row segment shares it), prel is the lane's row relative to that row
  const int32_t B = BUF ? (Lat<Q>::D == 3 ? int32_t(a.zpop) : 0) + a.ypop + 1 : 0;
  const int yb = __builtin_amdgcn_readfirstlane(y), zb = __builtin_amdgcn_readfirstlane(z);
  const uint32_t prel = uint32_t(a.zpop) * uint32_t(z - zb) + uint32_t(a.ypop) * uint32_t(y - yb);
  const auto rs = [&] {
    if constexpr (BUF)
      return pop_rsrc(a, sizeof(E), int64_t(kFrontPad) + a.zpop * zb + int64_t(a.ypop) * yb - B);
    else return 0;
  }();
  const uint8_t fl = a.flags[s0];
  float cs = 0.0f;
  const int dyn = match_task(a, fl, cs);
  if (dyn == DYN_NONE) return;
  unsigned ln = 0u, lm = 0u;
  if (!EVEN && (fl & a.bcany)) {
    // halfway links of a boundary cell, precomputed from the flags (k_links)
    const uint64_t m = a.links[s0];
    ln = uint32_t(m);
    lm = uint32_t(m >> 32);
  }
  if constexpr (BUF && EVEN && !PER) {
    // the even step touches the cell's own slots only: one vector offset,
    // the direction in the scalar offset
    constexpr uint32_t el = sizeof(E);
    const uint32_t vo = (prel + uint32_t(c0)) * el;
    const int32_t dp = int32_t(a.dpop);
    R f[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = pop_dec<Q, S>(bld_pop_s<PAR, E>(rs, vo, uint32_t(B + i * dp) * el), i);
    const R om = R(a.omega);
    collide<Q, R, FAM, kDev<Q, S>>(a, s0, dyn, cs, f, om, R(1) - om);
#pragma unroll
    for (int i = 0; i < Q; ++i)
      bst_pop_s<PAR, E>(rs, vo, uint32_t(B + opp(i) * dp) * el, pop_enc<Q, S>(f[i], i));
    return;
  }
  if constexpr (BUF && !EVEN && !PER) {
    // rows of 256 slots and more start on line boundaries: their c_x = 0
    // directions are whole lines of this wavefront alone
    if (FLB_NT_ODD_ALIGNED && a.pitch >= 256) odd_buf<Q, S, FAM, true>(a, rs, prel, c0, s0, dyn, cs, ln, lm);
    else odd_buf<Q, S, FAM, false>(a, rs, prel, c0, s0, dyn, cs, ln, lm);
    return;
  }
  const unsigned lk = ln | lm;
  R f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (EVEN) {
      f[i] = pop_dec<Q, S>(load(int64_t(i) * a.dpop + c0), i);
    } else {
      // upstream neighbour x - c_i a wall: own f*_{opp i}, left in A[i][x]
      const int o = opp(i);
      const bool wall = i != 0 && ((lk >> o) & 1u);
      const int64_t rel = wall ? int64_t(i) * a.dpop + c0
                               : int64_t(o) * a.dpop + c0 + nb_off<Q, PER, true>(a, o, x, y, z);
      f[i] = pop_dec<Q, S>(load(rel), i);
    }
  }
  if (!EVEN) wall_terms_in<Q>(a, s0, lm, f);
  const R om = R(a.omega);
  collide<Q, R, FAM, kDev<Q, S>>(a, s0, dyn, cs, f, om, R(1) - om);
  if (!EVEN) wall_terms_out<Q>(a, s0, lm, f);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (EVEN) {
      store(int64_t(opp(i)) * a.dpop + c0, pop_enc<Q, S>(f[i], i));
    } else {
      // downstream neighbour x + c_i a wall: f_{opp i}(x, t+1) = f*_i (+ wall term)
      const bool wall = i != 0 && ((lk >> i) & 1u);
      const int64_t rel = wall ? int64_t(opp(i)) * a.dpop + c0
                               : int64_t(i) * a.dpop + c0 + nb_off<Q, PER, true>(a, i, x, y, z);
      store(rel, pop_enc<Q, S>(f[i], i));
    }
  }
}

// The even step of the 16-bit storage kinds with two cells per lane: each
// direction is one 32-bit word per lane, so a wavefront moves 256 bytes per
// load or store instruction, as the FP32 step does with one cell per lane
// (one cell per lane moves 128). The even step has no halfway handling (its
// walls are the odd step's), so the two cells of a lane are independent:
// each is collided on its own, and a pair whose cells are not both updated
// stores its live half alone. Needs an even row pitch (every row and
// direction starts on a 4-byte boundary) and row segments, a.wpr counting
// wavefronts of 128 cells. Below 4 GiB of populations the words go through
// the buffer descriptor: one lane offset for the pair, the direction's offset
// in the scalar operand (no per-direction address arithmetic per lane).
#ifndef FLB_EVEN_PAIR
#define FLB_EVEN_PAIR 1
#endif
template <int Q, class S, int PAR, int FAM>
__global__ __launch_bounds__(kMaxThreads) void k_step_even_pair(KArgs a) {
  using R = typename S::R;
  using E = typename S::E;
  static_assert(sizeof(E) == 2, "two 16-bit populations per 32-bit word");
  const int lane = int(threadIdx.x) & (kWave - 1);
  const int wib = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  const int gw = int(xcd_block(blockIdx.x, unsigned(a.nblocks))) * a.wpb + wib;
  const int row = gw / a.wpr;
  if (row >= a.ey * a.ez) return;
  const int wv = gw - row * a.wpr;
  const int y = a.ys + row % a.ey;
  const int z = a.zs + row / a.ey;
  const int c0 = 2 * (wv * kWave + lane);
  if (c0 >= a.ex) return;
  const bool two = c0 + 1 < a.ex;
  const int s0 = c0 + a.pitch * y + int(a.plane) * z;
  E* const P = static_cast<E*>(a.pops) + int64_t(a.zpop) * z + int64_t(a.ypop) * y;
  const bool bufa = a.buf != 0;
  constexpr int aux = (PAR == 0 && FLB_NT_EVEN) ? 2 : 0;  // 2: non-temporal
  // descriptor from the row's first slot (the row is wavefront-uniform)
  const auto rs = pop_rsrc(a, sizeof(E), int64_t(kFrontPad) + a.zpop * z + int64_t(a.ypop) * y);
  const uint32_t vo = uint32_t(c0) * 2u;
  const int32_t dp = int32_t(a.dpop);
  float cs0 = 0.0f, cs1 = 0.0f;
  const int d0 = match_task(a, a.flags[s0], cs0);
  const int d1 = two ? match_task(a, a.flags[s0 + 1], cs1) : int(DYN_NONE);
  if (d0 == DYN_NONE && d1 == DYN_NONE) return;
  R f0[Q], f1[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    uint32_t w;
    if (bufa) {
      w = __builtin_amdgcn_raw_buffer_load_b32(rs, vo, uint32_t(i * dp) * 2u, aux);
    } else {
      w = ld_pop<PAR>(reinterpret_cast<const uint32_t*>((P + int64_t(i) * a.dpop) + c0));
    }
    f0[i] = pop_dec<Q, S>(E(w & 0xffffu), i);
    f1[i] = pop_dec<Q, S>(E(w >> 16), i);
  }
  const R om = R(a.omega);
  if (d0 != DYN_NONE) collide<Q, R, FAM, kDev<Q, S>>(a, s0, d0, cs0, f0, om, R(1) - om);
  if (d1 != DYN_NONE) collide<Q, R, FAM, kDev<Q, S>>(a, s0 + 1, d1, cs1, f1, om, R(1) - om);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    E* const q = (P + int64_t(opp(i)) * a.dpop) + c0;
    const E e0 = pop_enc<Q, S>(f0[i], i), e1 = pop_enc<Q, S>(f1[i], i);
    if (bufa && d0 != DYN_NONE && d1 != DYN_NONE) {
      __builtin_amdgcn_raw_buffer_store_b32(uint32_t(e0) | (uint32_t(e1) << 16), rs, vo,
                                            uint32_t(opp(i) * dp) * 2u, aux);
    } else if (d0 != DYN_NONE && d1 != DYN_NONE) {
      st_pop<PAR>(reinterpret_cast<uint32_t*>(q), uint32_t(e0) | (uint32_t(e1) << 16));
    } else if (d0 != DYN_NONE) {
      st_pop<PAR>(q, e0);
    } else {
      st_pop<PAR>(q + 1, e1);
    }
  }
}

template <int Q, class S, int PAR, int FAM, bool PER, bool BUF = false>
__global__ __launch_bounds__(kMaxThreads) void k_step_cell(KArgs a) {
  const int lane = int(threadIdx.x) & (kWave - 1);
  const int wib = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  const int gw = int(xcd_block(blockIdx.x, unsigned(a.nblocks))) * a.wpb + wib;
  cell_update<Q, S, PAR, FAM, PER, BUF>(a, gw, lane);
}

// ---------------------------------------------------------------------------
// The batched step: many blocks of one lattice and storage kind in one launch.
// tab[e] holds block e's kernel arguments and wave0[e] the number of
// wavefronts (64-cell row segments) of blocks 0..e-1; wave0[nb] is the total.
// Wavefronts are taken wpb at a time ("groups"); workgroup b walks groups
// xcd_block(b) + k * gridDim.x, so a grid smaller than the group count runs
// persistently and a grid of one workgroup per group runs each once. Every
// wavefront reads its block from wblk (one scalar load).
// ---------------------------------------------------------------------------
template <int Q, class S, int PAR, int FAM, bool PER>
__global__ __launch_bounds__(kMaxThreads) void k_step_batch(const KArgs* __restrict__ tab,
                                                            const int64_t* __restrict__ wave0,
                                                            const int* __restrict__ wblk,
                                                            int nb, int64_t groups, int wpb) {
  const int lane = int(threadIdx.x) & (kWave - 1);
  const int wib = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  if (wib >= wpb) return;
  const int64_t total = wave0[nb];
  for (int64_t g = xcd_block(blockIdx.x, gridDim.x); g < groups; g += gridDim.x) {
    const int64_t w = g * wpb + wib;
    if (w >= total) break;
    const int e = __builtin_amdgcn_readfirstlane(wblk[w]);
    cell_update<Q, S, PAR, FAM, PER>(tab[e], int(w - wave0[e]), lane);
  }
}

// ---------------------------------------------------------------------------
// The step kernel, row form: one workgroup per row, VEC cells per thread, no
// periodic axis. Every store is an aligned vector and, at the odd step, the
// stores of each direction cover whole 128-byte lines: a cell whose source
// lies outside the row (the overlap layer, the row padding) contributes an
// arbitrary value, written only into slots that no cell ever reads (ghost
// slots of directions pointing out of the block, row padding) or that the same
// thread overwrites right after with its halfway value. Partial-line writes
// would otherwise cost a read-modify-write in the memory system.
// ---------------------------------------------------------------------------
template <int Q, class S, int PAR, int FAM, int VEC>
__global__ __launch_bounds__(kMaxThreads) void k_step_row(KArgs a) {
  using R = typename S::R;
  using E = typename S::E;
  constexpr unsigned full = (1u << VEC) - 1u;
  constexpr int L = 128 / int(sizeof(E));  // elements of one cache line
  constexpr int kMaxW = kMaxThreads / kWave;
  __shared__ R xfirst[kMaxW][Q];  // each wavefront's lane-0 cell 0 values
  __shared__ R xlast[kMaxW][Q];   // each wavefront's lane-63 cell VEC-1 values
  __shared__ unsigned xmask[kMaxW][2];
  const int lane = int(threadIdx.x) & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  const int W = a.wpr;  // wavefronts per row (= per workgroup)
  const int row = int(xcd_block(blockIdx.x, unsigned(a.nblocks)));
  const int y = a.ys + row % a.ey;
  const int z = a.zs + row / a.ey;
  const int c0 = (wv * kWave + lane) * VEC;  // x - xs of this thread's first cell
  const int s0 = c0 + a.pitch * y + int(a.plane) * z;
  E* const P = static_cast<E*>(a.pops) + int64_t(a.zpop) * z + int64_t(a.ypop) * y;

  // --- dynamics ---------------------------------------------------------------------
  uint8_t fl[VEC];
  load_flags<VEC>(a.flags + s0, fl);
  int dyn[VEC];
  float cs[VEC];
  unsigned live = 0u, beyond = 0u;
#pragma unroll
  for (int j = 0; j < VEC; ++j) {
    cs[j] = 0.0f;
    dyn[j] = DYN_NONE;
    if (c0 + j < a.ex) dyn[j] = match_task(a, fl[j], cs[j]);
    else beyond |= 1u << j;
    if (dyn[j] != DYN_NONE) live |= 1u << j;
  }
  // slots this thread may write: live cells, and cells past the row end
  // (their slots are the ghost cell and padding, never read)
  const unsigned wmask = live | beyond;

  unsigned ln[VEC], lm[VEC];
  unsigned anylink = 0u;
#pragma unroll
  for (int j = 0; j < VEC; ++j) {
    ln[j] = 0u;
    lm[j] = 0u;
    if (PAR == 1 && a.nbc && ((live >> j) & 1u))
      wall_links<Q, false>(a, a.xs + c0 + j, y, z, s0 + j, fl[j], ln[j], lm[j]);
    anylink |= ln[j] | lm[j];
  }

  // --- read ---------------------------------------------------------------------------
  R f[VEC][Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    R v[VEC];
    if (PAR == 0) {
      load_row<Q, S, VEC>((P + int64_t(i) * a.dpop) + c0, i, v);
    } else {
      const int o = opp(i);
      const int cx = Lat<Q>::cx(o);  // source x + c_o = x - c_i
      const E* base = (P + int64_t(o) * a.dpop) + c0 + yz_off<Q, false, true>(a, o, y, z);
      R w[VEC];
      load_row<Q, S, VEC>(base, i, w);
      if (cx == 0) {
#pragma unroll
        for (int j = 0; j < VEC; ++j) v[j] = w[j];
      } else if (cx > 0) {
        R nxt = wave_shl(w[0]);
        if (lane == kWave - 1) nxt = pop_dec<Q, S>(base[VEC], i);
#pragma unroll
        for (int j = 0; j < VEC - 1; ++j) v[j] = w[j + 1];
        v[VEC - 1] = nxt;
      } else {
        R prv = wave_shr(w[VEC - 1]);
        if (lane == 0) prv = pop_dec<Q, S>(base[-1], i);
        v[0] = prv;
#pragma unroll
        for (int j = 1; j < VEC; ++j) v[j] = w[j - 1];
      }
    }
#pragma unroll
    for (int j = 0; j < VEC; ++j) f[j][i] = v[j];
  }
  if (PAR == 1 && anylink) {
    // upstream neighbour x - c_i a wall: own f*_{opp i}, left in A[i][x]
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const unsigned lk = ln[j] | lm[j];
      if (!lk) continue;
      R rho_f = R(1), u_f[3] = {R(0), R(0), R(0)};
      if (lm[j]) load_fields<Q>(a, s0 + j, rho_f, u_f);
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        if ((lk >> opp(i)) & 1u) {
          R v = pop_dec<Q, S>((P + int64_t(i) * a.dpop)[c0 + j], i);
          if ((lm[j] >> opp(i)) & 1u) v = v + movingwall_term<Q>(i, rho_f, u_f);
          f[j][i] = v;
        }
      }
    }
  }

  // --- collide --------------------------------------------------------------------------
  const R om = R(a.omega);
  const R _om = R(1) - om;
#pragma unroll
  for (int j = 0; j < VEC; ++j)
    if ((live >> j) & 1u) collide<Q, R, FAM, kDev<Q, S>>(a, s0 + j, dyn[j], cs[j], f[j], om, _om);

  // --- write ------------------------------------------------------------------------------
  if (PAR == 0) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      R v[VEC];
#pragma unroll
      for (int j = 0; j < VEC; ++j) v[j] = f[j][i];
      store_row<Q, S, VEC>((P + int64_t(opp(i)) * a.dpop) + c0, i, v, wmask);
    }
    return;
  }
  // odd step: the values crossing a wavefront boundary go through LDS
  if (W > 1) {
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < Q; ++i) xfirst[wv][i] = f[0][i];
      xmask[wv][0] = wmask & 1u;
    }
    if (lane == kWave - 1) {
#pragma unroll
      for (int i = 0; i < Q; ++i) xlast[wv][i] = f[VEC - 1][i];
      xmask[wv][1] = (wmask >> (VEC - 1)) & 1u;
    }
    __syncthreads();
  }
  const bool row_tail = (wv == W - 1) && (a.ex == W * kWave * VEC);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = Lat<Q>::cx(i);  // destination x + c_i
    E* base = (P + int64_t(i) * a.dpop) + c0 + yz_off<Q, false, true>(a, i, y, z);
    R w[VEC];
    unsigned m;
    if (cx == 0) {
#pragma unroll
      for (int j = 0; j < VEC; ++j) w[j] = f[j][i];
      m = wmask;
    } else if (cx > 0) {
      // slot j receives cell j-1; slot 0 the previous lane's last cell
      R prv = wave_shr(f[VEC - 1][i]);
      unsigned pm = (wave_shr(wmask) >> (VEC - 1)) & 1u;
      if (lane == 0) {
        if (wv > 0) {
          prv = xlast[wv - 1][i];
          pm = xmask[wv - 1][1];
        } else {
          pm = 1u;  // source is the overlap cell: slot x = xs is never read unfixed
        }
      }
      w[0] = prv;
#pragma unroll
      for (int j = 1; j < VEC; ++j) w[j] = f[j - 1][i];
      m = pm | ((wmask << 1) & full);
      if (row_tail) {
        // the row's last cell moves into the next line: write that whole line
        const R last = __shfl(f[VEC - 1][i], kWave - 1, kWave);
        if (lane < L / VEC) {
          R t[VEC];
#pragma unroll
          for (int j = 0; j < VEC; ++j) t[j] = (lane == 0 && j == 0) ? last : R(0);
          store_row<Q, S, VEC>(base - c0 + a.ex + lane * VEC, i, t, full);
        }
      }
    } else {
      // slot j receives cell j+1; the last slot the next lane's first cell
      R nxt = wave_shl(f[0][i]);
      unsigned nm = wave_shl(wmask) & 1u;
      if (lane == kWave - 1) {
        if (wv < W - 1) {
          nxt = xfirst[wv + 1][i];
          nm = xmask[wv + 1][0];
        } else {
          nm = 1u;  // past the row end: ghost / padding slot
        }
      }
#pragma unroll
      for (int j = 0; j < VEC - 1; ++j) w[j] = f[j + 1][i];
      w[VEC - 1] = nxt;
      m = (wmask >> 1) | (nm << (VEC - 1));
      if (wv == 0) {
        // the row's first cell moves to the overlap cell x = xs - 1, the last
        // slot of the line before the row: write that whole line
        const R first = __shfl(f[0][i], 0, kWave);
        if (lane < L / VEC) {
          R t[VEC];
#pragma unroll
          for (int j = 0; j < VEC; ++j) t[j] = (lane == L / VEC - 1 && j == VEC - 1) ? first : R(0);
          store_row<Q, S, VEC>(base - c0 - L + lane * VEC, i, t, full);
        }
      }
    }
    store_row<Q, S, VEC>(base, i, w, m);
  }
  if (anylink) {
    // downstream neighbour x + c_i a wall: f_{opp i}(x, t+1) = f*_i (+ wall term)
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const unsigned lk = ln[j] | lm[j];
      if (!lk) continue;
      R rho_f = R(1), u_f[3] = {R(0), R(0), R(0)};
      if (lm[j]) load_fields<Q>(a, s0 + j, rho_f, u_f);
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        if ((lk >> i) & 1u) {
          R v = f[j][i];
          if ((lm[j] >> i) & 1u) v = v + movingwall_term<Q>(opp(i), rho_f, u_f);
          (P + int64_t(opp(i)) * a.dpop)[c0 + j] = pop_enc<Q, S>(v, i);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Per-cell companions (one thread per host cell, dense host index).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool host_cell(const KArgs& a, int64_t id, int& x, int& y, int& z) {
  x = int(id % a.nx);
  const int64_t r = id / a.nx;
  y = int(r % a.ny);
  z = int(r / a.ny);
  return true;
}

// Slot of cell (x, y, z) in the per-cell arrays (flags, fields) ...
__device__ __forceinline__ int slot_of(const KArgs& a, int x, int y, int z) {
  return (x - a.xs) + a.pitch * y + int(a.plane) * z;
}
// ... and in the population arrays (64-bit: interleaved planes of a large
// block pass 2^31 slots).
__device__ __forceinline__ int64_t pslot_of(const KArgs& a, int x, int y, int z) {
  return int64_t(x - a.xs) + int64_t(a.ypop) * y + a.zpop * z;
}

__device__ __forceinline__ bool updated(const KArgs& a, int x, int y, int z) {
  return x >= a.xs && x < a.xs + a.ex && y >= a.ys && y < a.ys + a.ey && z >= a.zs &&
         z < a.zs + a.ez;
}

// Logical (post-stream) populations of one updated cell for the given parity
// (slot: flags / fields slot, pslot: population slot of the cell).
template <int Q, class S, int PAR>
__device__ __forceinline__ void read_logical(const KArgs& a, int x, int y, int z, int slot,
                                             int64_t pslot, typename S::R (&f)[Q]) {
  using R = typename S::R;
  if (PAR == 0) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = S::template from<Q>(arr<S>(a, i)[pslot], i);
    return;
  }
  unsigned ln = 0u, lm = 0u;
  R rho_f = R(1), u_f[3] = {R(0), R(0), R(0)};
  if (a.nbc) {
    wall_links<Q, true>(a, x, y, z, slot, a.flags[slot], ln, lm);
    if (lm) load_fields<Q>(a, slot, rho_f, u_f);
  }
  const unsigned lk = ln | lm;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int o = opp(i);
    if (i != 0 && ((lk >> o) & 1u)) {
      R v = S::template from<Q>(arr<S>(a, i)[pslot], i);
      if ((lm >> o) & 1u) {
        if constexpr (kScalarLat<Q>) v = anti_term<Q>(i, rho_f, v);
        else v = v + movingwall_term<Q>(i, rho_f, u_f);
      }
      f[i] = v;
    } else {
      f[i] = S::template from<Q>(arr<S>(a, o)[pslot + nb_off<Q, true, true>(a, o, x, y, z)], i);
    }
  }
}

template <int Q, class S, int PAR>
__global__ __launch_bounds__(kMaxThreads) void k_gather(KArgs a, typename S::R* out) {
  using R = typename S::R;
  const int64_t N = int64_t(a.nx) * a.ny * a.nz;
  const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= N) return;
  int x, y, z;
  host_cell(a, id, x, y, z);
  R f[Q];
  if (updated(a, x, y, z)) {
    read_logical<Q, S, PAR>(a, x, y, z, slot_of(a, x, y, z), pslot_of(a, x, y, z), f);
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = R(0);
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) out[i * N + id] = f[i];
}

template <int Q, class S, int PAR>
__global__ __launch_bounds__(kMaxThreads) void k_moments(KArgs a, uint8_t mask,
                                                         typename S::R* out) {
  using R = typename S::R;
  const int64_t N = int64_t(a.nx) * a.ny * a.nz;
  const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= N) return;
  int x, y, z;
  host_cell(a, id, x, y, z);
  R rho = R(0), u[3] = {R(0), R(0), R(0)};
  if (updated(a, x, y, z)) {
    const int slot = slot_of(a, x, y, z);
    if (a.flags[slot] & mask) {
      R f[Q];
      read_logical<Q, S, PAR>(a, x, y, z, slot, pslot_of(a, x, y, z), f);
      rhou(f, rho, u);
    }
  }
  out[id] = rho;
#pragma unroll
  for (int d = 0; d < Lat<Q>::D; ++d) out[(d + 1) * N + id] = u[d];
}

template <int Q, class S, int PAR>
__global__ __launch_bounds__(kMaxThreads) void k_reduce(KArgs a, uint8_t mask, double* partials) {
  using R = typename S::R;
  __shared__ double sh[4][kMaxThreads];
  const int tid = threadIdx.x;
  const int64_t N = int64_t(a.nx) * a.ny * a.nz;
  const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  if (id < N) {
    int x, y, z;
    host_cell(a, id, x, y, z);
    if (updated(a, x, y, z)) {
      const int slot = slot_of(a, x, y, z);
      if (a.flags[slot] & mask) {
        R f[Q];
        read_logical<Q, S, PAR>(a, x, y, z, slot, pslot_of(a, x, y, z), f);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
          const double v = double(f[i]);
          m[0] += v;
          m[1] += double(Lat<Q>::cx(i)) * v;
          m[2] += double(Lat<Q>::cy(i)) * v;
          m[3] += double(Lat<Q>::cz(i)) * v;
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) sh[c][tid] = m[c];
  __syncthreads();
  for (int s = kMaxThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sh[c][tid] += sh[c][tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) partials[int64_t(blockIdx.x) * 4 + c] = sh[c][0];
  }
}

// Sums the per-workgroup partials in a fixed order (deterministic result).
// (A template only so each storage kind's translation unit owns its copy.)
template <class S>
__global__ __launch_bounds__(kMaxThreads) void k_reduce_final(const double* partials, int n,
                                                             double* out) {
  __shared__ double sh[4][kMaxThreads];
  const int tid = threadIdx.x;
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = tid; b < n; b += kMaxThreads) {
#pragma unroll
    for (int c = 0; c < 4; ++c) m[c] += partials[int64_t(b) * 4 + c];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) sh[c][tid] = m[c];
  __syncthreads();
  for (int s = kMaxThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sh[c][tid] += sh[c][tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] = sh[c][0];
  }
}

// Natural-layout store of every host cell (overlap layer included).
template <int Q, class S>
__global__ __launch_bounds__(kMaxThreads) void k_scatter(KArgs a, const typename S::R* in) {
  const int64_t N = int64_t(a.nx) * a.ny * a.nz;
  const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= N) return;
  int x, y, z;
  host_cell(a, id, x, y, z);
  const int64_t slot = pslot_of(a, x, y, z);
#pragma unroll
  for (int i = 0; i < Q; ++i) arr<S>(a, i)[slot] = S::template to<Q>(in[i * N + id], i);
}

template <int Q, class S>
__global__ __launch_bounds__(kMaxThreads) void k_init_feq(KArgs a, const typename S::R* rhou_in) {
  using R = typename S::R;
  const int64_t N = int64_t(a.nx) * a.ny * a.nz;
  const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= N) return;
  int x, y, z;
  host_cell(a, id, x, y, z);
  const int64_t slot = pslot_of(a, x, y, z);
  R u[3] = {R(0), R(0), R(0)};
#pragma unroll
  for (int d = 0; d < Lat<Q>::D; ++d) u[d] = rhou_in[(d + 1) * N + id];
  R e[Q];
  feq(e, rhou_in[id], u);
#pragma unroll
  for (int i = 0; i < Q; ++i) arr<S>(a, i)[slot] = S::template to<Q>(e[i], i);
}

// ---------------------------------------------------------------------------
// Halfway wall links of every updated cell, bit k of the low (high) word set
// when the neighbour x + c_k is a wall of a normal (moving) halfway manager
// (0 for cells no manager applies to). Computed once per flag or manager
// change, so the odd step reads one word per boundary cell instead of
// probing its q - 1 neighbours' flags.
// ---------------------------------------------------------------------------
template <int Q, bool PER>
__global__ __launch_bounds__(kMaxThreads) void k_links(KArgs a, uint64_t* __restrict__ links) {
  const int64_t k = int64_t(blockIdx.x) * kMaxThreads + threadIdx.x;
  const int64_t n = int64_t(a.ex) * a.ey * a.ez;
  if (k >= n) return;
  const int c0 = int(k % a.ex);
  const int64_t r = k / a.ex;
  const int y = a.ys + int(r % a.ey), z = a.zs + int(r / a.ey);
  const int s0 = c0 + a.pitch * y + int(a.plane) * z;
  unsigned ln = 0u, lm = 0u;
  const uint8_t fl = a.flags[s0];
  if (fl & a.bcany) wall_links<Q, PER>(a, a.xs + c0, y, z, s0, fl, ln, lm);
  links[s0] = (uint64_t(lm) << 32) | ln;
}

// ---------------------------------------------------------------------------
// Launch helpers for one storage kind.
// ---------------------------------------------------------------------------
// Buffer addressing (pop_rsrc) for the odd step and the plain even step
// (measured on the 256^3 cavity: the FP16-shift odd step 4 % faster, the
// refined cavity's compact-row steps 4 %); the non-temporal even step of
// wide rows keeps 64-bit global addressing, which measured 1 % faster there.
#ifndef FLB_BUFFER_ADDR
#define FLB_BUFFER_ADDR 1
#endif
template <int Q, class S, int PAR>
hipError_t step_cell_q(int family, bool per, const KArgs& a, hipStream_t s) {
  const dim3 grid(a.nblocks), block(kWave * a.wpb);
  const bool buf = FLB_BUFFER_ADDR && PAR != 0 && a.buf != 0 && !per;
  if (family == FAM_FULL) {
    if (per) hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_FULL, true>), grid, block, 0, s, a);
    else if (buf) hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_FULL, false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_FULL, false>), grid, block, 0, s, a);
  } else {
    if (per) hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_BGK, true>), grid, block, 0, s, a);
    else if (buf) hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_BGK, false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_BGK, false>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

template <int Q, class S, int PAR>
hipError_t step_q(int family, const KArgs& a, hipStream_t s) {
  const dim3 grid(a.nblocks), block(kWave * a.wpb);
  const bool per = a.px || a.py || a.pz;
  if constexpr (PAR == 2) {
    return step_cell_q<Q, S, 2>(family, per, a, s);
  } else {
    if (per || a.vec <= 0) return step_cell_q<Q, S, PAR>(family, per, a, s);
    if (family == FAM_FULL) {
      if (a.vec == 1) hipLaunchKernelGGL((k_step_row<Q, S, PAR, FAM_FULL, 1>), grid, block, 0, s, a);
      else hipLaunchKernelGGL((k_step_row<Q, S, PAR, FAM_FULL, 2>), grid, block, 0, s, a);
    } else if (a.vec == 1) {
      hipLaunchKernelGGL((k_step_row<Q, S, PAR, FAM_BGK, 1>), grid, block, 0, s, a);
    } else if (a.vec == 2) {
      hipLaunchKernelGGL((k_step_row<Q, S, PAR, FAM_BGK, 2>), grid, block, 0, s, a);
    } else {
      hipLaunchKernelGGL((k_step_row<Q, S, PAR, FAM_BGK, 4>), grid, block, 0, s, a);
    }
    return hipGetLastError();
  }
}

template <int Q, class S, int PAR>
hipError_t batch_q(int family, bool per, const KArgs* tab, const int64_t* wave0, const int* wblk, int nb,
                   int64_t groups, int wpb, int grid, hipStream_t s) {
  const dim3 g(grid), b(kWave * wpb);
  if (family == FAM_FULL) {
    if (per) hipLaunchKernelGGL((k_step_batch<Q, S, PAR, FAM_FULL, true>), g, b, 0, s, tab, wave0, wblk, nb, groups, wpb);
    else hipLaunchKernelGGL((k_step_batch<Q, S, PAR, FAM_FULL, false>), g, b, 0, s, tab, wave0, wblk, nb, groups, wpb);
  } else {
    if (per) hipLaunchKernelGGL((k_step_batch<Q, S, PAR, FAM_BGK, true>), g, b, 0, s, tab, wave0, wblk, nb, groups, wpb);
    else hipLaunchKernelGGL((k_step_batch<Q, S, PAR, FAM_BGK, false>), g, b, 0, s, tab, wave0, wblk, nb, groups, wpb);
  }
  return hipGetLastError();
}

// batches: the even step without non-temporal accesses (PAR 2)
template <class S>
hipError_t launch_batch(int q, int parity, int family, bool per, const KArgs* tab,
                        const int64_t* wave0, const int* wblk, int nb, int64_t groups, int wpb,
                        int grid, hipStream_t s) {
  if (q == 19)
    return parity ? batch_q<19, S, 1>(family, per, tab, wave0, wblk, nb, groups, wpb, grid, s)
                  : batch_q<19, S, 2>(family, per, tab, wave0, wblk, nb, groups, wpb, grid, s);
  if (q == 9)
    return parity ? batch_q<9, S, 1>(family, per, tab, wave0, wblk, nb, groups, wpb, grid, s)
                  : batch_q<9, S, 2>(family, per, tab, wave0, wblk, nb, groups, wpb, grid, s);
  return hipErrorInvalidValue;
}

// Storage kinds of the advection-diffusion (D2Q5 / D3Q7) blocks: the 16-bit
// kinds store f - w_i for populations near the rest state of unit density,
// which a scalar field of arbitrary magnitude is not.
template <class S>
constexpr bool kPlainStore = std::is_same<S, StoreF32>::value || std::is_same<S, StoreF64>::value;

// Calls f(std::integral_constant<int, Q>) for the block's lattice; false for a
// lattice this storage kind does not support.
template <class S, class F>
bool with_q(int q, F&& f) {
  switch (q) {
    case 19: f(std::integral_constant<int, 19>{}); return true;
    case 9: f(std::integral_constant<int, 9>{}); return true;
    default: break;
  }
  if constexpr (kPlainStore<S>) {
    if (q == 7) {
      f(std::integral_constant<int, 7>{});
      return true;
    }
    if (q == 5) {
      f(std::integral_constant<int, 5>{});
      return true;
    }
  }
  return false;
}

// Scalar lattices step in the cell form only (no row form, no batches).
template <int Q, class S, int PAR>
hipError_t step_scalar_q(const KArgs& a, hipStream_t s) {
  const dim3 grid(a.nblocks), block(kWave * a.wpb);
  const bool per = a.px || a.py || a.pz;
  const bool buf = FLB_BUFFER_ADDR && PAR != 0 && a.buf != 0 && !per;
  if (per) hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_SCALAR, true>), grid, block, 0, s, a);
  else if (buf) hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_SCALAR, false, true>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((k_step_cell<Q, S, PAR, FAM_SCALAR, false>), grid, block, 0, s, a);
  return hipGetLastError();
}

// The even step of a 16-bit block in the cell form over row segments, with
// an even pitch: two cells per lane (k_step_even_pair).
template <int Q, class S>
hipError_t even_pair_q(int family, bool nt, const KArgs& a0, hipStream_t s) {
  KArgs a = a0;
  a.wpr = (a.ex + 2 * kWave - 1) / (2 * kWave);
  a.nblocks = int((int64_t(a.wpr) * a.ey * a.ez + a.wpb - 1) / a.wpb);
  const dim3 grid(a.nblocks), block(kWave * a.wpb);
  if (family == FAM_FULL) {
    if (nt) hipLaunchKernelGGL((k_step_even_pair<Q, S, 0, FAM_FULL>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_step_even_pair<Q, S, 2, FAM_FULL>), grid, block, 0, s, a);
  } else {
    if (nt) hipLaunchKernelGGL((k_step_even_pair<Q, S, 0, FAM_BGK>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_step_even_pair<Q, S, 2, FAM_BGK>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

// Non-temporal even steps for blocks with wide rows (pitch of whole
// 256-slot groups) and for the row form; compact rows take PAR 2.
template <class S>
hipError_t launch_step(int q, int parity, int family, const KArgs& a, hipStream_t s) {
  const bool per = a.px || a.py || a.pz;
  const bool nt = a.pitch >= 256 || !(per || a.vec <= 0);
  if constexpr (sizeof(typename S::E) == 2 && FLB_EVEN_PAIR) {
    if (!parity && !per && (a.vec == 0 || a.vec == -1) && a.pitch % 2 == 0 && (q == 19 || q == 9))
      return q == 19 ? even_pair_q<19, S>(family, a.pitch >= 256, a, s)
                     : even_pair_q<9, S>(family, a.pitch >= 256, a, s);
  }
  if (q == 19)
    return parity ? step_q<19, S, 1>(family, a, s)
                  : (nt ? step_q<19, S, 0>(family, a, s) : step_q<19, S, 2>(family, a, s));
  if (q == 9)
    return parity ? step_q<9, S, 1>(family, a, s)
                  : (nt ? step_q<9, S, 0>(family, a, s) : step_q<9, S, 2>(family, a, s));
  if constexpr (kPlainStore<S>) {
    const bool ntc = a.pitch >= 256;
    if (q == 7)
      return parity ? step_scalar_q<7, S, 1>(a, s)
                    : (ntc ? step_scalar_q<7, S, 0>(a, s) : step_scalar_q<7, S, 2>(a, s));
    if (q == 5)
      return parity ? step_scalar_q<5, S, 1>(a, s)
                    : (ntc ? step_scalar_q<5, S, 0>(a, s) : step_scalar_q<5, S, 2>(a, s));
  }
  return hipErrorInvalidValue;
}

template <class S>
hipError_t launch_gather(int q, int parity, const KArgs& a, void* out, hipStream_t s) {
  const unsigned nb = cell_blocks(int64_t(a.nx) * a.ny * a.nz);
  using R = typename S::R;
  R* o = static_cast<R*>(out);
  const bool ok = with_q<S>(q, [&](auto QC) {
    constexpr int Q = decltype(QC)::value;
    if (parity) hipLaunchKernelGGL((k_gather<Q, S, 1>), dim3(nb), dim3(kMaxThreads), 0, s, a, o);
    else hipLaunchKernelGGL((k_gather<Q, S, 0>), dim3(nb), dim3(kMaxThreads), 0, s, a, o);
  });
  return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template <class S>
hipError_t launch_scatter(int q, const KArgs& a, const void* in, hipStream_t s) {
  const unsigned nb = cell_blocks(int64_t(a.nx) * a.ny * a.nz);
  using R = typename S::R;
  const R* p = static_cast<const R*>(in);
  const bool ok = with_q<S>(q, [&](auto QC) {
    constexpr int Q = decltype(QC)::value;
    hipLaunchKernelGGL((k_scatter<Q, S>), dim3(nb), dim3(kMaxThreads), 0, s, a, p);
  });
  return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template <class S>
hipError_t launch_init_feq(int q, const KArgs& a, const void* rhou, hipStream_t s) {
  const unsigned nb = cell_blocks(int64_t(a.nx) * a.ny * a.nz);
  using R = typename S::R;
  const R* p = static_cast<const R*>(rhou);
  const bool ok = with_q<S>(q, [&](auto QC) {
    constexpr int Q = decltype(QC)::value;
    hipLaunchKernelGGL((k_init_feq<Q, S>), dim3(nb), dim3(kMaxThreads), 0, s, a, p);
  });
  return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template <class S>
hipError_t launch_moments(int q, int parity, const KArgs& a, uint8_t mask, void* out,
                          hipStream_t s) {
  const unsigned nb = cell_blocks(int64_t(a.nx) * a.ny * a.nz);
  using R = typename S::R;
  R* o = static_cast<R*>(out);
  const bool ok = with_q<S>(q, [&](auto QC) {
    constexpr int Q = decltype(QC)::value;
    if (parity) hipLaunchKernelGGL((k_moments<Q, S, 1>), dim3(nb), dim3(kMaxThreads), 0, s, a, mask, o);
    else hipLaunchKernelGGL((k_moments<Q, S, 0>), dim3(nb), dim3(kMaxThreads), 0, s, a, mask, o);
  });
  return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template <class S>
hipError_t launch_reduce(int q, int parity, const KArgs& a, uint8_t mask, double* partials,
                         double* out, hipStream_t s) {
  const unsigned nb = cell_blocks(int64_t(a.nx) * a.ny * a.nz);
  const bool ok = with_q<S>(q, [&](auto QC) {
    constexpr int Q = decltype(QC)::value;
    if (parity) hipLaunchKernelGGL((k_reduce<Q, S, 1>), dim3(nb), dim3(kMaxThreads), 0, s, a, mask, partials);
    else hipLaunchKernelGGL((k_reduce<Q, S, 0>), dim3(nb), dim3(kMaxThreads), 0, s, a, mask, partials);
  });
  if (!ok) return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce_final<S>, dim3(1), dim3(kMaxThreads), 0, s, partials, int(nb), out);
  return hipGetLastError();
}

}  // namespace kern
}  // namespace flb

// Defines the launchers of flb_internal.hpp for one storage kind.
#define FLB_DEFINE_LAUNCHERS(SUFFIX, STORE)                                                     \
  namespace flb {                                                                              \
  hipError_t launch_step_##SUFFIX(int q, int parity, int family, const KArgs& a,              \
                                  hipStream_t s) {                                            \
    return kern::launch_step<STORE>(q, parity, family, a, s);                                 \
  }                                                                                            \
  hipError_t launch_gather_##SUFFIX(int q, int parity, const KArgs& a, void* out,             \
                                    hipStream_t s) {                                          \
    return kern::launch_gather<STORE>(q, parity, a, out, s);                                  \
  }                                                                                            \
  hipError_t launch_scatter_##SUFFIX(int q, const KArgs& a, const void* in, hipStream_t s) {  \
    return kern::launch_scatter<STORE>(q, a, in, s);                                          \
  }                                                                                            \
  hipError_t launch_init_feq_##SUFFIX(int q, const KArgs& a, const void* rhou, hipStream_t s) { \
    return kern::launch_init_feq<STORE>(q, a, rhou, s);                                       \
  }                                                                                            \
  hipError_t launch_moments_##SUFFIX(int q, int parity, const KArgs& a, uint8_t mask,         \
                                     void* out, hipStream_t s) {                              \
    return kern::launch_moments<STORE>(q, parity, a, mask, out, s);                           \
  }                                                                                            \
  hipError_t launch_reduce_##SUFFIX(int q, int parity, const KArgs& a, uint8_t mask,          \
                                    double* partials, double* out, hipStream_t s) {           \
    return kern::launch_reduce<STORE>(q, parity, a, mask, partials, out, s);                  \
  }                                                                                            \
  hipError_t launch_batch_##SUFFIX(int q, int parity, int family, bool per, const KArgs* tab, \
                                   const int64_t* wave0, const int* wblk, int nb,             \
                                   int64_t groups, int wpb, int grid, hipStream_t s) {        \
    return kern::launch_batch<STORE>(q, parity, family, per, tab, wave0, wblk, nb, groups,    \
                                     wpb, grid, s);                                           \
  }                                                                                            \
  }
