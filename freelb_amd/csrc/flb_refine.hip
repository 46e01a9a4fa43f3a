// Two-level grid refinement on the device: a coarse block (level 0) and a
// fine block at twice its resolution (level 1) exchange populations after
// every step, as FreeLB's BlockLatticeManager::FullCommunicate(count) does
// for refined block pairs (data_struct/block_lattice.hh:1579-1591, 1724-1811),
// behind the C ABI (include/flb_abi.h, flb_refine_*).
//
// Geometry. The fine block covers the coarse cells [o, o + m) on every axis
// and carries two fine ghost layers (one coarse cell) around them, so its
// extent is 2m + 4; fine local cell 2 + 2 (c - o) + b, b in {0, 1}, lies in
// coarse cell c. This is FreeLB's level-1 block with useHelperOlap = false
// (geometry/block_geometry3d.hh:174-190).
//
// Average (fine -> coarse, averCommunicate): every coarse cell of the outer
// coarse layer of the patch (the coarse blocks' overlap cells inside the fine
// block, BlockGeometry3D::AddtoSharedAverComm, block_geometry3d.hh:1153-1174)
// takes the mean of its 2^d fine cells -- populations, RHO and VELOCITY,
// summed in the order of BasicBlock::getRefinedCellIdx
// (geometry/basic_geometry.hh:1211-1239, x fastest) and scaled by 1/2^d
// (getAverage, parallel/communicator.h:608-620) -- converted with
// f_C = f_F + (omega_F / 2)(f_F - feq) (RefineConverter::getPopC,
// lbm/unit_converter.h:657-659).
//
// Interpolation (coarse -> fine, intpCommunicate / popIntpCommunicate): each
// fine ghost cell is sub-cell D = bx + 2 by + 4 bz of its coarse unit u and
// takes sum_j w[D][j] g(corner + j), corner = u - 1 + b, j over the 2^d
// cube cells with x fastest (weights getIntpWeight, parallel/communicator.h:575-600;
// stencils as BlockGeometry2D::AddtoSharedIntpComm, geometry/block_geometry2d.hh:988-1059),
// converted f_F = f_C - (omega_C / 4)(f_C - feq) (getPopF, unit_converter.h:646-648).
// In 3D FreeLB's builder (block_geometry3d.hh:1177-1340) places the targets of
// sub-cells 5, 6 and 7 at Fid + 2 XY (+1) instead of Fid + XY + 1, Fid + XY + Nx
// and Fid + XY + Nx + 1, which writes past the fine block; the sub-cell of a
// target here is its own position, the placement its weights and stencil
// describe. Only the inner fine ghost layer is written: the outer one is
// never read by an updated fine cell.
//
// feq of both conversions is Equilibrium::SecondOrder (lbm/equilibrium.h:49-75)
// of the averaged / interpolated RHO and VELOCITY fields, which the fine and
// coarse cells write when they collide (FLB_DYN_BGK_WRITE).
//
// AA layout. A post-stream population f_i(x) lives at A[i][x] when the block
// is at parity 0 and at A[opp i][x - c_i] at parity 1; sources are read and
// targets written there, so the next step of each block sees the exchanged
// values as its own streamed-in populations.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/flb_abi.h"
#include "flb_host.hpp"
#include "flb_kernels.hpp"

namespace flb {
namespace refine {

using kern::arr;

constexpr int kThreads = 256;

// Equilibrium::Order2 of direction k with u2 = |u|^2 (lbm/equilibrium.h:49-53,
// 69-75): w rho (1 + 3 uc + uc uc 0.5 * 9 - 3 u2 0.5), uc and u2 summed from 0
// in component order.
template <int Q, class R>
__device__ __forceinline__ R feq_order2(int k, R rho, const R (&u)[3]) {
  constexpr int D = Lat<Q>::D;
  R u2 = R(0);
#pragma unroll
  for (int d = 0; d < D; ++d) u2 += u[d] * u[d];
  R uc = R(0);
  uc += u[0] * R(Lat<Q>::cx(k));
  uc += u[1] * R(Lat<Q>::cy(k));
  if (D == 3) uc += u[2] * R(Lat<Q>::cz(k));
  return Lat<Q>::template w<R>(k) * rho * (R(1) + R(3) * uc + uc * uc * R(0.5) * R(9) - R(3) * u2 * R(0.5));
}

// getIntpWeight2D / 3D (parallel/communicator.h:575-600): w[D][j] =
// prod over axes of 3/4 where bit a of j differs from bit a of D (the nearer
// coarse cell of the stencil along that axis), else 1/4 -- exact binary
// fractions, the table's values.
template <int DIM, class R>
__device__ __forceinline__ R intp_weight(int D, int j) {
  R w = R(1);
#pragma unroll
  for (int a = 0; a < DIM; ++a) w *= (((D >> a) & 1) != ((j >> a) & 1)) ? R(0.75) : R(0.25);
  return w;
}

template <int Q, class R>
__device__ __forceinline__ void load_rhou(const KArgs& a, int slot, R& rho, R (&u)[3]) {
  rho = static_cast<const R*>(a.rho_f)[slot];
  const R* uf = static_cast<const R*>(a.u_f);
#pragma unroll
  for (int d = 0; d < 3; ++d) u[d] = d < Lat<Q>::D ? uf[d * a.fstride + slot] : R(0);
}

// Slot offset of the neighbour x + c_o in the population arrays (the
// refined blocks are not periodic).
template <int Q>
__device__ __forceinline__ int64_t pop_off(const KArgs& a, int o) {
  return int64_t(Lat<Q>::cx(o)) + int64_t(Lat<Q>::cy(o)) * a.ypop + int64_t(Lat<Q>::cz(o)) * a.zpop;
}

// Post-stream population i of the cell at population slot ps / field slot
// slot for the block's parity (the single-direction form of read_logical).
// LINKS: some source cell may carry halfway links (read from the
// precomputed link words); without them no flag is read.
template <int Q, class S, int PAR, bool LINKS>
__device__ __forceinline__ typename S::R read_post(const KArgs& a, int64_t ps, int slot, int i) {
  using R = typename S::R;
  if (PAR == 0 || i == 0) return S::template from<Q>(arr<S>(a, i)[ps], i);
  const int o = opp(i);
  if (LINKS && (a.flags[slot] & a.bcany)) {
    const uint64_t lk = a.links[slot];
    const unsigned ln = unsigned(lk), lm = unsigned(lk >> 32);
    if (((ln | lm) >> o) & 1u) {
      R v = S::template from<Q>(arr<S>(a, i)[ps], i);
      if ((lm >> o) & 1u) {
        R rho_f, u_f[3];
        kern::load_fields<Q>(a, slot, rho_f, u_f);
        v = v + movingwall_term<Q>(i, rho_f, u_f);
      }
      return v;
    }
  }
  return S::template from<Q>(arr<S>(a, o)[ps + pop_off<Q>(a, o)], i);
}

// Post-stream population i of the cell at population slot ps stored where
// the block's next step reads it (target cells carry no halfway links).
template <int Q, class S, int PAR>
__device__ __forceinline__ void write_post(const KArgs& a, int64_t ps, int i, typename S::R v) {
  if (PAR == 0 || i == 0) {
    arr<S>(a, i)[ps] = S::template to<Q>(v, i);
  } else {
    const int o = opp(i);
    arr<S>(a, o)[ps + pop_off<Q>(a, o)] = S::template to<Q>(v, i);
  }
}

// One coarse target and the lower corner of its fine cells, as population
// and field slots (cps / csl coarse, fps / fsl fine).
struct AvgItem {
  int64_t cps, fps;
  int csl, fsl;
};
// One fine target, the lower corner of its coarse stencil (slots as above)
// and its sub-cell.
struct IntpItem {
  int64_t fps, cps;
  int fsl, csl;
  int sub, pad;
};

// The 2^d cube cells of a corner, x fastest: population and field slots.
template <int Q>
__device__ __forceinline__ int64_t cube_ps(const KArgs& a, int64_t ps, int j) {
  return ps + (j & 1) + ((j >> 1) & 1) * int64_t(a.ypop) + ((j >> 2) & 1) * a.zpop;
}
__device__ __forceinline__ int cube_sl(const KArgs& a, int sl, int j) {
  return sl + (j & 1) + ((j >> 1) & 1) * a.pitch + ((j >> 2) & 1) * int(a.plane);
}

// One thread per (item, direction): the 2^d sources' fields are read by all
// q threads of an item (cached), each thread sums its own direction.
template <int Q, class S, int PC, int PF, bool LINKS>
__global__ __launch_bounds__(kThreads) void k_average(KArgs c, KArgs f, const AvgItem* __restrict__ items,
                                                      int n, typename S::R omega_f) {
  using R = typename S::R;
  constexpr int D = Lat<Q>::D;
  constexpr int NS = 1 << D;
  const int t = int(blockIdx.x) * kThreads + int(threadIdx.x);
  if (t >= n * Q) return;
  const int k = t / Q, i = t - k * Q;
  const AvgItem it = items[k];
  R acc = R(0), rho = R(0), u[3] = {R(0), R(0), R(0)};
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int sl = cube_sl(f, it.fsl, j);
    const R g = read_post<Q, S, PF, LINKS>(f, cube_ps<Q>(f, it.fps, j), sl, i);
    R r, v[3];
    load_rhou<Q>(f, sl, r, v);
    if (j == 0) {
      acc = g;
      rho = r;
#pragma unroll
      for (int d = 0; d < 3; ++d) u[d] = v[d];
    } else {
      acc = acc + g;
      rho = rho + r;
#pragma unroll
      for (int d = 0; d < 3; ++d) u[d] = u[d] + v[d];
    }
  }
  const R scale = D == 3 ? R(0.125) : R(0.25);
  rho = rho * scale;
#pragma unroll
  for (int d = 0; d < 3; ++d) u[d] = u[d] * scale;
  const R e = feq_order2<Q>(i, rho, u);
  const R p = acc * scale;
  write_post<Q, S, PC>(c, it.cps, i, p + R(0.5) * omega_f * (p - e));
}

template <int Q, class S, int PC, int PF, bool LINKS>
__global__ __launch_bounds__(kThreads) void k_interp(KArgs c, KArgs f, const IntpItem* __restrict__ items,
                                                     int n, typename S::R omega_c) {
  using R = typename S::R;
  constexpr int D = Lat<Q>::D;
  constexpr int NS = 1 << D;
  const int t = int(blockIdx.x) * kThreads + int(threadIdx.x);
  if (t >= n * Q) return;
  const int k = t / Q, i = t - k * Q;
  const IntpItem it = items[k];
  R acc = R(0), rho = R(0), u[3] = {R(0), R(0), R(0)};
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int sl = cube_sl(c, it.csl, j);
    const R g = read_post<Q, S, PC, LINKS>(c, cube_ps<Q>(c, it.cps, j), sl, i);
    R r, v[3];
    load_rhou<Q>(c, sl, r, v);
    const R w = intp_weight<D, R>(it.sub, j);
    if (j == 0) {
      acc = g * w;
      rho = r * w;
#pragma unroll
      for (int d = 0; d < 3; ++d) u[d] = v[d] * w;
    } else {
      acc = acc + g * w;
      rho = rho + r * w;
#pragma unroll
      for (int d = 0; d < 3; ++d) u[d] = u[d] + v[d] * w;
    }
  }
  const R e = feq_order2<Q>(i, rho, u);
  write_post<Q, S, PF>(f, it.fps, i, acc - R(0.25) * omega_c * (acc - e));
}

template <int Q, class S, bool LINKS>
hipError_t launch_average_q(const KArgs& c, const KArgs& f, int pc, int pf, const AvgItem* items, int n,
                            double omega_f, hipStream_t s) {
  using R = typename S::R;
  const dim3 g(unsigned((int64_t(n) * Q + kThreads - 1) / kThreads)), b(kThreads);
  const R om = R(omega_f);
  if (pc == 0 && pf == 0) hipLaunchKernelGGL((k_average<Q, S, 0, 0, LINKS>), g, b, 0, s, c, f, items, n, om);
  else if (pc == 0) hipLaunchKernelGGL((k_average<Q, S, 0, 1, LINKS>), g, b, 0, s, c, f, items, n, om);
  else if (pf == 0) hipLaunchKernelGGL((k_average<Q, S, 1, 0, LINKS>), g, b, 0, s, c, f, items, n, om);
  else hipLaunchKernelGGL((k_average<Q, S, 1, 1, LINKS>), g, b, 0, s, c, f, items, n, om);
  return hipGetLastError();
}

template <int Q, class S, bool LINKS>
hipError_t launch_interp_q(const KArgs& c, const KArgs& f, int pc, int pf, const IntpItem* items, int n,
                           double omega_c, hipStream_t s) {
  using R = typename S::R;
  const dim3 g(unsigned((int64_t(n) * Q + kThreads - 1) / kThreads)), b(kThreads);
  const R om = R(omega_c);
  if (pc == 0 && pf == 0) hipLaunchKernelGGL((k_interp<Q, S, 0, 0, LINKS>), g, b, 0, s, c, f, items, n, om);
  else if (pc == 0) hipLaunchKernelGGL((k_interp<Q, S, 0, 1, LINKS>), g, b, 0, s, c, f, items, n, om);
  else if (pf == 0) hipLaunchKernelGGL((k_interp<Q, S, 1, 0, LINKS>), g, b, 0, s, c, f, items, n, om);
  else hipLaunchKernelGGL((k_interp<Q, S, 1, 1, LINKS>), g, b, 0, s, c, f, items, n, om);
  return hipGetLastError();
}

template <class S, bool LINKS>
hipError_t launch_pair(int q, bool average, const KArgs& c, const KArgs& f, int pc, int pf, const void* items,
                       int n, double omega, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (average) {
    const auto* it = static_cast<const AvgItem*>(items);
    return q == 19 ? launch_average_q<19, S, LINKS>(c, f, pc, pf, it, n, omega, s)
                   : launch_average_q<9, S, LINKS>(c, f, pc, pf, it, n, omega, s);
  }
  const auto* it = static_cast<const IntpItem*>(items);
  return q == 19 ? launch_interp_q<19, S, LINKS>(c, f, pc, pf, it, n, omega, s)
                 : launch_interp_q<9, S, LINKS>(c, f, pc, pf, it, n, omega, s);
}

template <class S>
hipError_t launch_kind(int q, bool average, bool links, const KArgs& c, const KArgs& f, int pc, int pf,
                       const void* items, int n, double omega, hipStream_t s) {
  return links ? launch_pair<S, true>(q, average, c, f, pc, pf, items, n, omega, s)
               : launch_pair<S, false>(q, average, c, f, pc, pf, items, n, omega, s);
}

hipError_t launch(int storage, int q, bool average, bool links, const KArgs& c, const KArgs& f, int pc, int pf,
                  const void* items, int n, double omega, hipStream_t s) {
  switch (storage) {
    case FLB_STORE_F32: return launch_kind<StoreF32>(q, average, links, c, f, pc, pf, items, n, omega, s);
    case FLB_STORE_F16S: return launch_kind<StoreF16S>(q, average, links, c, f, pc, pf, items, n, omega, s);
    case FLB_STORE_BF16S: return launch_kind<StoreBF16S>(q, average, links, c, f, pc, pf, items, n, omega, s);
    case FLB_STORE_F64: return launch_kind<StoreF64>(q, average, links, c, f, pc, pf, items, n, omega, s);
    case FLB_STORE_F16C: return launch_kind<StoreF16C>(q, average, links, c, f, pc, pf, items, n, omega, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace refine
}  // namespace flb

using namespace flbh;
namespace RF = flb::refine;

struct flb_refine_s {
  int device = 0;
  hipStream_t stream = nullptr;
  flb_block coarse = nullptr;
  flb_block fine = nullptr;
  int origin[3] = {0, 0, 0};
  int m[3] = {1, 1, 1};
  int navg = 0, nintp = 0;
  bool avg_links = false, intp_links = false;  // some source cell is a halfway boundary cell
  RF::AvgItem* d_avg = nullptr;
  RF::IntpItem* d_intp = nullptr;
  int coarse_level = 0, max_level = 1;  // level gates of the exchanges
};

namespace {

bool inside_update(const flb_block_s* b, int x, int y, int z) {
  const int ov = b->desc.overlap;
  auto in = [&](int v, int n, int axis) {
    return b->desc.periodic[axis] ? (v >= 0 && v < n) : (v >= ov && v < n - ov);
  };
  return in(x, b->desc.nx, 0) && in(y, b->desc.ny, 1) && (b->d == 2 ? z == 0 : in(z, b->desc.nz, 2));
}

uint8_t flag_at(const flb_block_s* b, int x, int y, int z) {
  return b->h_flags[size_t(x) + size_t(b->desc.nx) * (size_t(y) + size_t(b->desc.ny) * z)];
}

// Population and field slots of host cell (x, y, z), as the kernels' pslot_of / slot_of.
int64_t pop_slot(const flb_block_s* b, int x, int y, int z) {
  return int64_t(x - b->desc.overlap) + int64_t(b->ypop) * y + b->zpop * z;
}
int field_slot(const flb_block_s* b, int x, int y, int z) {
  return int(int64_t(x - b->desc.overlap) + int64_t(b->pitch) * y + b->plane * z);
}

// Whether any cell of the 2^d cube at (x, y, z) is a halfway boundary cell.
bool cube_links(const flb_block_s* b, int x, int y, int z) {
  uint8_t any = 0;
  for (int k = 0; k < b->nbc; ++k) any |= b->bcs[k].cell_mask;
  if (!any) return false;
  for (int j = 0; j < (b->d == 3 ? 8 : 4); ++j)
    if (flag_at(b, x + (j & 1), y + ((j >> 1) & 1), z + ((j >> 2) & 1)) & any) return true;
  return false;
}

// A target cell must be updated (so its next step reads what is written) and
// must carry no halfway link (its written slots are the plain streaming ones).
int check_target(const flb_block_s* b, int x, int y, int z, const char* what) {
  if (!inside_update(b, x, y, z))
    return fail(FLB_ERR_LAYOUT, "%s cell (%d,%d,%d) lies outside the block's updated cells", what, x, y, z);
  const uint8_t fl = flag_at(b, x, y, z);
  for (int k = 0; k < b->nbc; ++k)
    if (fl & b->bcs[k].cell_mask)
      return fail(FLB_ERR_LAYOUT, "%s cell (%d,%d,%d) is a halfway boundary cell (flag 0x%02x)", what, x, y, z, fl);
  return FLB_OK;
}

int launch_comm(flb_refine_s* r, bool average) {
  flb_block_s* c = r->coarse;
  flb_block_s* f = r->fine;
  // sources read their halfway links from the precomputed link words
  int rc = FLB_OK;
  if ((rc = ensure_links(c)) || (rc = ensure_links(f))) return rc;
  const KArgs ca = make_args(c), fa = make_args(f);
  const int n = average ? r->navg : r->nintp;
  const void* items = average ? static_cast<const void*>(r->d_avg) : static_cast<const void*>(r->d_intp);
  const double om = average ? f->desc.omega : c->desc.omega;
  const bool links = average ? r->avg_links : r->intp_links;
  HIP_TRY(RF::launch(c->desc.storage, c->desc.q, average, links, ca, fa, c->parity, f->parity, items, n, om,
                     r->stream));
  return FLB_OK;
}

// Average into the coarse block (gated by its level), interpolation into the
// fine block (gated by the fine level).
int comm_average(flb_refine_s* r, int64_t count) {
  return level_due(count, r->coarse_level, r->max_level) ? launch_comm(r, true) : FLB_OK;
}
int comm_interpolate(flb_refine_s* r, int64_t count) {
  return level_due(count, r->coarse_level + 1, r->max_level) ? launch_comm(r, false) : FLB_OK;
}

int communicate(flb_refine_s* r, int64_t count) {
  const int rc = comm_average(r, count);
  return rc ? rc : comm_interpolate(r, count);
}

int check_comm(flb_refine_s* r, int64_t count) {
  if (!r) return fail(FLB_ERR_ARG, "null refinement handle");
  if (count < 0) return fail(FLB_ERR_ARG, "negative count");
  if (r->coarse->pending || r->fine->pending)
    return fail(FLB_ERR_STATE, "communication between ApplyCellDynamics and Stream");
  return FLB_OK;
}

}  // namespace

extern "C" {

int flb_refine_create(flb_block coarse, flb_block fine, const int* origin, void* stream, flb_refine* out) {
  if (!coarse || !fine || !origin || !out) return fail(FLB_ERR_ARG, "null argument");
  *out = nullptr;
  if (coarse == fine) return fail(FLB_ERR_ARG, "coarse and fine block are the same block");
  if (coarse->desc.q != fine->desc.q) return fail(FLB_ERR_ARG, "blocks of different lattices");
  if (coarse->desc.q != FLB_D2Q9 && coarse->desc.q != FLB_D3Q19)
    return fail(FLB_ERR_UNSUPPORTED, "refinement takes D2Q9 / D3Q19 blocks");
  if (coarse->desc.storage != fine->desc.storage) return fail(FLB_ERR_ARG, "blocks of different storage kinds");
  if (coarse->device != fine->device) return fail(FLB_ERR_ARG, "blocks on different devices");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (coarse->stream != s || fine->stream != s) return fail(FLB_ERR_ARG, "both blocks must run on the refinement's stream");
  if (!coarse->flags_set || !fine->flags_set) return fail(FLB_ERR_STATE, "flags not uploaded");
  if (!coarse->d_rho || !coarse->d_u || !fine->d_rho || !fine->d_u)
    return fail(FLB_ERR_STATE, "both blocks need RHO and VELOCITY fields (tasks with FLB_DYN_BGK_WRITE)");
  for (int a = 0; a < 3; ++a)
    if (fine->desc.periodic[a] || coarse->desc.periodic[a])
      return fail(FLB_ERR_UNSUPPORTED, "refined blocks are not periodic");
  if (fine->desc.overlap != 1) return fail(FLB_ERR_ARG, "the fine block's overlap must be 1 (its outer ghost layer)");
  const int d = coarse->d;
  const int fdim[3] = {fine->desc.nx, fine->desc.ny, fine->desc.nz};
  int m[3] = {1, 1, 1}, o[3] = {origin[0], origin[1], d == 3 ? origin[2] : 0};
  for (int a = 0; a < d; ++a) {
    if (fdim[a] < 6 || (fdim[a] - 4) % 2 != 0)
      return fail(FLB_ERR_ARG, "fine extent %d on axis %d is not 2m + 4 with m >= 1", fdim[a], a);
    m[a] = (fdim[a] - 4) / 2;
  }
  if (d == 2 && fine->desc.nz != 1) return fail(FLB_ERR_ARG, "D2Q9 fine block needs nz == 1");
  // the coarse stencils reach from o - 2 to o + m + 1, all updated coarse cells
  for (int a = 0; a < d; ++a) {
    const int cn = a == 0 ? coarse->desc.nx : (a == 1 ? coarse->desc.ny : coarse->desc.nz);
    if (o[a] - 2 < coarse->desc.overlap || o[a] + m[a] + 1 >= cn - coarse->desc.overlap)
      return fail(FLB_ERR_LAYOUT, "patch [%d, %d) on axis %d needs two updated coarse cells on each side", o[a],
                  o[a] + m[a], a);
  }
  std::vector<RF::AvgItem> avg;
  std::vector<RF::IntpItem> intp;
  bool avg_links = false, intp_links = false;
  int rc = FLB_OK;
  // outer coarse layer of the patch
  for (int z = 0; z < (d == 3 ? m[2] : 1); ++z)
    for (int y = 0; y < m[1]; ++y)
      for (int x = 0; x < m[0]; ++x) {
        const bool shell = x == 0 || x == m[0] - 1 || y == 0 || y == m[1] - 1 ||
                           (d == 3 && (z == 0 || z == m[2] - 1));
        if (!shell) continue;
        const int cx = o[0] + x, cy = o[1] + y, cz = o[2] + z;
        const int fx = 2 + 2 * x, fy = 2 + 2 * y, fz = d == 3 ? 2 + 2 * z : 0;
        if ((rc = check_target(coarse, cx, cy, cz, "average target"))) return rc;
        avg_links = avg_links || cube_links(fine, fx, fy, fz);
        avg.push_back(RF::AvgItem{pop_slot(coarse, cx, cy, cz), pop_slot(fine, fx, fy, fz),
                                  field_slot(coarse, cx, cy, cz), field_slot(fine, fx, fy, fz)});
      }
  // inner fine ghost layer: fine cells 1 .. 2m + 2 with one coordinate on 1 or 2m + 2
  auto unit = [](int fcell, int oo) { return oo + ((fcell - 2) >= 0 ? (fcell - 2) / 2 : -1); };
  for (int z = (d == 3 ? 1 : 0); z < (d == 3 ? 2 * m[2] + 3 : 1); ++z)
    for (int y = 1; y < 2 * m[1] + 3; ++y)
      for (int x = 1; x < 2 * m[0] + 3; ++x) {
        const bool ghost = x == 1 || x == 2 * m[0] + 2 || y == 1 || y == 2 * m[1] + 2 ||
                           (d == 3 && (z == 1 || z == 2 * m[2] + 2));
        if (!ghost) continue;
        const int fc[3] = {x, y, z};
        int corner[3] = {0, 0, 0}, sub = 0;
        for (int a = 0; a < d; ++a) {
          const int bbit = (fc[a] - 2) & 1;  // 0 on the lower, 1 on the upper half of the unit
          corner[a] = unit(fc[a], o[a]) - 1 + bbit;
          sub |= bbit << a;
        }
        if ((rc = check_target(fine, x, y, z, "interpolation target"))) return rc;
        intp_links = intp_links || cube_links(coarse, corner[0], corner[1], corner[2]);
        intp.push_back(RF::IntpItem{pop_slot(fine, x, y, z), pop_slot(coarse, corner[0], corner[1], corner[2]),
                                    field_slot(fine, x, y, z), field_slot(coarse, corner[0], corner[1], corner[2]),
                                    sub, 0});
      }
  flb_refine_s* r = new flb_refine_s();
  r->device = coarse->device;
  r->stream = s;
  r->coarse = coarse;
  r->fine = fine;
  for (int a = 0; a < 3; ++a) {
    r->origin[a] = o[a];
    r->m[a] = m[a];
  }
  r->navg = int(avg.size());
  r->nintp = int(intp.size());
  r->avg_links = avg_links;
  r->intp_links = intp_links;
  DeviceGuard g(r->device);
  if (hipMalloc(reinterpret_cast<void**>(&r->d_avg), std::max<size_t>(1, avg.size()) * sizeof(RF::AvgItem)) !=
          hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&r->d_intp), std::max<size_t>(1, intp.size()) * sizeof(RF::IntpItem)) !=
          hipSuccess) {
    flb_refine_destroy(r);
    return fail(FLB_ERR_NOMEM, "refinement item allocation failed");
  }
  if ((!avg.empty() && hipMemcpyAsync(r->d_avg, avg.data(), avg.size() * sizeof(RF::AvgItem), hipMemcpyHostToDevice,
                                      s) != hipSuccess) ||
      (!intp.empty() && hipMemcpyAsync(r->d_intp, intp.data(), intp.size() * sizeof(RF::IntpItem),
                                       hipMemcpyHostToDevice, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess) {
    flb_refine_destroy(r);
    return fail(FLB_ERR_HIP, "refinement item upload failed");
  }
  *out = r;
  return FLB_OK;
}

int flb_refine_destroy(flb_refine r) {
  if (!r) return FLB_OK;
  DeviceGuard g(r->device);
  (void)hipStreamSynchronize(r->stream);
  if (r->d_avg) (void)hipFree(r->d_avg);
  if (r->d_intp) (void)hipFree(r->d_intp);
  delete r;
  return FLB_OK;
}

int flb_refine_info(flb_refine r, int64_t* n_average, int64_t* n_interp) {
  if (!r) return fail(FLB_ERR_ARG, "null refinement handle");
  if (n_average) *n_average = r->navg;
  if (n_interp) *n_interp = r->nintp;
  return FLB_OK;
}

int flb_refine_set_levels(flb_refine r, int coarse_level, int max_level) {
  if (!r) return fail(FLB_ERR_ARG, "null refinement handle");
  if (coarse_level < 0 || max_level < coarse_level + 1 || max_level > 30)
    return fail(FLB_ERR_ARG, "coarse level %d / max level %d", coarse_level, max_level);
  r->coarse_level = coarse_level;
  r->max_level = max_level;
  return FLB_OK;
}

int flb_refine_average(flb_refine r, int64_t count) {
  int rc = check_comm(r, count);
  if (rc) return rc;
  DeviceGuard g(r->device);
  return comm_average(r, count);
}

int flb_refine_interpolate(flb_refine r, int64_t count) {
  int rc = check_comm(r, count);
  if (rc) return rc;
  DeviceGuard g(r->device);
  return comm_interpolate(r, count);
}

int flb_refine_communicate(flb_refine r, int64_t count) {
  int rc = check_comm(r, count);
  if (rc) return rc;
  DeviceGuard g(r->device);
  return communicate(r, count);
}

int flb_refine_run(flb_refine r, int64_t count0, int n) {
  if (!r) return fail(FLB_ERR_ARG, "null refinement handle");
  if (count0 < 0 || n < 0) return fail(FLB_ERR_ARG, "negative count");
  if (r->coarse_level != 0 || r->max_level != 1)
    return fail(FLB_ERR_STATE, "flb_refine_run steps a two-level pair; a deeper hierarchy steps its blocks by count");
  if (r->coarse->pending || r->fine->pending) return fail(FLB_ERR_STATE, "step while a Stream is pending");
  DeviceGuard g(r->device);
  Range range("flb_refine_run");
  int rc = FLB_OK;
  for (int64_t count = count0; count < count0 + n; ++count) {
    if ((rc = flb_step(r->fine, 1))) return rc;
    if (count % 2 == 0 && (rc = flb_step(r->coarse, 1))) return rc;
    if ((rc = communicate(r, count))) return rc;
  }
  return FLB_OK;
}

}  // extern "C"
