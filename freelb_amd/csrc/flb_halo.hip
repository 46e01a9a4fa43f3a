// Same-level block halo on the device: the overlap layer of each block
// mirrors the cells of its neighbours (FreeLB's normalCommunicate /
// normalFullCommunicate, data_struct/block_lattice.hh:1669-1703, with the
// cell pairs of BlockGeometry3D::InitComm, geometry/block_geometry3d.hh:531-605).
//
// In the AA pattern a block's update touches its overlap ("ghost") cells only
// at the odd step, and only in the directions j whose reader cell g - c_j is
// one of its own updated cells: it reads slot A[j][g] (the neighbour's
// post-collision value) and, in the same update, pushes its own value into
// that same slot. So the exchange has two phases, one after each step:
//   after an even step  neighbour's A[j][g]   -> this block's ghost A[j][g]
//   after an odd step   this block's ghost A[j][g] -> neighbour's A[j][g]
// over exactly those directions. Every block must be at the same parity.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/flb_abi.h"
#include "flb_cell.hpp"
#include "flb_host.hpp"
#include "flb_internal.hpp"

using flb::KArgs;
using namespace flbh;

namespace flb {
namespace halo {

using Item = HaloItem;

template <int Q>
__device__ __forceinline__ int64_t pslot(const KArgs& k, int j, int x, int y, int z) {
  return int64_t(x - k.xs) + int64_t(k.ypop) * y + k.zpop * z + k.dpop * j;
}

template <int Q, class E>
__global__ __launch_bounds__(kMaxThreads) void k_halo(const KArgs* __restrict__ tab,
                                                      const Item* __restrict__ items,
                                                      const int64_t* __restrict__ cell0, int nitems,
                                                      int to_ghost) {
  const int64_t k = int64_t(blockIdx.x) * kMaxThreads + threadIdx.x;
  if (k >= cell0[nitems]) return;
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cell0[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  const Item it = items[lo];
  const int64_t li = k - cell0[lo];
  const int bx = int(li % it.nx);
  const int64_t r = li / it.nx;
  const int by = int(r % it.ny), bz = int(r / it.ny);
  const int xa = it.gx + bx, ya = it.gy + by, za = it.gz + bz;
  const int xb = xa + it.sx, yb = ya + it.sy, zb = za + it.sz;
  const KArgs& A = tab[it.a];
  const KArgs& B = tab[it.b];
  E* const pa = static_cast<E*>(A.pops);
  E* const pb = static_cast<E*>(B.pops);
#pragma unroll
  for (int j = 1; j < Q; ++j) {
    // reader of slot j at the ghost cell: g - c_j, one of a's updated cells?
    const int xr = xa - Lat<Q>::cx(j), yr = ya - Lat<Q>::cy(j), zr = za - Lat<Q>::cz(j);
    if (xr < A.xs || xr >= A.xs + A.ex || yr < A.ys || yr >= A.ys + A.ey || zr < A.zs || zr >= A.zs + A.ez)
      continue;
    const int64_t ia = pslot<Q>(A, j, xa, ya, za), ib = pslot<Q>(B, j, xb, yb, zb);
    if (to_ghost) pa[ia] = pb[ib];
    else pb[ib] = pa[ia];
  }
}

template <int Q, class E>
hipError_t launch_q(const KArgs* tab, const Item* items, const int64_t* cell0, int nitems, int64_t total,
                    int to_ghost, hipStream_t s) {
  const dim3 grid(unsigned((total + kMaxThreads - 1) / kMaxThreads)), block(kMaxThreads);
  hipLaunchKernelGGL((k_halo<Q, E>), grid, block, 0, s, tab, items, cell0, nitems, to_ghost);
  return hipGetLastError();
}

}  // namespace halo
}  // namespace flb

namespace H = flb::halo;

namespace flbh {

// Enqueues the exchange that follows a step which left the blocks at `parity`.
hipError_t halo_launch(const flb_halo_s* h, int parity, hipStream_t s) {
  const int n = int(h->h_items.size());
  if (n == 0) return hipSuccess;
  const int64_t total = h->h_cell0.back();
  const int to_ghost = parity == 1;  // after an even step
  if (h->q == 19) {
    switch (h->elem) {
      case 2: return H::launch_q<19, uint16_t>(h->d_tab, h->d_items, h->d_cell0, n, total, to_ghost, s);
      case 8: return H::launch_q<19, double>(h->d_tab, h->d_items, h->d_cell0, n, total, to_ghost, s);
      default: return H::launch_q<19, float>(h->d_tab, h->d_items, h->d_cell0, n, total, to_ghost, s);
    }
  }
  switch (h->elem) {
    case 2: return H::launch_q<9, uint16_t>(h->d_tab, h->d_items, h->d_cell0, n, total, to_ghost, s);
    case 8: return H::launch_q<9, double>(h->d_tab, h->d_items, h->d_cell0, n, total, to_ghost, s);
    default: return H::launch_q<9, float>(h->d_tab, h->d_items, h->d_cell0, n, total, to_ghost, s);
  }
}

// Re-uploads the blocks' kernel arguments when a block changed.
int halo_refresh(flb_halo_s* h) {
  bool dirty = false;
  for (size_t k = 0; k < h->blocks.size(); ++k) dirty |= h->blocks[k]->version != h->versions[k];
  if (!dirty) return FLB_OK;
  for (size_t k = 0; k < h->blocks.size(); ++k) {
    h->h_tab[k] = make_args(h->blocks[k]);
    h->versions[k] = h->blocks[k]->version;
  }
  HIP_TRY(hipMemcpyAsync(h->d_tab, h->h_tab.data(), h->h_tab.size() * sizeof(KArgs), hipMemcpyHostToDevice,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return FLB_OK;
}

int halo_parity(const flb_halo_s* h, int* parity) {
  const int p = h->blocks[0]->parity;
  for (size_t k = 0; k < h->blocks.size(); ++k) {
    if (h->blocks[k]->pending) return fail(FLB_ERR_STATE, "halo block %zu: a Stream is pending", k);
    if (h->blocks[k]->parity != p)
      return fail(FLB_ERR_STATE, "halo block %zu is at parity %d, block 0 at %d", k, h->blocks[k]->parity, p);
  }
  *parity = p;
  return FLB_OK;
}

}  // namespace flbh

extern "C" {

int flb_halo_create(const flb_block* blocks, int n, const int* origins, const int* periodic,
                    const int* extent, void* stream, flb_halo* out) {
  if (!blocks || !origins || !out || n < 1) return fail(FLB_ERR_ARG, "halo needs blocks, origins and an output");
  *out = nullptr;
  const flb_block_s* b0 = blocks[0];
  if (b0 && b0->desc.q != FLB_D2Q9 && b0->desc.q != FLB_D3Q19)
    return fail(FLB_ERR_UNSUPPORTED, "block halos take D2Q9 / D3Q19 blocks");
  for (int k = 0; k < n; ++k) {
    const flb_block_s* b = blocks[k];
    if (!b) return fail(FLB_ERR_ARG, "null block %d", k);
    if (b->desc.q != b0->desc.q || b->elem != b0->elem || b->device != b0->device)
      return fail(FLB_ERR_ARG, "block %d differs from block 0 in lattice, storage size or device", k);
    if (b->desc.periodic[0] || b->desc.periodic[1] || b->desc.periodic[2])
      return fail(FLB_ERR_ARG, "block %d is periodic itself; periodic domains wrap through the halo", k);
    if (b->stream != static_cast<hipStream_t>(stream))
      return fail(FLB_ERR_ARG, "block %d runs on another stream than the halo", k);
  }
  const int d = b0->d;
  int per[3] = {0, 0, 0}, ext[3] = {0, 0, 0};
  for (int a = 0; a < d; ++a) {
    per[a] = periodic ? periodic[a] != 0 : 0;
    ext[a] = extent ? extent[a] : 0;
    if (per[a] && ext[a] <= 0) return fail(FLB_ERR_ARG, "periodic axis %d needs the domain extent", a);
  }
  auto lo = [&](const flb_block_s* b, int a) { return a == 0 ? b->desc.overlap : (a == 1 ? b->desc.overlap : (d == 3 ? b->desc.overlap : 0)); };
  auto full = [&](const flb_block_s* b, int a) { return a == 0 ? b->desc.nx : (a == 1 ? b->desc.ny : b->desc.nz); };
  std::vector<H::Item> items;
  const int imgs = d == 3 ? 27 : 9;
  for (int ia = 0; ia < n; ++ia) {
    for (int ib = 0; ib < n; ++ib) {
      for (int im = 0; im < imgs; ++im) {
        int s[3] = {im % 3 - 1, (im / 3) % 3 - 1, d == 3 ? im / 9 - 1 : 0};
        bool skip = false;
        for (int a = 0; a < 3; ++a)
          if (s[a] != 0 && (a >= d || !per[a])) skip = true;
        if (skip || (ia == ib && s[0] == 0 && s[1] == 0 && s[2] == 0)) continue;
        const flb_block_s* A = blocks[ia];
        const flb_block_s* B = blocks[ib];
        int g0[3], g1[3];
        bool empty = false;
        for (int a = 0; a < 3; ++a) {
          if (a >= d) {
            g0[a] = 0;
            g1[a] = 1;
            continue;
          }
          // B's updated cells in global coordinates, shifted by the image
          const int b0g = origins[3 * ib + a] + lo(B, a) + s[a] * ext[a];
          const int b1g = origins[3 * ib + a] + full(B, a) - lo(B, a) + s[a] * ext[a];
          // A's whole box (updated cells and overlap layer)
          const int a0g = origins[3 * ia + a], a1g = origins[3 * ia + a] + full(A, a);
          g0[a] = std::max(a0g, b0g);
          g1[a] = std::min(a1g, b1g);
          if (g1[a] <= g0[a]) empty = true;
        }
        if (empty) continue;
        // the mirrored cells must lie in A's overlap layer, not among its updated cells
        bool inner = true;
        for (int a = 0; a < d; ++a) {
          const int u0 = origins[3 * ia + a] + lo(A, a), u1 = origins[3 * ia + a] + full(A, a) - lo(A, a);
          if (g1[a] <= u0 || g0[a] >= u1) inner = false;
        }
        if (inner)
          return fail(FLB_ERR_LAYOUT, "updated cells of blocks %d and %d overlap", ia, ib);
        H::Item it{};
        it.a = ia;
        it.b = ib;
        it.gx = g0[0] - origins[3 * ia + 0];
        it.gy = g0[1] - origins[3 * ia + 1];
        it.gz = d == 3 ? g0[2] - origins[3 * ia + 2] : 0;
        it.nx = g1[0] - g0[0];
        it.ny = g1[1] - g0[1];
        it.nz = d == 3 ? g1[2] - g0[2] : 1;
        it.sx = origins[3 * ia + 0] - origins[3 * ib + 0] - s[0] * ext[0];
        it.sy = origins[3 * ia + 1] - origins[3 * ib + 1] - s[1] * ext[1];
        it.sz = d == 3 ? origins[3 * ia + 2] - origins[3 * ib + 2] - s[2] * ext[2] : 0;
        items.push_back(it);
      }
    }
  }
  DeviceGuard g(b0->device);
  flb_halo_s* h = new flb_halo_s();
  h->device = b0->device;
  h->stream = static_cast<hipStream_t>(stream);
  h->blocks.assign(blocks, blocks + n);
  h->versions.assign(size_t(n), ~uint64_t(0));
  h->q = b0->desc.q;
  h->elem = b0->elem;
  h->h_tab.resize(size_t(n));
  h->h_items = items;
  h->h_cell0.assign(1, 0);
  for (const H::Item& it : items) h->h_cell0.push_back(h->h_cell0.back() + int64_t(it.nx) * it.ny * it.nz);
  int rc = FLB_OK;
  const size_t ni = std::max<size_t>(items.size(), 1);
  if (hipMalloc(reinterpret_cast<void**>(&h->d_tab), size_t(n) * sizeof(KArgs)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&h->d_items), ni * sizeof(H::Item)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&h->d_cell0), (ni + 1) * sizeof(int64_t)) != hipSuccess)
    rc = fail(FLB_ERR_NOMEM, "halo table allocation failed");
  if (rc == FLB_OK && !items.empty() &&
      (hipMemcpyAsync(h->d_items, items.data(), items.size() * sizeof(H::Item), hipMemcpyHostToDevice,
                      h->stream) != hipSuccess ||
       hipMemcpyAsync(h->d_cell0, h->h_cell0.data(), h->h_cell0.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice, h->stream) != hipSuccess))
    rc = fail(FLB_ERR_HIP, "halo table upload failed");
  if (rc == FLB_OK) rc = halo_refresh(h);
  if (rc != FLB_OK) {
    flb_halo_destroy(h);
    return rc;
  }
  *out = h;
  return FLB_OK;
}

int flb_halo_destroy(flb_halo h) {
  if (!h) return FLB_OK;
  DeviceGuard g(h->device);
  (void)hipStreamSynchronize(h->stream);
  if (h->d_tab) (void)hipFree(h->d_tab);
  if (h->d_items) (void)hipFree(h->d_items);
  if (h->d_cell0) (void)hipFree(h->d_cell0);
  delete h;
  return FLB_OK;
}

int flb_halo_info(flb_halo h, int* n_items, int64_t* n_ghost_cells) {
  if (!h) return fail(FLB_ERR_ARG, "null halo handle");
  if (n_items) *n_items = int(h->h_items.size());
  if (n_ghost_cells) *n_ghost_cells = h->h_cell0.back();
  return FLB_OK;
}

int flb_halo_exchange(flb_halo h) {
  if (!h) return fail(FLB_ERR_ARG, "null halo handle");
  DeviceGuard g(h->device);
  int parity = 0;
  int rc = halo_parity(h, &parity);
  if (rc) return rc;
  if ((rc = halo_refresh(h))) return rc;
  HIP_TRY(halo_launch(h, parity, h->stream));
  return FLB_OK;
}

}  // extern "C"
