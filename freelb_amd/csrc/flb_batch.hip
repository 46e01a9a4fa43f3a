// C-ABI batch layer: many blocks stepped by one launch per time step.
//
// A batch keeps a device table of the blocks' kernel arguments and the
// prefix counts of their wavefronts; k_step_batch walks that table. The
// blocks stay ordinary flb_block objects (their own allocations, flags,
// tasks, fields, read-out); the batch only replaces the per-block launch loop
// of FreeLB's BlockLatticeManager::ApplyCellDynamics / CuDevApplyCellDynamics
// (data_struct/block_lattice.hh:744-778, 919-936) with a single launch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/flb_abi.h"
#include "flb_host.hpp"
#include "flb_internal.hpp"

using flb::KArgs;
using namespace flbh;

struct flb_batch_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<flb_block> blocks;
  std::vector<uint64_t> versions;
  std::vector<hipStream_t> other_streams;  // distinct block streams other than `stream`
  std::vector<hipEvent_t> events;          // one per other stream, plus one for the batch
  int q = 19;
  int storage = FLB_STORE_F32;
  int family = flb::FAM_BGK;
  bool per = false;
  std::vector<KArgs> h_tab;
  std::vector<int64_t> h_wave0;
  KArgs* d_tab = nullptr;
  int64_t* d_wave0 = nullptr;
  std::vector<int> h_wblk;     // block of every wavefront
  int* d_wblk = nullptr;
  size_t wblk_cap = 0;
  int64_t cells = 0;
  int wpb = 4;
  int grid = 0;  // 0: one workgroup per group of wpb wavefronts
  flb_halo halo = nullptr;  // exchanged after every step when set
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  int graph_parity = 0;
  int graph_family = 0;
  bool graph_per = false;
};

namespace {

hipError_t launch_batch(const flb_batch_s* t, int parity, hipStream_t s) {
  const int64_t groups = (t->h_wave0.back() + t->wpb - 1) / t->wpb;
  int grid = t->grid;
  if (grid <= 0 || grid > groups) grid = int(std::min<int64_t>(groups, int64_t(1) << 30));
  const int nb = int(t->blocks.size());
  switch (t->storage) {
    case FLB_STORE_F32:
      return flb::launch_batch_f32(t->q, parity, t->family, t->per, t->d_tab, t->d_wave0, t->d_wblk, nb, groups, t->wpb, grid, s);
    case FLB_STORE_F16S:
      return flb::launch_batch_f16s(t->q, parity, t->family, t->per, t->d_tab, t->d_wave0, t->d_wblk, nb, groups, t->wpb, grid, s);
    case FLB_STORE_BF16S:
      return flb::launch_batch_bf16s(t->q, parity, t->family, t->per, t->d_tab, t->d_wave0, t->d_wblk, nb, groups, t->wpb, grid, s);
    case FLB_STORE_F16C:
      return flb::launch_batch_f16c(t->q, parity, t->family, t->per, t->d_tab, t->d_wave0, t->d_wblk, nb, groups, t->wpb, grid, s);
    default:
      return flb::launch_batch_f64(t->q, parity, t->family, t->per, t->d_tab, t->d_wave0, t->d_wblk, nb, groups, t->wpb, grid, s);
  }
}

// The kernel arguments of block b in the batch's cell form (row segments, or
// linear runs for short rows, whatever form the block itself would launch).
KArgs batch_args(const flb_block_s* b) {
  KArgs a = make_args(b);
  a.groups = a.ex;
  a.wpr = (a.ex + flb::kWave - 1) / flb::kWave;
  a.vec = linear_form(a.ex) ? -3 : 0;
  a.buf = buffer_form_ok(a, int(b->elem), b->d);
  return a;
}

// Wavefronts of one block in the batch kernel.
int64_t waves_of(const KArgs& a) {
  if (a.vec == -3) return (int64_t(a.ex) * a.ey * a.ez + flb::kWave - 1) / flb::kWave;
  return int64_t(a.wpr) * a.ey * a.ez;
}

// Common state of the blocks: same parity, nothing pending, flags present.
int check_state(const flb_batch_s* t, int* parity) {
  const int p = t->blocks[0]->parity;
  for (size_t k = 0; k < t->blocks.size(); ++k) {
    const flb_block_s* b = t->blocks[k];
    if (!b->flags_set) return fail(FLB_ERR_STATE, "batch block %zu: flags not uploaded", k);
    if (b->pending) return fail(FLB_ERR_STATE, "batch block %zu: a Stream is pending", k);
    if (b->parity != p)
      return fail(FLB_ERR_STATE, "batch block %zu is at parity %d, block 0 at %d: step them to a common parity",
                  k, b->parity, p);
  }
  *parity = p;
  return FLB_OK;
}

// Rebuilds the device table when a block's settings changed since the last build.
int refresh(flb_batch_s* t, bool force) {
  bool dirty = force;
  for (size_t k = 0; k < t->blocks.size(); ++k) dirty |= t->blocks[k]->version != t->versions[k];
  if (!dirty) return FLB_OK;
  t->family = flb::FAM_BGK;
  t->per = false;
  t->h_wave0.assign(1, 0);
  t->cells = 0;
  for (size_t k = 0; k < t->blocks.size(); ++k) {
    flb_block_s* b = t->blocks[k];
    const int rc = ensure_links(b);
    if (rc) return rc;
    t->h_tab[k] = batch_args(b);
    const KArgs& a = t->h_tab[k];
    t->h_wave0.push_back(t->h_wave0.back() + waves_of(a));
    t->cells += int64_t(a.ex) * a.ey * a.ez;
    t->family = std::max(t->family, family_of(b));
    t->per = t->per || a.px || a.py || a.pz;
    t->versions[k] = b->version;
  }
  const size_t waves = size_t(t->h_wave0.back());
  t->h_wblk.resize(waves);
  for (size_t k = 0; k < t->blocks.size(); ++k)
    std::fill(t->h_wblk.begin() + t->h_wave0[k], t->h_wblk.begin() + t->h_wave0[k + 1], int(k));
  if (t->wblk_cap < waves) {
    if (t->d_wblk) HIP_TRY(hipFree(t->d_wblk));
    t->d_wblk = nullptr;
    t->wblk_cap = 0;
    if (hipMalloc(reinterpret_cast<void**>(&t->d_wblk), std::max<size_t>(waves, 1) * sizeof(int)) != hipSuccess)
      return fail(FLB_ERR_NOMEM, "batch wave table allocation failed");
    t->wblk_cap = waves;
  }
  HIP_TRY(hipMemcpyAsync(t->d_tab, t->h_tab.data(), t->h_tab.size() * sizeof(KArgs), hipMemcpyHostToDevice,
                         t->stream));
  HIP_TRY(hipMemcpyAsync(t->d_wave0, t->h_wave0.data(), t->h_wave0.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, t->stream));
  HIP_TRY(hipMemcpyAsync(t->d_wblk, t->h_wblk.data(), waves * sizeof(int), hipMemcpyHostToDevice, t->stream));
  // the host vectors are rebuilt only here, after the copies have completed
  HIP_TRY(hipStreamSynchronize(t->stream));
  return FLB_OK;
}

// Work the blocks' own streams enqueued before now runs before the batch.
int join_in(flb_batch_s* t) {
  for (size_t k = 0; k < t->other_streams.size(); ++k) {
    HIP_TRY(hipEventRecord(t->events[k], t->other_streams[k]));
    HIP_TRY(hipStreamWaitEvent(t->stream, t->events[k], 0));
  }
  return FLB_OK;
}

// Work enqueued later on the blocks' own streams runs after the batch.
int join_out(flb_batch_s* t) {
  if (t->other_streams.empty()) return FLB_OK;
  hipEvent_t done = t->events.back();
  HIP_TRY(hipEventRecord(done, t->stream));
  for (hipStream_t s : t->other_streams) HIP_TRY(hipStreamWaitEvent(s, done, 0));
  return FLB_OK;
}

void advance(flb_batch_s* t, int n) {
  for (flb_block_s* b : t->blocks) {
    b->parity ^= (n & 1);
    b->step += n;
  }
}

void drop_graph(flb_batch_s* t) {
  if (t->exec) (void)hipGraphExecDestroy(t->exec);
  if (t->graph) (void)hipGraphDestroy(t->graph);
  t->exec = nullptr;
  t->graph = nullptr;
  t->graph_steps = 0;
}

#define CHECK_BATCH(t) \
  if (!(t)) return fail(FLB_ERR_ARG, "null batch handle")

}  // namespace

extern "C" {

int flb_batch_create(const flb_block* blocks, int n, void* stream, flb_batch* out) {
  if (!blocks || !out || n < 1) return fail(FLB_ERR_ARG, "batch needs a non-empty block list and an output");
  *out = nullptr;
  const flb_block_s* b0 = blocks[0];
  if (!b0) return fail(FLB_ERR_ARG, "null block 0");
  if (b0->desc.q != FLB_D2Q9 && b0->desc.q != FLB_D3Q19)
    return fail(FLB_ERR_UNSUPPORTED, "batches take D2Q9 / D3Q19 blocks");
  for (int k = 0; k < n; ++k) {
    const flb_block_s* b = blocks[k];
    if (!b) return fail(FLB_ERR_ARG, "null block %d", k);
    if (b->desc.q != b0->desc.q || b->desc.storage != b0->desc.storage || b->device != b0->device)
      return fail(FLB_ERR_ARG, "block %d differs from block 0 in lattice, storage kind or device", k);
    for (int j = 0; j < k; ++j)
      if (blocks[j] == b) return fail(FLB_ERR_ARG, "block %d appears twice", k);
  }
  DeviceGuard g(b0->device);
  flb_batch_s* t = new flb_batch_s();
  t->device = b0->device;
  t->stream = static_cast<hipStream_t>(stream);
  t->q = b0->desc.q;
  t->storage = b0->desc.storage;
  t->blocks.assign(blocks, blocks + n);
  t->versions.assign(size_t(n), ~uint64_t(0));
  t->h_tab.resize(size_t(n));
  for (int k = 0; k < n; ++k) {
    hipStream_t s = blocks[k]->stream;
    if (s != t->stream && std::find(t->other_streams.begin(), t->other_streams.end(), s) == t->other_streams.end())
      t->other_streams.push_back(s);
  }
  int rc = FLB_OK;
  t->events.resize(t->other_streams.size() + 1, nullptr);
  for (auto& e : t->events)
    if (rc == FLB_OK && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      rc = fail(FLB_ERR_HIP, "event creation failed");
  if (rc == FLB_OK && (hipMalloc(reinterpret_cast<void**>(&t->d_tab), size_t(n) * sizeof(KArgs)) != hipSuccess ||
                       hipMalloc(reinterpret_cast<void**>(&t->d_wave0), size_t(n + 1) * sizeof(int64_t)) != hipSuccess))
    rc = fail(FLB_ERR_NOMEM, "batch table allocation failed");
  if (rc == FLB_OK) rc = refresh(t, true);
  if (rc != FLB_OK) {
    flb_batch_destroy(t);
    return rc;
  }
  *out = t;
  return FLB_OK;
}

int flb_batch_destroy(flb_batch t) {
  if (!t) return FLB_OK;
  DeviceGuard g(t->device);
  (void)hipStreamSynchronize(t->stream);
  drop_graph(t);
  for (hipEvent_t e : t->events)
    if (e) (void)hipEventDestroy(e);
  if (t->d_tab) (void)hipFree(t->d_tab);
  if (t->d_wave0) (void)hipFree(t->d_wave0);
  if (t->d_wblk) (void)hipFree(t->d_wblk);
  delete t;
  return FLB_OK;
}

int flb_batch_info(flb_batch t, int* n_blocks, int64_t* n_cells, flb_launch_cfg* cfg) {
  CHECK_BATCH(t);
  if (n_blocks) *n_blocks = int(t->blocks.size());
  if (n_cells) *n_cells = t->cells;
  if (cfg) {
    cfg->waves_per_group = t->wpb;
    cfg->grid = t->grid;
  }
  return FLB_OK;
}

int flb_batch_set_launch(flb_batch t, const flb_launch_cfg* cfg) {
  CHECK_BATCH(t);
  if (!cfg) return fail(FLB_ERR_ARG, "null launch configuration");
  if (cfg->waves_per_group < 1 || cfg->waves_per_group * flb::kWave > flb::kMaxThreads)
    return fail(FLB_ERR_ARG, "waves_per_group %d outside [1, %d]", cfg->waves_per_group, flb::kMaxThreads / flb::kWave);
  if (cfg->grid < 0) return fail(FLB_ERR_ARG, "negative grid");
  t->wpb = cfg->waves_per_group;
  t->grid = cfg->grid;
  return FLB_OK;
}

int flb_batch_step(flb_batch t, int n) {
  CHECK_BATCH(t);
  if (n < 0) return fail(FLB_ERR_ARG, "negative step count");
  DeviceGuard g(t->device);
  Range range("flb_batch_step");
  int parity = 0;
  int rc = check_state(t, &parity);
  if (rc) return rc;
  if ((rc = refresh(t, false))) return rc;
  if (t->halo && (rc = halo_refresh(t->halo))) return rc;
  if ((rc = join_in(t))) return rc;
  for (int s = 0; s < n; ++s) {
    HIP_TRY(launch_batch(t, (parity + s) & 1, t->stream));
    if (t->halo) HIP_TRY(halo_launch(t->halo, (parity + s + 1) & 1, t->stream));
  }
  advance(t, n);
  return join_out(t);
}

int flb_batch_set_halo(flb_batch t, flb_halo h) {
  CHECK_BATCH(t);
  if (h && h->stream != t->stream) return fail(FLB_ERR_ARG, "halo and batch run on different streams");
  if (h) {
    for (flb_block_s* b : h->blocks)
      if (std::find(t->blocks.begin(), t->blocks.end(), b) == t->blocks.end())
        return fail(FLB_ERR_ARG, "halo block is not part of the batch");
  }
  drop_graph(t);
  t->halo = h;
  return FLB_OK;
}

int flb_batch_capture(flb_batch t, int n) {
  CHECK_BATCH(t);
  if (n < 1) return fail(FLB_ERR_ARG, "capture needs at least one step");
  DeviceGuard g(t->device);
  int parity = 0;
  int rc = check_state(t, &parity);
  if (rc) return rc;
  if ((rc = refresh(t, false))) return rc;
  drop_graph(t);
  // the halo's tables are refreshed (host copies and a stream sync) before
  // the capture opens, as flb_batch_step does before its launches
  if (t->halo && (rc = halo_refresh(t->halo))) return rc;
  hipStream_t cs = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return fail(FLB_ERR_HIP, "graph capture failed to begin: %s", hipGetErrorString(e));
  }
  for (int s = 0; e == hipSuccess && s < n; ++s) {
    e = launch_batch(t, (parity + s) & 1, cs);
    if (e == hipSuccess && t->halo) e = halo_launch(t->halo, (parity + s + 1) & 1, cs);
  }
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(cs, &graph);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(FLB_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  }
  t->graph = graph;
  if (hipGraphInstantiate(&t->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    drop_graph(t);
    return fail(FLB_ERR_HIP, "graph instantiation failed");
  }
  t->graph_steps = n;
  t->graph_parity = parity;
  t->graph_family = t->family;
  t->graph_per = t->per;
  return FLB_OK;
}

int flb_batch_replay(flb_batch t) {
  CHECK_BATCH(t);
  if (!t->exec) return fail(FLB_ERR_STATE, "no captured graph");
  DeviceGuard g(t->device);
  Range range("flb_batch_replay");
  int parity = 0;
  int rc = check_state(t, &parity);
  if (rc) return rc;
  if (parity != t->graph_parity)
    return fail(FLB_ERR_STATE, "graph was captured at parity %d, blocks are at %d (capture an even step count to replay repeatedly)",
                t->graph_parity, parity);
  if ((rc = refresh(t, false))) return rc;
  if (t->halo && (rc = halo_refresh(t->halo))) return rc;
  if (t->family != t->graph_family || t->per != t->graph_per)
    return fail(FLB_ERR_STATE, "block dynamics changed kernel family since capture: capture again");
  if ((rc = join_in(t))) return rc;
  HIP_TRY(hipGraphLaunch(t->exec, t->stream));
  advance(t, t->graph_steps);
  return join_out(t);
}

int flb_batch_synchronize(flb_batch t) {
  CHECK_BATCH(t);
  DeviceGuard g(t->device);
  HIP_TRY(hipStreamSynchronize(t->stream));
  return FLB_OK;
}

int flb_batch_autotune(flb_batch t, int steps, flb_launch_cfg* best, double* best_ms) {
  CHECK_BATCH(t);
  if (steps < 2 || (steps & 1)) return fail(FLB_ERR_ARG, "autotune needs an even step count >= 2");
  DeviceGuard g(t->device);
  int parity = 0;
  int rc = check_state(t, &parity);
  if (rc) return rc;
  if ((rc = refresh(t, false))) return rc;
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, t->device));
  // snapshot of every block's populations and of the RHO / VELOCITY fields
  // that field-writing dynamics update each step, restored after the trials
  struct Region {
    void* p;
    size_t bytes;
  };
  std::vector<Region> regions;
  for (flb_block_s* b : t->blocks) {
    regions.push_back({b->d_pops, size_t(b->pop_elems) * b->elem});
    if (b->d_rho) regions.push_back({b->d_rho, size_t(b->stride) * b->real});
    if (b->d_u) regions.push_back({b->d_u, size_t(b->stride) * b->real * size_t(b->d)});
  }
  size_t total = 0;
  for (const Region& r : regions) total += r.bytes;
  char* snap = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&snap), total) != hipSuccess)
    return fail(FLB_ERR_NOMEM, "autotune snapshot of %zu bytes failed", total);
  if ((rc = join_in(t))) {
    (void)hipFree(snap);
    return rc;
  }
  size_t off = 0;
  for (const Region& r : regions) {
    if (hipMemcpyAsync(snap + off, r.p, r.bytes, hipMemcpyDeviceToDevice, t->stream) != hipSuccess)
      rc = fail(FLB_ERR_HIP, "autotune snapshot copy failed");
    off += r.bytes;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == FLB_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(FLB_ERR_HIP, "event creation failed");
  const int saved_wpb = t->wpb, saved_grid = t->grid;
  double best_t = 1e300;
  flb_launch_cfg chosen{saved_wpb, saved_grid};
  const int wpbs[] = {1, 2, 4};
  const int mults[] = {0, 1, 2, 4, 8};
  for (int wpb : wpbs) {
    for (int m : mults) {
      if (rc != FLB_OK) break;
      t->wpb = wpb;
      t->grid = m == 0 ? 0 : cus * m * (4 / wpb);
      hipError_t e = hipSuccess;
      for (int s = 0; s < 2 + steps && e == hipSuccess; ++s) {
        if (s == 2) e = hipEventRecord(e0, t->stream);
        if (e == hipSuccess) e = launch_batch(t, (parity + s) & 1, t->stream);
        if (e == hipSuccess && t->halo) e = halo_launch(t->halo, (parity + s + 1) & 1, t->stream);
      }
      if (e == hipSuccess) e = hipEventRecord(e1, t->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(FLB_ERR_HIP, "autotune trial failed: %s", hipGetErrorString(e));
        break;
      }
      const double per_step = double(ms) / steps;
      if (per_step < best_t) {
        best_t = per_step;
        chosen = flb_launch_cfg{t->wpb, t->grid};
      }
    }
  }
  off = 0;
  for (const Region& r : regions) {
    if (hipMemcpyAsync(r.p, snap + off, r.bytes, hipMemcpyDeviceToDevice, t->stream) != hipSuccess && rc == FLB_OK)
      rc = fail(FLB_ERR_HIP, "autotune restore copy failed");
    off += r.bytes;
  }
  if (hipStreamSynchronize(t->stream) != hipSuccess && rc == FLB_OK) rc = fail(FLB_ERR_HIP, "autotune sync failed");
  (void)hipFree(snap);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc != FLB_OK) {
    t->wpb = saved_wpb;
    t->grid = saved_grid;
    return rc;
  }
  t->wpb = chosen.waves_per_group;
  t->grid = chosen.grid;
  if (best) *best = chosen;
  if (best_ms) *best_ms = best_t;
  return join_out(t);
}

}  // extern "C"
