// Host-side internals shared by the C-ABI translation units: the block
// object behind flb_block and the helpers that build kernel arguments.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/flb_abi.h"
#include "flb_internal.hpp"

struct flb_block_s {
  int device = 0;
  hipStream_t stream = nullptr;
  flb_block_desc desc{};
  int d = 3;
  int64_t N = 0;        // host cells nx*ny*nz
  int64_t stride = 0;   // slots between two per-cell arrays (pads included)
  int pitch = 0;        // slots per row
  int64_t plane = 0;    // slots per plane
  int64_t dpop = 0, zpop = 0;  // population strides (see flb_internal.hpp)
  int ypop = 0;
  int64_t pop_elems = 0;       // elements allocated for the populations
  size_t elem = 4;   // bytes per stored population
  size_t real = 4;   // bytes of the compute type
  void* d_pops = nullptr;
  uint8_t* d_flags = nullptr;
  void* d_rho = nullptr;
  void* d_u = nullptr;
  double* d_partials = nullptr;
  double* d_red = nullptr;
  void* d_stage = nullptr;
  size_t stage_bytes = 0;
  std::vector<uint8_t> h_flags;
  bool flags_set = false;
  flb_task tasks[flb::kMaxTasks]{};
  int ntasks = 0;
  flb_hwbc bcs[flb::kMaxBc]{};
  int nbc = 0;
  double rates[19]{};
  double force[3]{};
  int64_t step = 0;
  int parity = 0;
  int pending = 0;  // 1: cell dynamics applied, stream not yet called
  // refinement level of the block and the hierarchy's finest level: the
  // count-gated calls step the block when count % 2^(max_level - level) == 0
  int level = 0, max_level = 0;
  int watchdog = 0;  // flb_step checks mass / momentum every `watchdog` steps (0: off)
  int vec = 0;  // step kernel form: 0 automatic cell form, -1 row segments, -3 linear runs,
                // 1/2/4 row form with that many cells per thread
  int wpb = 4;   // wavefronts per workgroup
  uint64_t version = 0;  // bumped by every change a batch table must see
  uint64_t* d_links = nullptr;  // precomputed halfway wall links (slot layout)
  bool links_dirty = true;      // flags or halfway managers changed since d_links was built
  bool ghost_overlap = false;   // the overlap layer mirrors neighbouring blocks (flb_halo)
  void* d_rho_old = nullptr;    // RhoOld / UOld of the convergence residual (field slot layout)
  void* d_u_old = nullptr;
  unsigned long long* d_res = nullptr;  // residual as the bits of a non-negative double
  // Coupled lattices (flb_share_velocity, flb_add_buoyancy). A scalar block
  // advects with vel_src's VELOCITY field; a flow block's DYN_BGK_SCALAR_FORCE
  // cells take the buoyancy of each entry of buoy. users lists the blocks that
  // hold a reference to this one, so destroying either side detaches it.
  struct Buoyancy {
    flb_block src;
    uint8_t mask;
    double gbeta, rho0;
  };
  flb_block vel_src = nullptr;
  std::vector<Buoyancy> buoy;
  std::vector<flb_block> users;
};

struct flb_halo_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<flb_block> blocks;
  std::vector<uint64_t> versions;
  int q = 19;
  size_t elem = 4;
  std::vector<flb::KArgs> h_tab;
  std::vector<flb::HaloItem> h_items;
  std::vector<int64_t> h_cell0;
  flb::KArgs* d_tab = nullptr;
  flb::HaloItem* d_items = nullptr;
  int64_t* d_cell0 = nullptr;
};

namespace flbh {

using flb::KArgs;

// Records the message for flb_last_error() (thread-local) and returns code.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(FLB_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

// A roctx range over a host call (its kernels and waits show under it in
// rocprofv3 traces with --marker-trace; a no-op without a profiler).
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) == hipSuccess && prev != dev) {
      (void)hipSetDevice(dev);
    } else {
      prev = -1;
    }
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

KArgs make_args(const flb_block_s* b);
uint32_t buffer_form_ok(const KArgs& a, int elem, int dims);
// True when a count-gated call acts on a block of `level` in a hierarchy whose
// finest level is max_level (BlockLatticeManager's 2^(maxlevel - level) rule).
inline bool level_due(int64_t count, int level, int max_level) {
  return count % (int64_t(1) << (max_level - level)) == 0;
}
// Builds the block's wall-link words when flags or halfway managers changed
// (enqueued on the block's stream); a no-op without halfway managers.
int ensure_links(flb_block_s* b);

// Block halo (flb_halo.hip): enqueue the exchange that follows a step leaving
// the blocks at `parity`; refresh the device copy of the blocks' arguments;
// the common parity of the blocks.
hipError_t halo_launch(const flb_halo_s* h, int parity, hipStream_t s);
int halo_refresh(flb_halo_s* h);
int halo_parity(const flb_halo_s* h, int* parity);
// True when rows of ex cells fill less than 90 % of their wavefronts' lanes.
bool linear_form(int ex);
int family_of(const flb_block_s* b);

}  // namespace flbh

#define CHECK_BLOCK(b) \
  if (!(b)) return fail(FLB_ERR_ARG, "null block handle")
