// C-ABI host layer: block handles, validation, staging and launches.
// Declarations and reference citations are in include/flb_abi.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flb_abi.h"
#include "flb_cell.hpp"
#include "flb_internal.hpp"
#include "flb_host.hpp"

using flb::KArgs;

namespace flbh {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

bool storage_ok(int s) {
  return s == FLB_STORE_F32 || s == FLB_STORE_F16S || s == FLB_STORE_BF16S || s == FLB_STORE_F64 ||
         s == FLB_STORE_F16C;
}

hipError_t launch_step(const flb_block_s* b, int family, const KArgs& a) {
  switch (b->desc.storage) {
    case FLB_STORE_F32: return flb::launch_step_f32(b->desc.q, b->parity, family, a, b->stream);
    case FLB_STORE_F16S: return flb::launch_step_f16s(b->desc.q, b->parity, family, a, b->stream);
    case FLB_STORE_BF16S: return flb::launch_step_bf16s(b->desc.q, b->parity, family, a, b->stream);
    case FLB_STORE_F16C: return flb::launch_step_f16c(b->desc.q, b->parity, family, a, b->stream);
    default: return flb::launch_step_f64(b->desc.q, b->parity, family, a, b->stream);
  }
}

hipError_t launch_gather(const flb_block_s* b, const KArgs& a, void* out) {
  switch (b->desc.storage) {
    case FLB_STORE_F32: return flb::launch_gather_f32(b->desc.q, b->parity, a, out, b->stream);
    case FLB_STORE_F16S: return flb::launch_gather_f16s(b->desc.q, b->parity, a, out, b->stream);
    case FLB_STORE_BF16S: return flb::launch_gather_bf16s(b->desc.q, b->parity, a, out, b->stream);
    case FLB_STORE_F16C: return flb::launch_gather_f16c(b->desc.q, b->parity, a, out, b->stream);
    default: return flb::launch_gather_f64(b->desc.q, b->parity, a, out, b->stream);
  }
}

hipError_t launch_scatter(const flb_block_s* b, const KArgs& a, const void* in) {
  switch (b->desc.storage) {
    case FLB_STORE_F32: return flb::launch_scatter_f32(b->desc.q, a, in, b->stream);
    case FLB_STORE_F16S: return flb::launch_scatter_f16s(b->desc.q, a, in, b->stream);
    case FLB_STORE_BF16S: return flb::launch_scatter_bf16s(b->desc.q, a, in, b->stream);
    case FLB_STORE_F16C: return flb::launch_scatter_f16c(b->desc.q, a, in, b->stream);
    default: return flb::launch_scatter_f64(b->desc.q, a, in, b->stream);
  }
}

hipError_t launch_init_feq(const flb_block_s* b, const KArgs& a, const void* rhou) {
  switch (b->desc.storage) {
    case FLB_STORE_F32: return flb::launch_init_feq_f32(b->desc.q, a, rhou, b->stream);
    case FLB_STORE_F16S: return flb::launch_init_feq_f16s(b->desc.q, a, rhou, b->stream);
    case FLB_STORE_BF16S: return flb::launch_init_feq_bf16s(b->desc.q, a, rhou, b->stream);
    case FLB_STORE_F16C: return flb::launch_init_feq_f16c(b->desc.q, a, rhou, b->stream);
    default: return flb::launch_init_feq_f64(b->desc.q, a, rhou, b->stream);
  }
}

hipError_t launch_moments(const flb_block_s* b, const KArgs& a, uint8_t mask, void* out) {
  switch (b->desc.storage) {
    case FLB_STORE_F32: return flb::launch_moments_f32(b->desc.q, b->parity, a, mask, out, b->stream);
    case FLB_STORE_F16S: return flb::launch_moments_f16s(b->desc.q, b->parity, a, mask, out, b->stream);
    case FLB_STORE_BF16S: return flb::launch_moments_bf16s(b->desc.q, b->parity, a, mask, out, b->stream);
    case FLB_STORE_F16C: return flb::launch_moments_f16c(b->desc.q, b->parity, a, mask, out, b->stream);
    default: return flb::launch_moments_f64(b->desc.q, b->parity, a, mask, out, b->stream);
  }
}

// partials: one double4 per workgroup of the per-cell kernels (cell_blocks(N)).
hipError_t launch_reduce(const flb_block_s* b, const KArgs& a, uint8_t mask, double* partials, double* out) {
  switch (b->desc.storage) {
    case FLB_STORE_F32: return flb::launch_reduce_f32(b->desc.q, b->parity, a, mask, partials, out, b->stream);
    case FLB_STORE_F16S: return flb::launch_reduce_f16s(b->desc.q, b->parity, a, mask, partials, out, b->stream);
    case FLB_STORE_BF16S: return flb::launch_reduce_bf16s(b->desc.q, b->parity, a, mask, partials, out, b->stream);
    case FLB_STORE_F16C: return flb::launch_reduce_f16c(b->desc.q, b->parity, a, mask, partials, out, b->stream);
    default: return flb::launch_reduce_f64(b->desc.q, b->parity, a, mask, partials, out, b->stream);
  }
}

void default_rates(flb_block b) {
  if (b->desc.q == 19) {
    for (int j = 0; j < 19; ++j) b->rates[j] = flb::m19_default_rate(j);
  } else {
    // collisionMRT.ur.h:51: {0, 1.1, 1.1, 0, 1.1, 0, 1.1, omega, omega}
    const double s[9] = {0.0, 1.1, 1.1, 0.0, 1.1, 0.0, 1.1, -1.0, -1.0};
    for (int j = 0; j < 9; ++j) b->rates[j] = s[j];
    for (int j = 9; j < 19; ++j) b->rates[j] = 0.0;
  }
}

int axis_lo(const flb_block_s* b, int ax) { return b->desc.periodic[ax] ? 0 : b->desc.overlap; }
int axis_ext(const flb_block_s* b, int ax) {
  const int n = ax == 0 ? b->desc.nx : (ax == 1 ? b->desc.ny : b->desc.nz);
  if (ax == 2 && b->d == 2) return 1;
  return b->desc.periodic[ax] ? n : n - 2 * b->desc.overlap;
}

int family_of(const flb_block_s* b);

// Step kernel form and cells per thread along x. 0 selects the cell form
// (one cell per thread, any periodic axis); 1, 2, 4 the row form, where one
// workgroup of at most 4 wavefronts covers a whole row. The row form is only
// taken without halfway boundaries (its whole-line odd-step stores of a row
// can land on a wall-redirect slot that another row writes in the same
// launch) and with the wide row pitch its stores past the row end need.
int vec_of(const flb_block_s* b) {
  const bool per = b->desc.periodic[0] || b->desc.periodic[1] || b->desc.periodic[2];
  // scalar (D2Q5 / D3Q7) blocks step in the cell form only
  if (per || b->nbc > 0 || b->desc.q == FLB_D2Q5 || b->desc.q == FLB_D3Q7) return 0;
  int v = b->vec;
  if (v <= 0) return 0;
  if (family_of(b) == flb::FAM_FULL && v > 2) v = 2;
  const int ex = b->desc.nx - 2 * b->desc.overlap;
  if (b->pitch < (ex + 255) / 256 * 256 + 64) return 0;
  while (v <= 4 && (ex + flb::kWave * v - 1) / (flb::kWave * v) > flb::kMaxThreads / flb::kWave) v *= 2;
  if (v > 4) return 0;
  return v;
}

// Cell-form wave mapping: row segments when rows fill their wavefronts well,
// else the linear form (wavefronts run across row ends), chosen when less
// than 90 % of the lanes of row-segment wavefronts would hold a cell.
bool linear_form(int ex) {
  const int lanes = (ex + flb::kWave - 1) / flb::kWave * flb::kWave;
  return 10 * ex < 9 * lanes;
}

// Whether the step kernels may address the block's populations through
// wavefront-based buffer descriptors: no periodic axis (the descriptor starts
// one plane and row before the wavefront's first row, which rows y, z >= 1
// keep inside the allocation) and every offset from that start below 2 GiB:
// the lane's row within the wavefront (row segments: 0; linear runs of 64
// cells: at most 64 rows and planes), its x, and the per-direction scalar
// offsets of at most two neighbour ranges.
uint32_t buffer_form_ok(const KArgs& a, int elem, int dims) {
  if (a.px || a.py || a.pz) return 0u;
  if (a.ys < 1 || (dims == 3 && a.zs < 1)) return 0u;
  const uint64_t zp = dims == 3 ? uint64_t(a.zpop) : 0u;
  const uint64_t B = zp + uint64_t(a.ypop) + 1u;
  const uint64_t rows = a.vec == -3 ? 64u * (uint64_t(a.ypop) + zp) : 0u;
  const uint64_t span = rows + uint64_t(a.pitch) + 2u * B + 2u * 18u * uint64_t(a.dpop) + 2u;
  return span * uint64_t(elem) < (uint64_t(1) << 31) ? 1u : 0u;
}

KArgs make_args(const flb_block_s* b) {
  KArgs a{};
  a.pops = static_cast<char*>(b->d_pops) + flb::kFrontPad * b->elem;
  a.pop_total = uint64_t(b->pop_elems) * b->elem;
  a.dpop = b->dpop;
  a.ypop = b->ypop;
  a.zpop = b->zpop;
  a.flags = b->d_flags + flb::kFrontPad;
  a.rho_f = b->d_rho ? static_cast<char*>(b->d_rho) + flb::kFrontPad * b->real : nullptr;
  a.u_f = b->d_u ? static_cast<char*>(b->d_u) + flb::kFrontPad * b->real : nullptr;
  a.fstride = b->stride;
  a.nx = b->desc.nx;
  a.ny = b->desc.ny;
  a.nz = b->desc.nz;
  a.xs = axis_lo(b, 0);
  a.ys = axis_lo(b, 1);
  a.zs = b->d == 3 ? axis_lo(b, 2) : 0;
  a.ex = axis_ext(b, 0);
  a.ey = axis_ext(b, 1);
  a.ez = axis_ext(b, 2);
  a.pitch = b->pitch;
  a.plane = b->plane;
  a.px = b->desc.periodic[0] != 0;
  a.py = b->desc.periodic[1] != 0;
  a.pz = b->d == 3 && b->desc.periodic[2] != 0;
  a.vec = vec_of(b);
  if (a.vec > 0) {
    // row form: one workgroup per row
    a.groups = (a.ex + a.vec - 1) / a.vec;
    a.wpr = (a.groups + flb::kWave - 1) / flb::kWave;
    a.wpb = a.wpr;
    a.nblocks = a.ey * a.ez;
  } else {
    a.groups = a.ex;
    a.wpr = (a.ex + flb::kWave - 1) / flb::kWave;
    a.wpb = b->wpb;
    int64_t waves = int64_t(a.wpr) * a.ey * a.ez;
    if (b->vec == -3 || (b->vec == 0 && linear_form(a.ex))) {
      a.vec = -3;
      waves = (int64_t(a.ex) * a.ey * a.ez + flb::kWave - 1) / flb::kWave;
    }
    a.nblocks = int((waves + a.wpb - 1) / a.wpb);
  }
  a.ntasks = b->ntasks;
  for (int t = 0; t < b->ntasks; ++t) {
    a.tmask[t] = b->tasks[t].mask;
    a.tdyn[t] = b->tasks[t].dynamics;
    a.tcs[t] = b->tasks[t].smagorinsky;
  }
  a.nbc = b->nbc;
  for (int k = 0; k < b->nbc; ++k) {
    a.bcmask[k] = b->bcs[k].cell_mask;
    a.bcvoid[k] = b->bcs[k].void_mask;
    a.bckind[k] = b->bcs[k].kind;
  }
  a.bcany = 0;
  for (int k = 0; k < b->nbc; ++k) a.bcany |= b->bcs[k].cell_mask;
  a.links = b->d_links ? b->d_links + flb::kFrontPad : nullptr;
  if (!a.links) a.bcany = 0;
  a.omega = b->desc.omega;
  for (int j = 0; j < 19; ++j) a.rates[j] = b->rates[j];
  a.mrt_shear = 0u;
  if (b->desc.q == 19) {
    for (int j = 0; j < 19; ++j) {
      const double r = b->rates[j];
      if (!(r >= 0.0)) {
        a.mrt_shear |= 1u << j;
        a.mrt_sn_f[j] = 0.0f;
        a.mrt_sn_d[j] = 0.0;
      } else {
        const double sn = r / double(flb::m19_norm(j));
        a.mrt_sn_f[j] = float(sn);
        a.mrt_sn_d[j] = sn;
      }
    }
  }
  for (int d = 0; d < 3; ++d) a.force[d] = b->force[d];
  const flb_block_s* v = b->vel_src;
  a.u_ext = v && v->d_u ? static_cast<char*>(v->d_u) + flb::kFrontPad * v->real : nullptr;
  a.ncpl = 0;
  for (const auto& c : b->buoy) {
    if (!c.src->d_rho || a.ncpl == flb::kMaxCouplings) continue;
    a.cpl_mask[a.ncpl] = c.mask;
    a.cpl_rho[a.ncpl] = static_cast<char*>(c.src->d_rho) + flb::kFrontPad * c.src->real;
    a.cpl_gbeta[a.ncpl] = c.gbeta;
    a.cpl_rho0[a.ncpl] = c.rho0;
    ++a.ncpl;
  }
  a.buf = buffer_form_ok(a, int(b->elem), b->d);
  return a;
}

// Slot of host cell (x, y, z) in the padded device layout.
int64_t slot_of(const flb_block_s* b, int x, int y, int z) {
  return int64_t(x - axis_lo(b, 0)) + int64_t(b->pitch) * y + b->plane * z;
}

// Host dense array -> padded slot layout (values outside cells stay `fill`).
template <class T, class U>
void to_slots(const flb_block_s* b, const U* dense, T* slots, T fill) {
  for (int64_t k = 0; k < b->stride; ++k) slots[k] = fill;
  const int nx = b->desc.nx, ny = b->desc.ny, nz = b->desc.nz;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        slots[flb::kFrontPad + slot_of(b, x, y, z)] =
            T(dense[size_t(x) + size_t(nx) * (size_t(y) + size_t(ny) * z)]);
}

bool scalar_lattice(const flb_block_s* b) { return b->desc.q == FLB_D2Q5 || b->desc.q == FLB_D3Q7; }

int family_of(const flb_block_s* b) {
  if (scalar_lattice(b)) return flb::FAM_SCALAR;
  for (int t = 0; t < b->ntasks; ++t) {
    const int d = b->tasks[t].dynamics;
    if (d == FLB_DYN_MRT || d == FLB_DYN_SMAG_BGK || d == FLB_DYN_SMAG_MRT || d == FLB_DYN_BGK_FORCE ||
        d == FLB_DYN_BGK_WRITE || d == FLB_DYN_BGK_SCALAR_FORCE)
      return flb::FAM_FULL;
  }
  return flb::FAM_BGK;
}

int upload_field(flb_block b, const double* host, int ncomp, void** dst);

// RHO / VELOCITY fields at FreeLB's initial values (rho 1, u 0) for a block
// that has none yet.
int ensure_fields(flb_block b) {
  int rc = FLB_OK;
  DeviceGuard g(b->device);
  if (!b->d_rho) {
    const std::vector<double> one(size_t(b->N), 1.0);
    if ((rc = upload_field(b, one.data(), 1, &b->d_rho))) return rc;
  }
  if (!b->d_u) {
    const std::vector<double> zero(size_t(b->N) * size_t(b->d), 0.0);
    if ((rc = upload_field(b, zero.data(), b->d, &b->d_u))) return rc;
  }
  return FLB_OK;
}

// Drops every coupling reference between b and other blocks (both directions).
void detach_couplings(flb_block b) {
  auto drop_user = [](flb_block from, flb_block user) {
    auto& u = from->users;
    for (size_t k = 0; k < u.size(); ++k)
      if (u[k] == user) {
        u.erase(u.begin() + std::ptrdiff_t(k));
        return;
      }
  };
  if (b->vel_src) drop_user(b->vel_src, b);
  for (const auto& c : b->buoy) drop_user(c.src, b);
  for (flb_block u : b->users) {
    if (u->vel_src == b) {
      u->vel_src = nullptr;
      u->version++;
    }
    for (size_t k = u->buoy.size(); k-- > 0;)
      if (u->buoy[k].src == b) {
        u->buoy.erase(u->buoy.begin() + std::ptrdiff_t(k));
        u->version++;
      }
  }
  b->vel_src = nullptr;
  b->buoy.clear();
  b->users.clear();
}

// Drops the velocity source of a scalar block.
void detach_couplings_one(flb_block b) {
  if (!b->vel_src) return;
  auto& u = b->vel_src->users;
  for (size_t k = 0; k < u.size(); ++k)
    if (u[k] == b) {
      u.erase(u.begin() + std::ptrdiff_t(k));
      break;
    }
  b->vel_src = nullptr;
  b->version++;
}

// Two blocks of one coupled pair: same cells, same slot layout, same compute
// type and device, and lattices of one dimension (flow D2Q9 / D3Q19 with
// scalar D2Q5 / D3Q7).
int check_pair(flb_block flow, flb_block scal) {
  if (scalar_lattice(flow) || !scalar_lattice(scal))
    return fail(FLB_ERR_ARG, "expected a flow block (D2Q9 / D3Q19) and a scalar block (D2Q5 / D3Q7)");
  if (flow->d != scal->d) return fail(FLB_ERR_ARG, "flow and scalar blocks of different dimension");
  const flb_block_desc &a = flow->desc, &c = scal->desc;
  if (a.nx != c.nx || a.ny != c.ny || a.nz != c.nz || a.overlap != c.overlap || a.periodic[0] != c.periodic[0] ||
      a.periodic[1] != c.periodic[1] || (flow->d == 3 && a.periodic[2] != c.periodic[2]) ||
      flow->pitch != scal->pitch || flow->stride != scal->stride)
    return fail(FLB_ERR_ARG, "flow and scalar blocks cover different cells");
  if (flow->real != scal->real) return fail(FLB_ERR_ARG, "flow and scalar blocks compute in different types");
  if (flow->device != scal->device) return fail(FLB_ERR_ARG, "flow and scalar blocks on different devices");
  // The scalar step reads the flow block's VELOCITY field while the flow step
  // writes it, and the flow step reads the scalar's RHO field: only one
  // stream orders the two blocks' launches.
  if (flow->stream != scal->stream)
    return fail(FLB_ERR_ARG, "coupled flow and scalar blocks must share one stream");
  return FLB_OK;
}

// A non-periodic axis's overlap layer is not visited by the kernels; a cell
// there that matches a task would be updated by the reference (its loop runs
// over every cell), so such a layout is rejected rather than half-served.
int check_layout(const flb_block_s* b) {
  if (!b->flags_set) return FLB_OK;
  const int nx = b->desc.nx, ny = b->desc.ny, nz = b->desc.nz;
  const int ov = b->desc.overlap;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const bool outer = (!b->desc.periodic[0] && (x < ov || x >= nx - ov)) ||
                           (!b->desc.periodic[1] && (y < ov || y >= ny - ov)) ||
                           (b->d == 3 && !b->desc.periodic[2] && (z < ov || z >= nz - ov));
        (void)outer;
        const uint8_t fl = b->h_flags[size_t(x) + size_t(nx) * (size_t(y) + size_t(ny) * z)];
        int dyn = FLB_DYN_NONE, task = -1;
        for (int t = 0; t < b->ntasks; ++t) {
          if (fl & b->tasks[t].mask) {
            dyn = b->tasks[t].dynamics;
            task = t;
            break;
          }
        }
        if (dyn == FLB_DYN_NONE) continue;
        if (outer && !b->ghost_overlap)
          return fail(FLB_ERR_LAYOUT,
                      "cell (%d,%d,%d) in the overlap layer has flag 0x%02x matching task %d",
                      x, y, z, fl, task);
        // a halfway wall neighbour must not be updated itself: its slots take
        // the ordinary writes of the links that bounce off it
        for (int k = 0; k < b->nbc; ++k)
          if (fl & b->bcs[k].void_mask)
            return fail(FLB_ERR_LAYOUT,
                        "cell (%d,%d,%d) with flag 0x%02x is updated by task %d but is a wall "
                        "(void mask 0x%02x) of halfway boundary %d",
                        x, y, z, fl, task, b->bcs[k].void_mask, k);
      }
  return FLB_OK;
}

int ensure_links(flb_block_s* b) {
  if (b->nbc == 0 || !b->flags_set) return FLB_OK;
  if (!b->links_dirty && b->d_links) return FLB_OK;
  if (!b->d_links) {
    if (hipMalloc(reinterpret_cast<void**>(&b->d_links), size_t(b->stride) * sizeof(uint64_t)) != hipSuccess)
      return fail(FLB_ERR_NOMEM, "wall-link allocation failed");
  }
  KArgs a = make_args(b);
  HIP_TRY(flb::launch_links(b->desc.q, a, b->d_links + flb::kFrontPad, b->stream));
  b->links_dirty = false;
  return FLB_OK;
}

int ensure_stage(flb_block b, size_t bytes) {
  if (b->stage_bytes >= bytes) return FLB_OK;
  if (b->d_stage) HIP_TRY(hipFree(b->d_stage));
  b->d_stage = nullptr;
  b->stage_bytes = 0;
  if (hipMalloc(&b->d_stage, bytes) != hipSuccess)
    return fail(FLB_ERR_NOMEM, "staging allocation of %zu bytes failed", bytes);
  b->stage_bytes = bytes;
  return FLB_OK;
}

// Host doubles (ncomp dense components of N cells) -> device compute-type
// arrays in the slot layout, one `stride` apart.
int upload_field(flb_block b, const double* host, int ncomp, void** dst) {
  const size_t bytes = size_t(b->stride) * ncomp * b->real;
  if (!*dst) {
    if (hipMalloc(dst, bytes) != hipSuccess)
      return fail(FLB_ERR_NOMEM, "field allocation of %zu bytes failed", bytes);
  }
  if (b->real == 8) {
    std::vector<double> tmp(size_t(b->stride) * ncomp);
    for (int c = 0; c < ncomp; ++c) to_slots<double, double>(b, host + c * b->N, tmp.data() + c * b->stride, 0.0);
    HIP_TRY(hipMemcpyAsync(*dst, tmp.data(), bytes, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  } else {
    std::vector<float> tmp(size_t(b->stride) * ncomp);
    for (int c = 0; c < ncomp; ++c) to_slots<float, double>(b, host + c * b->N, tmp.data() + c * b->stride, 0.0f);
    HIP_TRY(hipMemcpyAsync(*dst, tmp.data(), bytes, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  }
  return FLB_OK;
}

}  // namespace flbh

using namespace flbh;

extern "C" {

const char* flb_last_error(void) { return g_err.c_str(); }

int flb_abi_version(void) { return FLB_ABI_VERSION; }

int flb_block_layout(const flb_block_desc* desc, flb_layout* out) {
  if (!desc || !out) return fail(FLB_ERR_ARG, "null descriptor or output");
  *out = flb_layout{};
  if (desc->q != FLB_D2Q9 && desc->q != FLB_D3Q19 && desc->q != FLB_D2Q5 && desc->q != FLB_D3Q7)
    return fail(FLB_ERR_ARG, "q must be 9 (D2Q9), 19 (D3Q19), 5 (D2Q5) or 7 (D3Q7), got %d", desc->q);
  if (!storage_ok(desc->storage)) return fail(FLB_ERR_ARG, "unknown storage kind %d", desc->storage);
  const bool scalar = desc->q == FLB_D2Q5 || desc->q == FLB_D3Q7;
  if (scalar && desc->storage != FLB_STORE_F32 && desc->storage != FLB_STORE_F64)
    return fail(FLB_ERR_ARG, "D2Q5 / D3Q7 blocks store FP32 or FP64 populations");
  if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1)
    return fail(FLB_ERR_ARG, "extents must be positive (%d,%d,%d)", desc->nx, desc->ny, desc->nz);
  if ((desc->q == FLB_D2Q9 || desc->q == FLB_D2Q5) && desc->nz != 1)
    return fail(FLB_ERR_ARG, "2-D lattices need nz == 1");
  if (desc->overlap < 1) return fail(FLB_ERR_ARG, "overlap must be >= 1");
  const int d = (desc->q == FLB_D3Q19 || desc->q == FLB_D3Q7) ? 3 : 2;
  const int n[3] = {desc->nx, desc->ny, desc->nz};
  for (int ax = 0; ax < d; ++ax) {
    const int need = desc->periodic[ax] ? 1 : 2 * desc->overlap + 1;
    if (n[ax] < need) return fail(FLB_ERR_ARG, "axis %d extent %d leaves no interior", ax, n[ax]);
  }
  // Slot of cell (x, y, z): (x - xs) + pitch y + plane z, xs the first updated
  // x. Cells x < xs (the left overlap) take the last slots of the previous
  // row, so a row of nx cells needs pitch >= nx for the slots to stay
  // distinct.
  const int xs = desc->periodic[0] ? 0 : desc->overlap;
  const int ex_x = desc->periodic[0] ? desc->nx : desc->nx - 2 * desc->overlap;
  // Short rows (at most 222 updated cells) take a compact pitch: the row's
  // own cells (the overlap cells included; two spare slots on a periodic
  // axis), rounded up to FLB_PITCH_ROUND slots (default 1: no rounding), so
  // that neighbouring rows share cache lines rather than each filling a
  // mostly empty 320-slot row. On 512 cavities of 16..96 cells per edge, 1
  // slot runs 3 % faster than 8 and 22 % faster than 16.
  // Longer rows: the updated cells rounded up to whole 256-slot groups plus
  // one cache line of padding that takes the odd step's whole-line stores
  // past the row ends, so every row's first updated cell sits on a 256-byte
  // boundary.
  int round = 1;
  if (const char* e = std::getenv("FLB_PITCH_ROUND")) round = std::max(1, std::min(256, std::atoi(e)));
  const int compact = desc->periodic[0] ? ex_x + 2 : desc->nx;
  const int pitch = ex_x + 2 <= 224 ? (compact + round - 1) / round * round
                                    : std::max((ex_x + 255) / 256 * 256 + 64, (desc->nx + 63) / 64 * 64);
  const int64_t plane = int64_t(pitch) * desc->ny;
  // tail: one extra row plus the widest vector run past a row end
  const int64_t tail = 2 * int64_t(pitch) + 1024;
  const int64_t stride = (flb::kFrontPad + plane * desc->nz + tail + 255) / 256 * 256;
  if (stride >= (int64_t(1) << 31))
    return fail(FLB_ERR_ARG, "block of %lld padded cells exceeds 2^31", (long long)stride);
  out->pitch = pitch;
  out->xs = xs;
  out->plane = plane;
  out->stride = stride;
  out->front_pad = flb::kFrontPad;
  // population layout: directions interleaved per row (default), per plane,
  // or one array per direction
  const char* lay = std::getenv("FLB_POP_LAYOUT");
  const std::string layout = lay ? lay : "row";
  const int64_t q = desc->q;
  if (layout == "array") {
    out->dpop = stride;
    out->ypop = pitch;
    out->zpop = plane;
    out->pop_elems = q * stride;
  } else if (layout == "plane") {
    out->dpop = plane;
    out->ypop = pitch;
    out->zpop = q * plane;
    out->pop_elems = (flb::kFrontPad + q * plane * desc->nz + tail + 255) / 256 * 256;
  } else {
    out->dpop = pitch;
    out->ypop = q * pitch;
    out->zpop = q * plane;
    out->pop_elems = (flb::kFrontPad + q * plane * desc->nz + tail + 255) / 256 * 256;
  }
  const int64_t elem = desc->storage == FLB_STORE_F64 ? 8 : (desc->storage == FLB_STORE_F32 ? 4 : 2);
  out->pop_bytes = out->pop_elems * elem;
  return FLB_OK;
}

int flb_block_create(const flb_block_desc* desc, flb_block* out) {
  if (!desc || !out) return fail(FLB_ERR_ARG, "null descriptor or output");
  *out = nullptr;
  flb_layout lay{};
  if (int rc = flb_block_layout(desc, &lay)) return rc;
  if (!(desc->omega > 0.0 && desc->omega < 2.0))
    return fail(FLB_ERR_ARG, "omega must lie in (0, 2), got %g", desc->omega);

  flb_block b = new flb_block_s();
  b->desc = *desc;
  b->d = (desc->q == FLB_D3Q19 || desc->q == FLB_D3Q7) ? 3 : 2;
  if (b->d == 2) b->desc.periodic[2] = 0;
  b->N = int64_t(desc->nx) * desc->ny * desc->nz;
  b->stride = lay.stride;
  b->pitch = lay.pitch;
  b->plane = lay.plane;
  b->dpop = lay.dpop;
  b->ypop = int(lay.ypop);
  b->zpop = lay.zpop;
  b->pop_elems = lay.pop_elems;
  b->elem = desc->storage == FLB_STORE_F64 ? 8 : (desc->storage == FLB_STORE_F32 ? 4 : 2);
  // one wavefront per workgroup for the 16-bit kinds (measured ~1 % faster on
  // the 256^3 cavity; FP32 / FP64 do not care)
  if (b->elem == 2) b->wpb = 1;
  b->real = desc->storage == FLB_STORE_F64 ? 8 : 4;
  b->stream = static_cast<hipStream_t>(desc->stream);
  if (hipGetDevice(&b->device) != hipSuccess) {
    delete b;
    return fail(FLB_ERR_HIP, "no HIP device");
  }
  if (const char* env = std::getenv("FLB_LAUNCH")) {
    // "vec,waves_per_workgroup", e.g. "4,4": vec 0 automatic cell form, -1 cell form over
    // row segments, -3 cell form over linear runs, 1/2/4 row form; invalid settings are
    // ignored
    int v = 0, w = 0;
    if (std::sscanf(env, "%d,%d", &v, &w) == 2 && v >= -3 && v <= 4 && v != 3 && v != -2 && w >= 1 &&
        w * flb::kWave <= flb::kMaxThreads) {
      b->vec = v;
      b->wpb = w;
    }
  }
  const size_t pop_bytes = size_t(b->pop_elems) * b->elem;
  if (hipMalloc(&b->d_pops, pop_bytes) != hipSuccess) {
    delete b;
    return fail(FLB_ERR_NOMEM, "population allocation of %zu bytes failed", pop_bytes);
  }
  if (hipMalloc(reinterpret_cast<void**>(&b->d_flags), size_t(b->stride)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_red), 4 * sizeof(double)) != hipSuccess) {
    flb_block_destroy(b);
    return fail(FLB_ERR_NOMEM, "flag allocation failed");
  }
  b->h_flags.assign(size_t(b->N), 0);
  if (hipMemsetAsync(b->d_flags, 0, size_t(b->stride), b->stream) != hipSuccess ||
      hipMemsetAsync(b->d_pops, 0, pop_bytes, b->stream) != hipSuccess) {
    flb_block_destroy(b);
    return fail(FLB_ERR_HIP, "initial clear failed");
  }
  default_rates(b);
  *out = b;
  return FLB_OK;
}

int flb_block_destroy(flb_block b) {
  if (!b) return FLB_OK;
  DeviceGuard g(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  else (void)hipDeviceSynchronize();
  detach_couplings(b);
  void* ptrs[] = {b->d_pops,  b->d_flags,   b->d_rho,   b->d_u,       b->d_partials, b->d_red,
                  b->d_stage, b->d_links, b->d_rho_old, b->d_u_old, b->d_res};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete b;
  return FLB_OK;
}

int flb_get_info(flb_block b, flb_block_info* info) {
  CHECK_BLOCK(b);
  if (!info) return fail(FLB_ERR_ARG, "null info");
  info->q = b->desc.q;
  info->storage = b->desc.storage;
  info->nx = b->desc.nx;
  info->ny = b->desc.ny;
  info->nz = b->desc.nz;
  info->overlap = b->desc.overlap;
  for (int a = 0; a < 3; ++a) info->periodic[a] = b->desc.periodic[a];
  info->n_cells = b->N;
  info->n_interior = int64_t(axis_ext(b, 0)) * axis_ext(b, 1) * axis_ext(b, 2);
  info->stride = b->stride;
  info->step = b->step;
  info->parity = b->parity;
  info->d_pops = b->d_pops;
  info->d_flags = b->d_flags;
  return FLB_OK;
}

int flb_synchronize(flb_block b) {
  CHECK_BLOCK(b);
  DeviceGuard g(b->device);
  HIP_TRY(hipStreamSynchronize(b->stream));
  return FLB_OK;
}

int flb_set_tasks(flb_block b, const flb_task* tasks, int n) {
  CHECK_BLOCK(b);
  b->version++;
  if (n < 0 || n > flb::kMaxTasks) return fail(FLB_ERR_ARG, "task count %d outside [0, %d]", n, flb::kMaxTasks);
  if (n > 0 && !tasks) return fail(FLB_ERR_ARG, "null task table");
  const bool scalar = scalar_lattice(b);
  for (int t = 0; t < n; ++t) {
    const int dyn = tasks[t].dynamics;
    if (dyn > FLB_DYN_SCALAR_BGK_FIELD)
      return fail(FLB_ERR_ARG, "task %d: unknown dynamics %d", t, dyn);
    const bool scalar_dyn = dyn == FLB_DYN_SCALAR_BGK || dyn == FLB_DYN_SCALAR_BGK_FIELD;
    if (dyn != FLB_DYN_NONE && scalar != scalar_dyn)
      return fail(FLB_ERR_ARG, "task %d: dynamics %d is for %s lattices", t, dyn,
                  scalar_dyn ? "D2Q5 / D3Q7" : "D2Q9 / D3Q19");
    if (tasks[t].mask == 0) return fail(FLB_ERR_ARG, "task %d: zero mask never matches", t);
  }
  flb_task saved[flb::kMaxTasks];
  const int saved_n = b->ntasks;
  std::memcpy(saved, b->tasks, sizeof saved);
  for (int t = 0; t < n; ++t) b->tasks[t] = tasks[t];
  b->ntasks = n;
  int rc = check_layout(b);
  if (rc != FLB_OK) {
    std::memcpy(b->tasks, saved, sizeof saved);
    b->ntasks = saved_n;
    return rc;
  }
  // rhoU<CELL, true>, forcerhoU<..., true> and the scalar dynamics read or
  // write the RHO / VELOCITY fields: give the block fields at FreeLB's
  // initial values (rho 1, u 0) when it has none yet
  bool fields = false;
  for (int t = 0; t < n; ++t) {
    const int dyn = tasks[t].dynamics;
    fields = fields || dyn == FLB_DYN_BGK_WRITE || dyn == FLB_DYN_BGK_SCALAR_FORCE || dyn == FLB_DYN_SCALAR_BGK ||
             dyn == FLB_DYN_SCALAR_BGK_FIELD;
  }
  if (fields) return ensure_fields(b);
  return FLB_OK;
}

int flb_set_halfway_bc(flb_block b, const flb_hwbc* bcs, int n) {
  CHECK_BLOCK(b);
  b->version++;
  if (n < 0 || n > flb::kMaxBc) return fail(FLB_ERR_ARG, "boundary count %d outside [0, %d]", n, flb::kMaxBc);
  if (n > 0 && !bcs) return fail(FLB_ERR_ARG, "null boundary table");
  for (int k = 0; k < n; ++k) {
    if (bcs[k].kind != FLB_HWBC_NORMAL && bcs[k].kind != FLB_HWBC_MOVING && bcs[k].kind != FLB_HWBC_ANTI)
      return fail(FLB_ERR_ARG, "boundary %d: unknown kind %d", k, bcs[k].kind);
    if (bcs[k].kind == FLB_HWBC_MOVING && scalar_lattice(b))
      return fail(FLB_ERR_ARG, "boundary %d: moving walls are for D2Q9 / D3Q19 lattices", k);
    if (bcs[k].kind == FLB_HWBC_ANTI && !scalar_lattice(b))
      return fail(FLB_ERR_ARG, "boundary %d: anti-bounce-back is for D2Q5 / D3Q7 lattices", k);
    if (bcs[k].void_mask == 0)
      return fail(FLB_ERR_ARG, "boundary %d: void mask 0 (isFlag needs a non-zero flag)", k);
  }
  for (int k = 0; k < n; ++k) b->bcs[k] = bcs[k];
  b->nbc = n;
  b->links_dirty = true;
  for (int k = 0; k < n; ++k)
    if (bcs[k].kind == FLB_HWBC_ANTI) return ensure_fields(b);  // the wall value is the RHO field
  return FLB_OK;
}

int flb_share_velocity(flb_block scalar, flb_block flow) {
  CHECK_BLOCK(scalar);
  if (!flow) {
    detach_couplings_one(scalar);
    return FLB_OK;
  }
  int rc = check_pair(flow, scalar);
  if (rc) return rc;
  if ((rc = ensure_fields(flow)) || (rc = ensure_fields(scalar))) return rc;
  detach_couplings_one(scalar);
  scalar->vel_src = flow;
  flow->users.push_back(scalar);
  scalar->version++;
  return FLB_OK;
}

int flb_add_buoyancy(flb_block flow, flb_block scalar, uint8_t mask, double gbeta, double rho_init) {
  CHECK_BLOCK(flow);
  CHECK_BLOCK(scalar);
  int rc = check_pair(flow, scalar);
  if (rc) return rc;
  if (mask == 0) return fail(FLB_ERR_ARG, "zero mask never matches");
  if (int(flow->buoy.size()) >= flb::kMaxCouplings)
    return fail(FLB_ERR_ARG, "a flow block takes at most %d buoyancy couplings", flb::kMaxCouplings);
  if ((rc = ensure_fields(flow)) || (rc = ensure_fields(scalar))) return rc;
  flow->buoy.push_back({scalar, mask, gbeta, rho_init});
  scalar->users.push_back(flow);
  flow->version++;
  return FLB_OK;
}

int flb_clear_buoyancy(flb_block flow) {
  CHECK_BLOCK(flow);
  for (const auto& c : flow->buoy) {
    auto& u = c.src->users;
    for (size_t k = 0; k < u.size(); ++k)
      if (u[k] == flow) {
        u.erase(u.begin() + std::ptrdiff_t(k));
        break;
      }
  }
  flow->buoy.clear();
  flow->version++;
  return FLB_OK;
}

int flb_set_mrt_rates(flb_block b, const double* rates, int q) {
  CHECK_BLOCK(b);
  b->version++;
  if (!rates) {
    default_rates(b);
    return FLB_OK;
  }
  if (q != b->desc.q) return fail(FLB_ERR_ARG, "expected %d rates, got %d", b->desc.q, q);
  for (int j = 0; j < q; ++j) b->rates[j] = std::isnan(rates[j]) ? -1.0 : rates[j];
  return FLB_OK;
}

int flb_set_ghost_overlap(flb_block b, int enable) {
  CHECK_BLOCK(b);
  b->version++;
  const bool saved = b->ghost_overlap;
  b->ghost_overlap = enable != 0;
  const int rc = check_layout(b);
  if (rc != FLB_OK) b->ghost_overlap = saved;
  return rc;
}

int flb_set_force(flb_block b, const double* force, int d) {
  CHECK_BLOCK(b);
  b->version++;
  if (!force) return fail(FLB_ERR_ARG, "null force");
  if (d != b->d) return fail(FLB_ERR_ARG, "expected %d force components, got %d", b->d, d);
  for (int a = 0; a < 3; ++a) b->force[a] = a < d ? force[a] : 0.0;
  return FLB_OK;
}

int flb_set_omega(flb_block b, double omega) {
  CHECK_BLOCK(b);
  b->version++;
  if (!(omega > 0.0 && omega < 2.0)) return fail(FLB_ERR_ARG, "omega must lie in (0, 2), got %g", omega);
  b->desc.omega = omega;
  return FLB_OK;
}

int flb_upload_flags(flb_block b, const uint8_t* flags) {
  CHECK_BLOCK(b);
  b->version++;
  if (!flags) return fail(FLB_ERR_ARG, "null flags");
  DeviceGuard g(b->device);
  std::vector<uint8_t> saved;
  saved.swap(b->h_flags);
  b->h_flags.assign(flags, flags + b->N);
  const bool was_set = b->flags_set;
  b->flags_set = true;
  const int rc = check_layout(b);
  if (rc != FLB_OK) {
    b->h_flags.swap(saved);
    b->flags_set = was_set;
    return rc;
  }
  std::vector<uint8_t> slots(size_t(b->stride));
  to_slots<uint8_t, uint8_t>(b, b->h_flags.data(), slots.data(), uint8_t(0));
  b->links_dirty = true;
  HIP_TRY(hipMemcpyAsync(b->d_flags, slots.data(), size_t(b->stride), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return FLB_OK;
}

int flb_set_field_rho(flb_block b, const double* rho) {
  CHECK_BLOCK(b);
  b->version++;
  if (!rho) return fail(FLB_ERR_ARG, "null rho");
  DeviceGuard g(b->device);
  return upload_field(b, rho, 1, &b->d_rho);
}

int flb_set_field_u(flb_block b, const double* u) {
  CHECK_BLOCK(b);
  b->version++;
  if (!u) return fail(FLB_ERR_ARG, "null u");
  DeviceGuard g(b->device);
  return upload_field(b, u, b->d, &b->d_u);
}

int flb_enable_tolerance(flb_block b) {
  CHECK_BLOCK(b);
  if (!b->d_rho || !b->d_u)
    return fail(FLB_ERR_STATE, "the block has no RHO / VELOCITY fields (FLB_DYN_BGK_WRITE tasks create them)");
  DeviceGuard g(b->device);
  const size_t one = size_t(b->stride) * b->real;
  if (!b->d_rho_old && hipMalloc(&b->d_rho_old, one) != hipSuccess)
    return fail(FLB_ERR_NOMEM, "residual field allocation failed");
  if (!b->d_u_old && hipMalloc(&b->d_u_old, one * size_t(b->d)) != hipSuccess)
    return fail(FLB_ERR_NOMEM, "residual field allocation failed");
  if (!b->d_res && hipMalloc(reinterpret_cast<void**>(&b->d_res), sizeof(unsigned long long)) != hipSuccess)
    return fail(FLB_ERR_NOMEM, "residual allocation failed");
  HIP_TRY(hipMemcpyAsync(b->d_rho_old, b->d_rho, one, hipMemcpyDeviceToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_u_old, b->d_u, one * size_t(b->d), hipMemcpyDeviceToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return FLB_OK;
}

int flb_tolerance(flb_block b, int which, int shift, double* res) {
  CHECK_BLOCK(b);
  if (!res) return fail(FLB_ERR_ARG, "null output");
  if (which != 0 && which != 1) return fail(FLB_ERR_ARG, "which must be 0 (RHO) or 1 (VELOCITY)");
  if (shift < 0) return fail(FLB_ERR_ARG, "negative shift");
  if (!b->d_res) return fail(FLB_ERR_STATE, "flb_enable_tolerance first");
  if (b->pending) return fail(FLB_ERR_STATE, "residual between ApplyCellDynamics and Stream");
  DeviceGuard g(b->device);
  const size_t pad = size_t(flb::kFrontPad) * b->real;
  const void* field = static_cast<const char*>(which == 0 ? b->d_rho : b->d_u) + pad;
  void* old = static_cast<char*>(which == 0 ? b->d_rho_old : b->d_u_old) + pad;
  HIP_TRY(hipMemsetAsync(b->d_res, 0, sizeof(unsigned long long), b->stream));
  HIP_TRY(flb::launch_tolerance(b->real == 8, field, old, b->stride, which == 0 ? 1 : b->d, b->desc.nx, b->desc.ny,
                                b->desc.nz, int(axis_lo(b, 0)), b->pitch, b->plane, shift, b->d_res, b->stream));
  unsigned long long bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, b->d_res, sizeof bits, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  double v;
  std::memcpy(&v, &bits, sizeof v);
  *res = v;
  return FLB_OK;
}

int flb_download_fields(flb_block b, double* rhou) {
  CHECK_BLOCK(b);
  if (!rhou) return fail(FLB_ERR_ARG, "null output");
  if (b->pending) return fail(FLB_ERR_STATE, "fields read between ApplyCellDynamics and Stream");
  DeviceGuard g(b->device);
  const int nx = b->desc.nx, ny = b->desc.ny, nz = b->desc.nz;
  for (int c = 0; c <= b->d; ++c) {
    void* src = c == 0 ? b->d_rho : b->d_u;
    double* out = rhou + size_t(c) * size_t(b->N);
    if (!src) {
      for (int64_t k = 0; k < b->N; ++k) out[k] = c == 0 ? 1.0 : 0.0;
      continue;
    }
    const size_t off = c == 0 ? 0 : size_t(c - 1) * size_t(b->stride);
    std::vector<char> tmp(size_t(b->stride) * b->real);
    HIP_TRY(hipMemcpyAsync(tmp.data(), static_cast<char*>(src) + off * b->real, tmp.size(), hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const int64_t sl = flb::kFrontPad + slot_of(b, x, y, z);
          const double v = b->real == 8 ? reinterpret_cast<const double*>(tmp.data())[sl]
                                        : double(reinterpret_cast<const float*>(tmp.data())[sl]);
          out[size_t(x) + size_t(nx) * (size_t(y) + size_t(ny) * z)] = v;
        }
  }
  return FLB_OK;
}

static int upload_pops(flb_block b, const void* host, size_t host_real) {
  if (host_real != b->real)
    return fail(FLB_ERR_ARG, "population precision does not match storage kind %d", b->desc.storage);
  DeviceGuard g(b->device);
  const size_t bytes = size_t(b->N) * size_t(b->desc.q) * b->real;
  int rc = ensure_stage(b, bytes);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(b->d_stage, host, bytes, hipMemcpyHostToDevice, b->stream));
  KArgs a = make_args(b);
  HIP_TRY(launch_scatter(b, a, b->d_stage));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->parity = 0;
  b->pending = 0;
  return FLB_OK;
}

int flb_upload_pops_f32(flb_block b, const float* pops) {
  CHECK_BLOCK(b);
  if (!pops) return fail(FLB_ERR_ARG, "null pops");
  return upload_pops(b, pops, 4);
}

int flb_upload_pops_f64(flb_block b, const double* pops) {
  CHECK_BLOCK(b);
  if (!pops) return fail(FLB_ERR_ARG, "null pops");
  return upload_pops(b, pops, 8);
}

static int download_pops(flb_block b, void* host, size_t host_real) {
  if (host_real != b->real)
    return fail(FLB_ERR_ARG, "population precision does not match storage kind %d", b->desc.storage);
  if (b->pending) return fail(FLB_ERR_STATE, "download between ApplyCellDynamics and Stream");
  DeviceGuard g(b->device);
  const size_t bytes = size_t(b->N) * size_t(b->desc.q) * b->real;
  int rc = ensure_stage(b, bytes);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(b->d_stage, 0, bytes, b->stream));
  KArgs a = make_args(b);
  HIP_TRY(launch_gather(b, a, b->d_stage));
  HIP_TRY(hipMemcpyAsync(host, b->d_stage, bytes, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return FLB_OK;
}

int flb_download_pops_f32(flb_block b, float* pops) {
  CHECK_BLOCK(b);
  if (!pops) return fail(FLB_ERR_ARG, "null pops");
  return download_pops(b, pops, 4);
}

int flb_download_pops_f64(flb_block b, double* pops) {
  CHECK_BLOCK(b);
  if (!pops) return fail(FLB_ERR_ARG, "null pops");
  return download_pops(b, pops, 8);
}

int flb_init_equilibrium(flb_block b, const double* rhou) {
  CHECK_BLOCK(b);
  if (!rhou) return fail(FLB_ERR_ARG, "null rhou");
  DeviceGuard g(b->device);
  const size_t count = size_t(b->N) * size_t(b->d + 1);
  int rc = ensure_stage(b, count * b->real);
  if (rc) return rc;
  if (b->real == 8) {
    HIP_TRY(hipMemcpyAsync(b->d_stage, rhou, count * 8, hipMemcpyHostToDevice, b->stream));
  } else {
    std::vector<float> tmp(count);
    for (size_t i = 0; i < count; ++i) tmp[i] = float(rhou[i]);
    HIP_TRY(hipMemcpyAsync(b->d_stage, tmp.data(), count * 4, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  }
  KArgs a = make_args(b);
  HIP_TRY(launch_init_feq(b, a, b->d_stage));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->parity = 0;
  b->pending = 0;
  return FLB_OK;
}

static int launch_one(flb_block b) {
  if (!b->flags_set) return fail(FLB_ERR_STATE, "flags not uploaded");
  const int rc = ensure_links(b);
  if (rc) return rc;
  KArgs a = make_args(b);
  HIP_TRY(launch_step(b, family_of(b), a));
  return FLB_OK;
}

int flb_set_level(flb_block b, int level, int max_level) {
  CHECK_BLOCK(b);
  if (level < 0 || max_level < level || max_level > 30)
    return fail(FLB_ERR_ARG, "level %d / max level %d", level, max_level);
  if (b->pending) return fail(FLB_ERR_STATE, "level change between ApplyCellDynamics and Stream");
  b->level = level;
  b->max_level = max_level;
  return FLB_OK;
}

int flb_apply_cell_dynamics(flb_block b, int64_t count) {
  CHECK_BLOCK(b);
  if (count < 0) return fail(FLB_ERR_ARG, "negative count");
  if (b->pending) return fail(FLB_ERR_STATE, "ApplyCellDynamics called twice without Stream");
  if (!level_due(count, b->level, b->max_level)) return FLB_OK;
  DeviceGuard g(b->device);
  const int rc = launch_one(b);
  if (rc) return rc;
  b->pending = 1;
  return FLB_OK;
}

int flb_stream(flb_block b, int64_t count) {
  CHECK_BLOCK(b);
  if (count < 0) return fail(FLB_ERR_ARG, "negative count");
  if (!level_due(count, b->level, b->max_level)) {
    if (b->pending) return fail(FLB_ERR_STATE, "Stream skipped at this count after an ApplyCellDynamics");
    return FLB_OK;
  }
  if (!b->pending) return fail(FLB_ERR_STATE, "Stream without a preceding ApplyCellDynamics");
  b->pending = 0;
  b->parity ^= 1;
  b->step += 1;
  return FLB_OK;
}

int flb_apply_halfway_bc(flb_block b) {
  CHECK_BLOCK(b);
  if (b->pending) return fail(FLB_ERR_STATE, "halfway boundaries apply after Stream");
  return FLB_OK;
}

int flb_set_watchdog(flb_block b, int every) {
  CHECK_BLOCK(b);
  if (every < 0) return fail(FLB_ERR_ARG, "negative watchdog interval");
  b->watchdog = every;
  return FLB_OK;
}

static int watchdog_check(flb_block b) {
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  const int rc = flb_reduce_mass_momentum(b, 0xFF, m);
  if (rc) return rc;
  for (double v : m)
    if (!std::isfinite(v))
      return fail(FLB_ERR_NONFINITE, "non-finite mass / momentum (%g, %g, %g, %g) after step %lld", m[0], m[1],
                  m[2], m[3], static_cast<long long>(b->step));
  return FLB_OK;
}

int flb_step(flb_block b, int n) {
  CHECK_BLOCK(b);
  if (n < 0) return fail(FLB_ERR_ARG, "negative step count");
  if (b->pending) return fail(FLB_ERR_STATE, "step while a Stream is pending");
  DeviceGuard g(b->device);
  Range range("flb_step");
  for (int s = 0; s < n; ++s) {
    int rc = launch_one(b);
    if (rc) return rc;
    b->parity ^= 1;
    b->step += 1;
    if (b->watchdog > 0 && b->step % b->watchdog == 0 && (rc = watchdog_check(b))) return rc;
  }
  return FLB_OK;
}

int flb_moments(flb_block b, uint8_t mask, void* rhou) {
  CHECK_BLOCK(b);
  if (!rhou) return fail(FLB_ERR_ARG, "null output");
  if (b->pending) return fail(FLB_ERR_STATE, "moments between ApplyCellDynamics and Stream");
  DeviceGuard g(b->device);
  const size_t bytes = size_t(b->N) * size_t(b->d + 1) * b->real;
  int rc = ensure_stage(b, bytes);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(b->d_stage, 0, bytes, b->stream));
  KArgs a = make_args(b);
  HIP_TRY(launch_moments(b, a, mask, b->d_stage));
  HIP_TRY(hipMemcpyAsync(rhou, b->d_stage, bytes, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return FLB_OK;
}

int flb_reduce_mass_momentum(flb_block b, uint8_t mask, double out[4]) {
  CHECK_BLOCK(b);
  if (!out) return fail(FLB_ERR_ARG, "null output");
  if (b->pending) return fail(FLB_ERR_STATE, "reduction between ApplyCellDynamics and Stream");
  DeviceGuard g(b->device);
  KArgs a = make_args(b);
  if (!b->d_partials) {
    // one double4 per workgroup of the per-cell kernels
    const size_t nb = flb::cell_blocks(b->N);
    if (hipMalloc(reinterpret_cast<void**>(&b->d_partials), nb * 4 * sizeof(double)) != hipSuccess)
      return fail(FLB_ERR_NOMEM, "reduction scratch allocation failed");
  }
  HIP_TRY(launch_reduce(b, a, mask, b->d_partials, b->d_red));
  HIP_TRY(hipMemcpyAsync(out, b->d_red, 4 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return FLB_OK;
}

}  // extern "C"

namespace {

// ---- scratch of the allocation-free reduction ---------------------------------
// One double4 of partial sums per workgroup of the per-cell kernels, for
// blocks of up to kNoAllocMaxWG * 256 = 2^28 host cells: 32 MiB, held by
// every device the library runs on from the code object's load until the
// process ends. The array's name is a host-side stand-in; kernels get the
// device address that hipGetSymbolAddress returns for the current device.
constexpr int64_t kNoAllocMaxWG = int64_t(1) << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[4 * kNoAllocMaxWG];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on device dev, looked up once per device
// with that device current.
int scratch_on(int dev, double** out) {
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(FLB_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    DeviceGuard g(dev);
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FLB_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FLB_OK;
}

}  // namespace

extern "C" {

int flb_reduce_mass_momentum_noalloc(flb_block b, uint8_t mask, double* d_out) {
  CHECK_BLOCK(b);
  if (!d_out) return fail(FLB_ERR_ARG, "null device output");
  if (b->pending) return fail(FLB_ERR_STATE, "reduction between ApplyCellDynamics and Stream");
  const int64_t nb = int64_t(flb::cell_blocks(b->N));
  if (nb > kNoAllocMaxWG)
    return fail(FLB_ERR_ARG, "block of %lld host cells exceeds the allocation-free reduction's maximum of %lld",
                static_cast<long long>(b->N), static_cast<long long>(kNoAllocMaxWG * flb::kMaxThreads));
  double* partials = nullptr;
  const int rc = scratch_on(b->device, &partials);
  if (rc) return rc;
  DeviceGuard g(b->device);
  const KArgs a = make_args(b);
  // every workgroup overwrites its own partials, so nothing needs clearing
  HIP_TRY(launch_reduce(b, a, mask, partials, d_out));
  return FLB_OK;
}

}  // extern "C"

namespace {
// Times `reps` launches of one probe form over two fresh buffers of `bytes`
// (rounded down to whole tiles); *gbs gets the bytes moved per launch (read +
// written) over the mean launch time.
int run_probe(int mode, int unroll, bool nt, int wg_per_cu, bool legacy, size_t bytes, int reps, double* gbs) {
  const size_t tile = size_t(unroll) * flb::kMaxThreads;
  const size_t n4 = legacy ? bytes / 16 : (bytes / 16) / tile * tile;
  if (n4 == 0) return fail(FLB_ERR_ARG, "probe buffer smaller than one tile");
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    return fail(FLB_ERR_HIP, "device query failed");
  const size_t ntile = n4 / tile;
  const size_t grid = wg_per_cu > 0 ? std::min(ntile, size_t(wg_per_cu) * size_t(cus)) : ntile;
  if (grid > size_t(0x7fffffff)) return fail(FLB_ERR_ARG, "probe grid too large");
  void *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, n4 * 16) != hipSuccess) return fail(FLB_ERR_NOMEM, "probe allocation failed");
  if (hipMalloc(&dst, n4 * 16) != hipSuccess) {
    (void)hipFree(src);
    return fail(FLB_ERR_NOMEM, "probe allocation failed");
  }
  auto launch = [&]() {
    return legacy ? flb::launch_copy_probe(dst, src, n4, nullptr)
                  : flb::launch_bw_probe(mode, unroll, nt, unsigned(grid), dst, src, n4, nullptr);
  };
  hipEvent_t e0, e1;
  int rc = FLB_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail(FLB_ERR_HIP, "event creation failed");
  } else if (hipMemset(src, 0, n4 * 16) != hipSuccess || hipMemset(dst, 0, n4 * 16) != hipSuccess ||
             launch() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    rc = fail(FLB_ERR_HIP, "probe warm-up failed");
  } else {
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) (void)launch();
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    if (hipGetLastError() != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f) {
      rc = fail(FLB_ERR_HIP, "probe timing failed");
    } else {
      const double moved = (mode == 0 ? 2.0 : 1.0) * double(n4 * 16);
      *gbs = moved * reps / (double(ms) * 1e-3) / 1e9;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(src);
  (void)hipFree(dst);
  return rc;
}
}  // namespace

extern "C" {

int flb_bandwidth_probe(size_t bytes, int reps, double* gbs) {
  if (!gbs || reps < 1 || bytes < 16) return fail(FLB_ERR_ARG, "bad probe arguments");
  return run_probe(0, 1, false, 0, true, bytes, reps, gbs);
}

int flb_bandwidth_probe_ex(int mode, int unroll, int nontemporal, int wg_per_cu, size_t bytes, int reps,
                           double* gbs) {
  if (!gbs || reps < 1 || mode < 0 || mode > 2 || wg_per_cu < 0 ||
      !(unroll == 1 || unroll == 2 || unroll == 4 || unroll == 8))
    return fail(FLB_ERR_ARG, "bad probe arguments");
  return run_probe(mode, unroll, nontemporal != 0, wg_per_cu, false, bytes, reps, gbs);
}

}  // extern "C"
