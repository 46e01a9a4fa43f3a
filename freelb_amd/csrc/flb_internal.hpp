// Internal interface between the C-ABI host layer and the kernels.
//
// Device layout ("slots"). Host arrays are dense, id = x + nx*(y + ny*z).
// On the device the flags and the RHO / VELOCITY fields use the padded slot
//     slot(x, y, z) = (x - xs) + pitch*y + plane*z,   plane = pitch*ny,
// where xs is the first updated x (the overlap width, 0 on a periodic x axis)
// and pitch >= nx is a multiple of 64. Row y, z therefore starts its updated
// cells on a 256-byte boundary for every element size, and the x = xs - 1
// overlap cell sits in the last, otherwise unused, slot of the row before.
// Populations use the same rows, placed by three strides:
//     pslot(i, x, y, z) = (x - xs) + ypop*y + zpop*z + dpop*i.
// The default interleaves the q directions of each row (dpop = pitch,
// ypop = q*pitch, zpop = q*plane): the q rows a cell update touches at the
// even step are one contiguous run, and the 3 x 3 x q rows of the odd step
// lie in nine such runs instead of q separate arrays.
// Every array reserves kFrontPad slots before slot 0 and a tail after the
// last row, so vector loads that run past a row end stay inside it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flb {

constexpr int kMaxTasks = 8;
constexpr int kMaxBc = 4;
constexpr int kMaxThreads = 256;
constexpr int kWave = 64;
constexpr int kFrontPad = 256;

// Dynamics families: the kernel for a family compiles only the collisions in
// it, so the common BGK / bounce-back step does not pay the registers of MRT.
// FAM_SCALAR is the advection-diffusion family of the D2Q5 / D3Q7 lattices.
enum Family { FAM_BGK = 0, FAM_FULL = 1, FAM_SCALAR = 2 };

// Buoyancy couplings one flow block can carry (FreeLB's examples couple a
// temperature and a concentration lattice).
constexpr int kMaxCouplings = 2;

// Everything a lattice kernel needs, passed by value (kernel argument space).
struct KArgs {
  void* pops;            // population slot 0 of direction 0
  int64_t dpop;          // elements between directions
  int ypop;              // population slots between rows
  int64_t zpop;          // population slots between planes
  const uint8_t* flags;  // slot 0
  const void* rho_f;     // slot 0 of the RHO field (compute type), or null (rho = 1)
  const void* u_f;       // slot 0 of VELOCITY x; component d at + d*fstride, or null (u = 0)
  int64_t fstride;
  int nx, ny, nz;
  int xs, ys, zs;        // first updated cell
  int ex, ey, ez;        // updated extents
  int pitch;
  int64_t plane;
  int px, py, pz;        // periodic axes
  int vec;               // cells per thread along x
  int groups;            // thread groups per row: ceil(ex / vec)
  int wpr;               // wavefronts per row: ceil(groups / 64)
  int wpb;               // wavefronts per workgroup
  int nblocks;           // workgroups
  int ntasks;
  int nbc;
  uint8_t tmask[kMaxTasks];
  uint8_t tdyn[kMaxTasks];
  float tcs[kMaxTasks];
  uint8_t bcmask[kMaxBc];
  uint8_t bcvoid[kMaxBc];
  uint8_t bckind[kMaxBc];
  uint8_t bcany;         // OR of bcmask: cells that may have halfway links
  const uint64_t* links; // slot 0 of the precomputed wall links (normal | moving << 32)
  uint64_t pop_total;    // bytes of the population allocation
  uint32_t buf;          // 1: the step may address populations through wavefront-based buffer descriptors
  double omega;
  double rates[19];      // MRT rates; < 0 or NaN marks the shear rates
  float mrt_sn_f[19];    // D3Q19 MRT: rate / |M_j|^2 of the fixed-rate rows (FP32 kinds)
  double mrt_sn_d[19];   // the same in FP64
  uint32_t mrt_shear;    // D3Q19 MRT: bit j set for the rows that take omega
  double force[3];       // constant body-force density of DYN_BGK_FORCE
  // advection-diffusion dynamics: slot 0 of the VELOCITY field that advects
  // the scalar (the coupled flow block's), or null for the block's own u_f
  const void* u_ext;
  // DYN_BGK_SCALAR_FORCE: force on the last axis = sum over the couplings k
  // whose mask matches the flag of gbeta_k (C_k - C0_k), C_k the coupled
  // block's RHO (scalar) field in slot layout
  int ncpl;
  uint8_t cpl_mask[kMaxCouplings];
  const void* cpl_rho[kMaxCouplings];
  double cpl_gbeta[kMaxCouplings];
  double cpl_rho0[kMaxCouplings];
};

// Launchers, one set per storage kind (defined in flb_inst_*.hip).
#define FLB_DECLARE_LAUNCHERS(SUFFIX)                                                         \
  hipError_t launch_step_##SUFFIX(int q, int parity, int family, const KArgs& a,             \
                                  hipStream_t s);                                           \
  hipError_t launch_gather_##SUFFIX(int q, int parity, const KArgs& a, void* out,            \
                                    hipStream_t s);                                         \
  hipError_t launch_scatter_##SUFFIX(int q, const KArgs& a, const void* in, hipStream_t s);  \
  hipError_t launch_init_feq_##SUFFIX(int q, const KArgs& a, const void* rhou, hipStream_t s); \
  hipError_t launch_moments_##SUFFIX(int q, int parity, const KArgs& a, uint8_t mask,        \
                                     void* out, hipStream_t s);                             \
  hipError_t launch_reduce_##SUFFIX(int q, int parity, const KArgs& a, uint8_t mask,         \
                                    double* partials, double* out, hipStream_t s);          \
  hipError_t launch_batch_##SUFFIX(int q, int parity, int family, bool per, const KArgs* tab, \
                                   const int64_t* wave0, const int* wblk, int nb,             \
                                   int64_t groups, int wpb, int grid, hipStream_t s);

FLB_DECLARE_LAUNCHERS(f32)
FLB_DECLARE_LAUNCHERS(f16s)
FLB_DECLARE_LAUNCHERS(bf16s)
FLB_DECLARE_LAUNCHERS(f64)
FLB_DECLARE_LAUNCHERS(f16c)

// Ghost cells of block a that mirror updated cells of block b (block halo):
// a box in a-local coordinates and the shift from a-local to b-local ones.
struct HaloItem {
  int a, b;
  int gx, gy, gz;
  int nx, ny, nz;
  int sx, sy, sz;
};

// Halfway wall links of every updated cell into links (slot layout), see k_links.
hipError_t launch_links(int q, const KArgs& a, uint64_t* links, hipStream_t s);

// float4 copy of n4 16-byte elements, one element per thread (bandwidth calibration).
hipError_t launch_copy_probe(void* dst, const void* src, size_t n4, hipStream_t s);
// Bandwidth probe forms (flb_probe.hip): mode 0 copy, 1 read-only, 2
// write-only; unroll 16-byte accesses per lane in flight (1, 2, 4, 8);
// non-temporal or plain; grid workgroups striding over n4 elements (n4 a
// multiple of unroll * 256).
hipError_t launch_bw_probe(int mode, int unroll, bool nt, unsigned grid, void* dst, const void* src, size_t n4,
                           hipStream_t s);

// Convergence residual of one field (ncomp components at fstride, field and
// old in slot layout from slot 0) over the host cells [shift, n - shift) per
// axis; folds the maximum into *res (bits of a non-negative double) and
// copies the field into old (flb_tol.hip).
hipError_t launch_tolerance(bool dbl, const void* field, void* old, int64_t fstride, int ncomp, int nx, int ny,
                            int nz, int xlo, int pitch, int64_t plane, int shift, unsigned long long* res,
                            hipStream_t s);

// Workgroups of the per-cell (non-step) kernels for a block of n_cells host cells.
inline unsigned cell_blocks(int64_t n_cells) {
  return unsigned((n_cells + kMaxThreads - 1) / kMaxThreads);
}

}  // namespace flb
