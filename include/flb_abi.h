/*
 * flb_abi.h -- C ABI of the MI355X (gfx950) lattice-update engine.
 *
 * One flb_block is one FreeLB BlockLattice living on one GPU: SoA population
 * storage, a uint8 flag field, a first-match task table (FreeLB's
 * TaskSelector<flag, Key_TypePair<mask, Dynamics>...>) and a list of halfway
 * bounce-back managers (BBLikeFixedBlockBdManager). The update is the fused
 * collide-and-stream step of the AA (esoteric) in-place pattern, with the
 * halfway boundary fix-up folded into the same kernel.
 *
 * Conventions
 *   - Every function returns FLB_OK (0) or an FLB_ERR_* code; the message of
 *     the last failure on the calling thread is flb_last_error().
 *   - Host arrays are plain C arrays: populations are SoA [q][N], fields are
 *     SoA [d][N], with N = nx*ny*nz and cell index id = x + nx*(y + ny*z)
 *     (FreeLB's BasicBlock::getIndex order).
 *   - Work is enqueued on the block's stream; functions that return data to
 *     the host synchronise that stream. Calls on one block are not
 *     re-entrant; different blocks may be driven from different threads.
 *
 * Reference interfaces replaced (paths relative to the FreeLB source tree):
 *   flb_apply_cell_dynamics  BlockLattice::ApplyCellDynamics<TaskSelector>(flagarr)
 *                              data_struct/block_lattice.hh:72-92 (CPU),
 *                              BlockLatticeManager::CuDevApplyCellDynamics
 *                              data_struct/block_lattice.hh:919-936 (CUDA)
 *   flb_stream               BlockLattice::Stream / CuDevStream
 *                              data_struct/block_lattice.hh:64-70, 915-917
 *   flb_apply_halfway_bc     BlockBoundaryManager::Apply -> BBLikeFixedBlockBdManager::Apply
 *                              boundary/basic_boundary.h:345-347, boundary/boundary.h:121-128
 *   flb_set_tasks            TaskSelector / Key_TypePair  utils/tmp.h:194-202, 283-309
 *   flb_set_halfway_bc       BBLikeFixedBlockBdManager ctor + BlockFixedBoundary::Setup
 *                              boundary/boundary.h:94-111, boundary/basic_boundary.h:198-264
 *   flb_block_create         BlockLattice ctor (Omega, 1-Omega, 1-Omega/2)
 *                              data_struct/block_lattice.hh:24-35; POP field storage
 *                              data_struct/field.h:1419-1575, utils/alias.h:296-308
 *   flb_upload_flags         BlockFieldManager<FLAG>::copyToDevice  data_struct/field.h:883-896
 *   flb_set_field_rho/u      RHO / VELOCITY fields read by moment::useFieldrhoU,
 *                              BounceBackMovingWall and bounceback::movingwall
 *                              lbm/moment.h:71-90, lbm/collision.ur.h:130-187,
 *                              boundary/bounce_back_boundary.h:240-256
 *   flb_moments              moment::rhoU<CELL, true> via TaskSelectorRhoU
 *                              lbm/moment.h:381-398, examples/cavity3d/cavity3d.cpp:192-194
 *   flb_init_equilibrium     BlockLatticeManager ctor pop init / equilibrium::SecondOrder
 *                              data_struct/block_lattice.hh:676-681, lbm/equilibrium.h:156-178
 *   flb_share_velocity,      BlockLatManagerCoupling with force::Buoyancy and the VELOCITY
 *   flb_add_buoyancy           field reference of a temperature / concentration lattice
 *                              data_struct/block_lattice.h:436-649, lbm/force.h:160-181,
 *                              examples/cazsblock2d/cazsblock2d.cpp:225-409
 */
#ifndef FLB_ABI_H
#define FLB_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLB_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
  FLB_OK = 0,
  FLB_ERR_ARG = 1,         /* bad argument (null pointer, bad size, unknown enum) */
  FLB_ERR_STATE = 2,       /* call out of order (e.g. ApplyCellDynamics twice without Stream) */
  FLB_ERR_HIP = 3,         /* a HIP runtime call failed */
  FLB_ERR_UNSUPPORTED = 4, /* combination not built (e.g. MRT on a storage kind without it) */
  FLB_ERR_LAYOUT = 5,      /* a cell on the outer layer of a non-periodic axis matches a task */
  FLB_ERR_NOMEM = 6,       /* device allocation failed */
  FLB_ERR_NONFINITE = 7    /* the watchdog found a non-finite mass or momentum (flb_set_watchdog) */
};

/* ---- enums ------------------------------------------------------------- */
/* Flow lattices D2Q9 / D3Q19; advection-diffusion (scalar) lattices D2Q5 /
 * D3Q7 (lbm/lattice_set.h:144-207), FP32 or FP64 storage only. */
enum { FLB_D2Q9 = 9, FLB_D3Q19 = 19, FLB_D2Q5 = 5, FLB_D3Q7 = 7 };

/* Population storage. The shifted forms store f_i - w_i (rho0 = 1) and do
 * all arithmetic in FP32: F16S as IEEE binary16, BF16S as bfloat16, F16C as
 * a 16-bit float with 4 exponent and 11 fraction bits (range 2^-28 .. 0.25,
 * one more significant bit than binary16). FLB_STORE_F64 stores and
 * computes in FP64. */
enum { FLB_STORE_F32 = 0, FLB_STORE_F16S = 1, FLB_STORE_BF16S = 2, FLB_STORE_F64 = 3, FLB_STORE_F16C = 4 };

/* Cell dynamics (the `Dynamics` of a Key_TypePair). */
enum {
  FLB_DYN_NONE = 0,          /* no-op: the cell is not updated (void / solid) */
  FLB_DYN_BGK = 1,           /* collision::BGK<moment::rhoU, equilibrium::SecondOrder>      lbm/collision.h:32-59 */
  FLB_DYN_BGK_FIELD = 2,     /* collision::BGK<moment::useFieldrhoU, SecondOrder>           lbm/moment.h:71-90 */
  FLB_DYN_BOUNCEBACK = 3,    /* collision::BounceBack (fullway wall as collision)           lbm/collision.ur.h:62-93 */
  FLB_DYN_BOUNCEBACK_MW = 4, /* collision::BounceBackMovingWall                             lbm/collision.ur.h:130-187 */
  FLB_DYN_MRT = 5,           /* collision::MRT_Feq_RhoU (D2Q9 as lbm/collisionMRT.ur.h:45-129;
                                D3Q19 on the d'Humieres 2002 basis, no reference form) */
  FLB_DYN_SMAG_BGK = 6,      /* collision::SmagorinskyBGK                                   lbm/collisionLES.h:29-67 */
  FLB_DYN_SMAG_MRT = 7,      /* MRT whose shear rates take the Smagorinsky omega (no reference form) */
  FLB_DYN_BGK_FORCE = 8,     /* collision::BGKForce<moment::forcerhoU, SecondOrder, force::ConstForce>
                                lbm/collision.h:61-90, lbm/force.h:29-36, lbm/moment.h:401-438;
                                the force is the block's constant (flb_set_force) */
  FLB_DYN_BGK_WRITE = 9,     /* collision::BGK<moment::rhoU<CELL, true>, SecondOrder>: BGK that also
                                stores rho and u in the RHO / VELOCITY fields (lbm/moment.h:359-377),
                                as the refined drivers do (examples/cavref2d/cavref2d.cpp:178-179) */
  FLB_DYN_BGK_SCALAR_FORCE = 10, /* collision::BGKForce<moment::forcerhoU<CELL, force::ScalarForce, true>,
                                SecondOrder, force::ScalarForce> (lbm/collision.h:61-90, lbm/force.h:109-133,
                                lbm/moment.h:401-435): the force on the last axis is the SCALARFORCE of
                                the block's buoyancy couplings (flb_add_buoyancy), and rho / u go to the
                                RHO / VELOCITY fields (examples/cazsblock2d/cazsblock2d.cpp:306-309) */
  FLB_DYN_SCALAR_BGK = 11,   /* D2Q5 / D3Q7: collision::BGK<moment::MomentaTuple<moment::rho<CELL, true>,
                                moment::useFieldU<CELL>>, SecondOrder>: the scalar (sum of the populations)
                                goes to the RHO field, advected by the VELOCITY field (lbm/moment.h:56-69,
                                159-182, 667-679) */
  FLB_DYN_SCALAR_BGK_FIELD = 12 /* D2Q5 / D3Q7: collision::BGK<moment::useFieldrhoU<CELL>, SecondOrder>:
                                relaxes toward the equilibrium of the RHO field's scalar (a Dirichlet wall
                                value) and the VELOCITY field (lbm/moment.h:71-90) */
};

/* Halfway bounce-back kinds: bounceback::normal / bounceback::movingwall
 * (boundary/bounce_back_boundary.h:169-181, 240-256), and for D2Q5 / D3Q7
 * blocks bounceback::anti_simple (183-193), f_k = 2 RHO w_k - f*_{opp k}
 * with RHO the cell's scalar field (a fixed-value wall). */
enum { FLB_HWBC_NORMAL = 0, FLB_HWBC_MOVING = 1, FLB_HWBC_ANTI = 2 };

/* ---- descriptors ------------------------------------------------------- */
typedef struct flb_block_s* flb_block;

typedef struct {
  int q;            /* FLB_D2Q9, FLB_D3Q19, FLB_D2Q5 or FLB_D3Q7 */
  int storage;      /* FLB_STORE_* */
  int nx, ny, nz;   /* extents including the overlap layer; nz = 1 in 2-D */
  int overlap;      /* overlap (ghost) width on non-periodic axes, >= 1 */
  int periodic[3];  /* non-zero: the axis wraps (no overlap layer on it) */
  double omega;     /* 1/tau of the block (BlockLattice::Omega) */
  void* stream;     /* hipStream_t to enqueue on; NULL = the null stream */
} flb_block_desc;

/* One Key_TypePair<mask, Dynamics>: the first entry whose (flag & mask) != 0 runs. */
typedef struct {
  uint8_t mask;
  uint8_t dynamics;  /* FLB_DYN_* */
  uint8_t pad[2];
  float smagorinsky; /* C_s for the Smagorinsky kinds (the SMAGORINSKY data field) */
} flb_task;

/* One halfway manager: boundary cells have (flag & cell_mask) != 0; a link to
 * a neighbour with (flag & void_mask) != 0 is a wall link. Entries apply in
 * order, so where two match the same cell the later one's value stands. */
typedef struct {
  uint8_t cell_mask;
  uint8_t void_mask;
  uint8_t kind;      /* FLB_HWBC_* */
  uint8_t pad;
} flb_hwbc;

typedef struct {
  int q, storage, nx, ny, nz, overlap;
  int periodic[3];
  int64_t n_cells;        /* nx*ny*nz */
  int64_t n_interior;     /* cells the kernel visits (overlap layer excluded) */
  int64_t stride;         /* elements between two population arrays */
  int64_t step;           /* completed time steps */
  int parity;             /* 0: natural layout, 1: AA-swapped layout */
  void* d_pops;           /* device base of the population arrays */
  const void* d_flags;    /* device flag field */
} flb_block_info;

/* Device layout of a block (HBM slots). Per-cell arrays (flags, each field
 * component) put cell (x, y, z) at slot (x - xs) + pitch*y + plane*z of an
 * array of `stride` slots that starts front_pad slots in; populations put
 * direction i at (x - xs) + ypop*y + zpop*z + dpop*i. Distinct cells get
 * distinct slots for every overlap width. */
typedef struct {
  int pitch, xs;
  int64_t plane, stride, front_pad;
  int64_t dpop, ypop, zpop;
  int64_t pop_elems;      /* population elements allocated */
  int64_t pop_bytes;
} flb_layout;

/* ---- lifecycle ---------------------------------------------------------- */
const char* flb_last_error(void);
int flb_abi_version(void);
/* The layout flb_block_create would give `desc`, checked the same way; host
 * only (no device call), e.g. to size device memory before creating blocks.
 * BlockField's storage of the reference is one plain array per field
 * (data_struct/field.h:1419-1575), so it has no counterpart to this query. */
int flb_block_layout(const flb_block_desc* desc, flb_layout* out);
int flb_block_create(const flb_block_desc* desc, flb_block* out);
int flb_block_destroy(flb_block b);
int flb_get_info(flb_block b, flb_block_info* info);
int flb_synchronize(flb_block b);

/* ---- set-up ------------------------------------------------------------- */
int flb_set_tasks(flb_block b, const flb_task* tasks, int n);      /* n <= 8 */
int flb_set_halfway_bc(flb_block b, const flb_hwbc* bcs, int n);   /* n <= 4 */

/* ---- coupled lattices (a flow block and its temperature / concentration
 * blocks, FreeLB's BlockLatManagerCoupling, data_struct/block_lattice.h:436-649) */

/* The scalar block is advected by the flow block's VELOCITY field: FreeLB's
 * &NSLattice.getField<VELOCITY>() passed to the scalar lattice's constructor
 * (examples/cazsblock2d/cazsblock2d.cpp:240-251). The blocks must cover the
 * same cells. flow = NULL detaches (the scalar block then uses its own
 * VELOCITY field). Both blocks get RHO / VELOCITY fields if they have none. */
int flb_share_velocity(flb_block scalar, flb_block flow);
/* Adds force::Buoyancy<NSCELL, CELL1> (lbm/force.h:160-181) run by
 * BlockLatManagerCoupling::ApplyCellDynamics with a CoupledTaskSelector on
 * `mask` (cazsblock2d.cpp:326-333, 385-392): every FLB_DYN_BGK_SCALAR_FORCE
 * cell of `flow` whose flag matches takes SCALARFORCE += gbeta (RHO_scalar -
 * rho_init) (GBETA and RHOINIT are the scalar lattice's constant fields),
 * starting from 0 each step (SCALARFORCE.InitValue(0)). The force is formed
 * inside the flow block's collision, from the scalar field as the scalar
 * block's last step left it. At most 2 couplings per flow block, applied in
 * the order added. */
int flb_add_buoyancy(flb_block flow, flb_block scalar, uint8_t mask, double gbeta, double rho_init);
int flb_clear_buoyancy(flb_block flow);
/* MRT relaxation rates, q values; NaN marks the shear-viscosity rates that
 * take omega. NULL restores the defaults. */
int flb_set_mrt_rates(flb_block b, const double* rates, int q);
int flb_set_omega(flb_block b, double omega);
/* Constant body-force density of FLB_DYN_BGK_FORCE cells, d values: FreeLB's
 * CONSTFORCE data field read by force::ConstForce (lbm/force.h:86-107). */
int flb_set_force(flb_block b, const double* force, int d);
int flb_upload_flags(flb_block b, const uint8_t* flags);           /* N bytes */
int flb_set_field_rho(flb_block b, const double* rho);             /* N values  */
int flb_set_field_u(flb_block b, const double* u);                 /* d*N values, SoA */
/* RHO and VELOCITY fields back to the host, rhou[(d+1)][N] (rho 1, u 0 where
 * the block has none): the fields FLB_DYN_BGK_WRITE cells wrote
 * (cell.get<RHO>() / get<VELOCITY>() of the reference). */
int flb_download_fields(flb_block b, double* rhou);
/* Convergence residual (BlockLattice::EnableToleranceRho/U, getToleranceRho/U,
 * getTolRho/getTolU(shift), data_struct/block_lattice.hh:239-367):
 * flb_enable_tolerance keeps the current RHO and VELOCITY fields as the old
 * ones; flb_tolerance(which 0 = RHO, 1 = VELOCITY) returns the maximum over
 * the cells [shift, n - shift) per axis of |rho - rho_old| (of
 * max_a |u_a - u_old_a|) and makes the current field the old one. */
int flb_enable_tolerance(flb_block b);
int flb_tolerance(flb_block b, int which, int shift, double* res);
/* Read-out moment::rhoU<CELL, true> on the cells with (flag & mask) != 0
 * (lbm/moment.h:359-377, run before output as examples/cavity3d/cavity3d.cpp:192-194
 * does): their rho and u go to the RHO / VELOCITY fields. */
int flb_store_moments(flb_block b, uint8_t mask);
/* FreeLB's VTK output, vtmo::vtmWriter::WriteBinary(step) (io/vtm_writer.h:288-430,
 * io/vti_writer.h:424-987): per block <dir>/<name>_T<step>_B<k>.vti (ImageData,
 * Origin = min_centres[3k..] + strips[k] * voxels[k], strips[k] layers cut off
 * each side; point data as base64 binary arrays: the RHO field named rho_name,
 * the VELOCITY field named u_name, the flags named flag_name, any of them NULL
 * to leave out) and <dir>/<name><step>.vtm listing them; step < 0 drops the
 * step from the names (WriteBinary()). strips may be NULL (no cut). */
int flb_write_vtm(const flb_block* blocks, int n, const double* min_centres, const double* voxels,
                  const int* strips, const char* dir, const char* name, int step, const char* rho_name,
                  const char* u_name, const char* flag_name);

/* ---- populations --------------------------------------------------------- */
/* Natural-layout (post-stream) populations [q][N]. _f32 serves the F32, F16S,
 * BF16S and F16C storage kinds, _f64 serves F64. Upload resets the layout to
 * natural; download resolves the AA layout of the current parity. Values of
 * cells outside the visited interior download as 0. */
int flb_upload_pops_f32(flb_block b, const float* pops);
int flb_upload_pops_f64(flb_block b, const double* pops);
int flb_download_pops_f32(flb_block b, float* pops);
int flb_download_pops_f64(flb_block b, double* pops);
/* f_i = feq_i(rho, u) (equilibrium::SecondOrder) from host rhou[d+1][N]. */
int flb_init_equilibrium(flb_block b, const double* rhou);

/* ---- the hot path ------------------------------------------------------- */
/* Collide + stream + halfway fix-up of one time step, fused (AA pattern),
 * gated by the block's refinement level as BlockLatticeManager::ApplyCellDynamics(
 * std::int64_t count, BFM) gates it (data_struct/block_lattice.hh:751-764): the
 * block steps when count % 2^(max_level - level) == 0 and the call is a no-op
 * otherwise. A block of level 0 in a hierarchy of one level steps at every count. */
int flb_apply_cell_dynamics(flb_block b, int64_t count);
/* Completes the step begun by flb_apply_cell_dynamics (AA parity flip, O(1)),
 * with the same gate (BlockLatticeManager::Stream(std::int64_t count),
 * block_lattice.hh:737-742). */
int flb_stream(flb_block b, int64_t count);
/* The block's level and the finest level of its hierarchy (BasicBlock::getLevel
 * and BlockLatticeManager::getMaxLevel, data_struct/block_lattice.h:364-372);
 * default 0 / 0. */
int flb_set_level(flb_block b, int level, int max_level);
/* Halfway managers are fused into flb_apply_cell_dynamics; this checks call
 * order only and keeps the reference's three-call step shape. */
int flb_apply_halfway_bc(flb_block b);
/* n complete steps (apply + stream) back to back. */
int flb_step(flb_block b, int n);
/* NaN / Inf watchdog of flb_step: after every `every` steps of the block
 * (counted by its step counter), the FP64 mass / momentum reduction over the
 * updated cells runs and flb_step returns FLB_ERR_NONFINITE, naming the step,
 * when any of the four sums is not finite. 0 (the default) switches it off.
 * The reference has none: it exits only on CUDA errors (utils/cuda_device.h:36-44). */
int flb_set_watchdog(flb_block b, int every);

/* ---- read-out ------------------------------------------------------------ */
/* rho and u of cells with (flag & mask) != 0, 0 elsewhere, into host
 * rhou[d+1][N] (float for F32/F16S/BF16S storage, double for F64). */
int flb_moments(flb_block b, uint8_t mask, void* rhou);
/* FP64 totals over cells with (flag & mask) != 0: out = {mass, Px, Py, Pz}. */
int flb_reduce_mass_momentum(flb_block b, uint8_t mask, double out[4]);
/* The same reduction for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No reference
 * counterpart. It calls no allocating routine: its partial sums live in a
 * file-scope device array of the library (32 MiB per device, held until the
 * process ends), so it takes blocks of at most 2^28 host cells (FLB_ERR_ARG
 * beyond), and it writes the four sums (mass, momentum x/y/z; the same bits as
 * flb_reduce_mass_momentum) to the caller's device memory d_out, enqueued on the
 * block's stream without synchronising. Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (different streams, or threads) race. */
int flb_reduce_mass_momentum_noalloc(flb_block b, uint8_t mask, double* d_out);

/* ---- batches ----------------------------------------------------------------
 * Many blocks of one lattice and storage kind stepped by one launch per time
 * step: a device table of the blocks' kernel arguments, walked by a grid of
 * workgroups that is either one workgroup per group of wavefronts or a
 * persistent grid; steps can be recorded once as a hipGraph and replayed.
 * Replaces the per-block launch loop of
 *   BlockLatticeManager::ApplyCellDynamics / Stream  data_struct/block_lattice.hh:744-778
 *   BlockLatticeManager::CuDevApplyCellDynamics       data_struct/block_lattice.hh:919-936
 * The blocks stay usable on their own (set-up, read-out) and must outlive the
 * batch; settings changed on a block after batch creation are picked up at
 * the next batch call. All blocks must be at the same parity with no Stream
 * pending. Work on the batch is ordered after earlier work on the blocks'
 * streams and before later work on them. */
typedef struct flb_batch_s* flb_batch;
typedef struct flb_halo_s* flb_halo;
typedef struct {
  int waves_per_group;  /* wavefronts per workgroup, 1..4 */
  int grid;             /* workgroups; 0 = one per group of wavefronts (no persistence) */
} flb_launch_cfg;
int flb_batch_create(const flb_block* blocks, int n, void* stream, flb_batch* out);
int flb_batch_destroy(flb_batch t);
int flb_batch_info(flb_batch t, int* n_blocks, int64_t* n_cells, flb_launch_cfg* cfg);
int flb_batch_set_launch(flb_batch t, const flb_launch_cfg* cfg);
/* n steps of every block, one launch each (enqueued, not synchronised). */
int flb_batch_step(flb_batch t, int n);
/* Records n steps from the current parity as a hipGraph (nothing runs). */
int flb_batch_capture(flb_batch t, int n);
/* Launches the recorded graph: n more steps. The blocks must be at the
 * parity of the capture (an even n replays any number of times). */
int flb_batch_replay(flb_batch t);
int flb_batch_synchronize(flb_batch t);
/* Times `steps` (even) launches of each candidate shape (1, 2, 4 wavefronts
 * per workgroup; one workgroup per group, or 1, 2, 4, 8 x 4 wavefronts per
 * compute unit, persistent) and keeps the fastest. Populations are restored
 * afterwards, so the blocks' state is unchanged. */
int flb_batch_autotune(flb_batch t, int steps, flb_launch_cfg* best, double* best_ms_per_step);

/* With a halo set, every batch step is followed by its exchange (and a
 * recorded graph holds both). NULL removes it. Drops a recorded graph. */
int flb_batch_set_halo(flb_batch t, flb_halo h);

/* ---- block halo ----------------------------------------------------------------
 * Blocks of one domain, each with a one-cell overlap layer that mirrors the
 * neighbouring blocks' cells: FreeLB's same-level communication
 * (BlockLatticeManager::normalCommunicate / normalFullCommunicate,
 * data_struct/block_lattice.hh:1669-1703; cell pairs as BlockGeometry3D::InitComm,
 * geometry/block_geometry3d.hh:531-605). origins[3k..3k+2] is the global
 * index of block k's local cell (0,0,0) (its overlap corner); periodic axes of
 * the domain (extent cells long) wrap from block to block. Blocks are not
 * periodic themselves, share the halo's stream, and need
 * flb_set_ghost_overlap (cells of the overlap layer may then match tasks;
 * they are not updated, only mirrored). flb_halo_exchange follows every step
 * (all blocks at one parity): in the AA pattern it fills the ghost slots the
 * next odd step reads, or returns the values an odd step pushed into them. */
int flb_set_ghost_overlap(flb_block b, int enable);
int flb_halo_create(const flb_block* blocks, int n, const int* origins, const int* periodic,
                    const int* extent, void* stream, flb_halo* out);
int flb_halo_destroy(flb_halo h);
int flb_halo_info(flb_halo h, int* n_items, int64_t* n_ghost_cells);
int flb_halo_exchange(flb_halo h);

/* ---- two-level refinement -------------------------------------------------------
 * A coarse block (level 0) and a fine block at twice its resolution (level 1)
 * whose interior covers the coarse cells [origin, origin + m) per axis; the
 * fine extent is 2m + 4 (two fine ghost layers, one coarse cell, as FreeLB's
 * level-1 blocks built with useHelperOlap = false, geometry/block_geometry3d.hh:174-190)
 * and fine local cell 2 + 2 (c - origin) + b, b in {0, 1}, lies in coarse cell c.
 * The fine block has overlap 1 (its outer ghost layer; the inner one is updated),
 * the coarse cells inside the patch beyond its outer coarse layer match no task,
 * and both blocks keep RHO / VELOCITY fields written by FLB_DYN_BGK_WRITE cells.
 * Replaces FreeLB's BlockLatticeManager::FullCommunicate(count) for the pair
 * (data_struct/block_lattice.hh:1579-1591): at even counts the outer coarse layer
 * of the patch takes the average of its 2^d fine cells converted by
 * RefineConverter::getPopC (averCommunicate, block_lattice.hh:1724-1755,
 * lbm/unit_converter.h:657-659); at every count the inner fine ghost layer takes
 * the 2^d-point interpolation of the coarse cells around it (weights
 * parallel/communicator.h:575-600) converted by getPopF (intpCommunicate,
 * block_lattice.hh:1757-1811, unit_converter.h:646-648); feq is
 * Equilibrium::SecondOrder of the averaged / interpolated fields. Both blocks
 * run on the refinement's stream. */
typedef struct flb_refine_s* flb_refine;
int flb_refine_create(flb_block coarse, flb_block fine, const int* origin, void* stream, flb_refine* out);
int flb_refine_destroy(flb_refine r);  /* the blocks stay */
int flb_refine_info(flb_refine r, int64_t* n_average, int64_t* n_interp);
/* Levels of a pair inside a deeper hierarchy: the coarse block is at
 * coarse_level, the fine one at coarse_level + 1, the finest level is
 * max_level (default 0 / 1). The average into the coarse block runs at counts
 * with count % 2^(max_level - coarse_level) == 0 and the interpolation into the
 * fine block at counts with count % 2^(max_level - coarse_level - 1) == 0,
 * each gated by its receiving block's level as FreeLB's AverageCommunicate(count)
 * and InterpolateCommunicate(count) gate them (block_lattice.hh:1520-1561). */
int flb_refine_set_levels(flb_refine r, int coarse_level, int max_level);
/* AverageCommunicate(count) / InterpolateCommunicate(count) of the pair alone,
 * for managers that run every pair's average before any interpolation. */
int flb_refine_average(flb_refine r, int64_t count);
int flb_refine_interpolate(flb_refine r, int64_t count);
/* FullCommunicate(count) of the pair: the average, then the interpolation. */
int flb_refine_communicate(flb_refine r, int64_t count);
/* n counts of the refined loop from count0 (examples/cavref2d/cavref2d.cpp:221-228)
 * for a pair that is the whole hierarchy (levels 0 / 1): the fine block steps
 * every count, the coarse block on even counts, then flb_refine_communicate(count). */
int flb_refine_run(flb_refine r, int64_t count0, int n);

/* ---- differentiation -------------------------------------------------------
 * A D3Q19 box (nx*ny*nz cells, FreeLB index order, any axis periodic; cells
 * outside a non-periodic axis and cells marked solid are halfway
 * bounce-back walls) stepped by BGK with a constant body-force density G
 * (BGKForce with ConstForce: lbm/collision.h:61-90, lbm/force.h:29-36,
 * lbm/moment.h:401-438) and a Brinkman drag -alpha*s(x)*rho*u per cell,
 * s in [0, 1], taken implicitly in the force-corrected velocity
 * u = (sum c f + G/2) / (rho (1 + alpha s / 2)). The objective is
 * J = (1/T) sum_{t=1..T} sum over the plane x = outlet_x of sum_i c_ix f_i,
 * summed in FP64. flb_adjoint_backward returns dJ/ds from the hand-written
 * discrete adjoint of the step, recomputing forward states from
 * checkpoints. No reference counterpart (FreeLB has no adjoint). A box holds
 * at most 2^32 bytes per population array (2^30 cells in FP32, 2^29 in
 * FP64); create fails with FLB_ERR_ARG beyond that. */
typedef struct flb_adjoint_s* flb_adjoint;
typedef struct {
  int nx, ny, nz;
  int periodic[3];
  int storage;            /* FLB_STORE_F32 or FLB_STORE_F64 */
  double omega;
  double force[3];        /* body-force density G */
  double brinkman_alpha;  /* drag coefficient at s = 1 */
  int outlet_x;           /* plane of the objective */
  void* stream;           /* hipStream_t; NULL = the null stream */
} flb_adjoint_desc;
int flb_adjoint_create(const flb_adjoint_desc* desc, flb_adjoint* out);
int flb_adjoint_destroy(flb_adjoint h);
int flb_adjoint_set_solid(flb_adjoint h, const uint8_t* solid);       /* N bytes, nonzero = solid */
int flb_adjoint_set_fraction(flb_adjoint h, const double* s);         /* N values */
int flb_adjoint_set_fraction_at(flb_adjoint h, int64_t cell, double s);
/* f^0 = feq(rho, u) from host rhou[4][N]. */
int flb_adjoint_init_equilibrium(flb_adjoint h, const double* rhou);
/* T steps from f^0; *J gets the objective. The adjoint steps read a forward
 * state only through its moments (rho - 1 and u: 4 values per cell, against
 * 19 populations). checkpoint_every: k > 0 keeps every k-th state and, in the
 * backward pass, recomputes each segment of k states from its checkpoint,
 * recording their moments (memory: ceil(T/k) + 6 states and k moment
 * records of 4/19 of a state each; k near sqrt(T) balances memory and
 * recompute); 0 records the moments of every state during the forward pass
 * and recomputes nothing (memory: 5 states and T moment records); -1 records
 * nothing (objective only, e.g. for finite differences). */
int flb_adjoint_forward(flb_adjoint h, int steps, int checkpoint_every, double* J);
/* dJ/ds for every cell (0 on solid cells) after a recorded forward run. */
int flb_adjoint_backward(flb_adjoint h, double* grad);
/* Final forward state f^T, [19][N]. */
int flb_adjoint_state(flb_adjoint h, double* pops);
/* Device time of the last forward and backward pass (ms, HIP events). */
int flb_adjoint_timing(flb_adjoint h, double* forward_ms, double* backward_ms);

/* Many boxes of any sizes stepped together: one launch per time step (and
 * per adjoint step) covers the cells of every member. Members keep their own
 * fields and buffers; all must use the same storage kind and device. After a
 * batch forward run each member holds its own result exactly as after
 * flb_adjoint_forward (its objective, final state and recorded run, so
 * flb_adjoint_backward on one member works too). No reference counterpart. */
typedef struct flb_adjoint_batch_s* flb_adjoint_batch;
int flb_adjoint_batch_create(const flb_adjoint* members, int n, void* stream, flb_adjoint_batch* out);
int flb_adjoint_batch_destroy(flb_adjoint_batch b);  /* the members stay */
/* J[n]: each member's objective. */
int flb_adjoint_batch_forward(flb_adjoint_batch b, int steps, int checkpoint_every, double* J);
/* grad: every member's dJ/ds one after another (sum of the members' N). */
int flb_adjoint_batch_backward(flb_adjoint_batch b, double* grad);
int flb_adjoint_batch_timing(flb_adjoint_batch b, double* forward_ms, double* backward_ms);

/* ---- diagnostics ------------------------------------------------------------ */
/* Device copy bandwidth: a float4 copy kernel over `bytes` (rounded down to a
 * multiple of 16) between two fresh buffers, `reps` timed launches; *gbs gets
 * (read + written bytes) / mean launch time. No reference counterpart: it
 * calibrates the roofline figures the benchmark reports. */
int flb_bandwidth_probe(size_t bytes, int reps, double* gbs);
/* The probe in the other shapes a streaming kernel can take: mode 0 copy,
 * 1 read-only, 2 write-only; `unroll` 16-byte accesses in flight per lane
 * (1, 2, 4 or 8); plain or non-temporal accesses; `wg_per_cu` workgroups per
 * CU striding over the buffer (0: one workgroup per tile of unroll * 256
 * elements, each visited once). *gbs counts read + written bytes. No
 * reference counterpart: it measures the achievable HBM rate. */
int flb_bandwidth_probe_ex(int mode, int unroll, int nontemporal, int wg_per_cu, size_t bytes, int reps,
                           double* gbs);

#ifdef __cplusplus
}
#endif

#endif /* FLB_ABI_H */
