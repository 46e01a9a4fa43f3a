/*
 * CPU restatement of FreeLB's per-block lattice update -- TEST INFRASTRUCTURE.
 *
 * This library is the oracle: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it to check and to time against. Nothing in the
 * product (freelb_amd/, include/) links it or falls back to it.
 *
 * It restates the reference algorithm with the reference's own data
 * semantics -- structure-of-arrays populations, periodic-shift (PS)
 * streaming over the flattened block with wrap modulo N, first-match flag
 * dispatch, halfway bounce-back applied after streaming from the pre-stream
 * values -- and the reference's floating-point expression order, so that an
 * FP32 run reproduces FreeLB's FP32 run bit for bit when both are built with
 * floating-point contraction switched off.
 *
 * Each function names the reference file:line it follows (paths relative to
 * the FreeLB source tree).
 */
#ifndef FLB_ORACLE_H
#define FLB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Dynamics kinds; numeric values agree with include/flb_abi.h. */
enum {
  FLBO_DYN_NONE = 0,
  FLBO_DYN_BGK = 1,            /* collision::BGK<moment::rhoU, SecondOrder>       collision.h:32-59 */
  FLBO_DYN_BGK_FIELD = 2,      /* collision::BGK<moment::useFieldrhoU, ...>       moment.h:71-90    */
  FLBO_DYN_BOUNCEBACK = 3,     /* collision::BounceBack                           collision.ur.h:62-93 */
  FLBO_DYN_BOUNCEBACK_MW = 4,  /* collision::BounceBackMovingWall                 collision.ur.h:130-187 */
  FLBO_DYN_MRT = 5,            /* collision::MRT_Feq_RhoU                         collisionMRT.h:132-190 */
  FLBO_DYN_SMAG_BGK = 6,       /* collision::SmagorinskyBGK                       collisionLES.h:29-67 */
  FLBO_DYN_SMAG_MRT = 7,       /* MRT with the Smagorinsky effective shear rate (no reference form) */
  FLBO_DYN_BGK_FORCE = 8,      /* collision::BGKForce<forcerhoU, SecondOrder, ConstForce>
                                  collision.h:61-90, moment.ur.h:564-574/639-663, force.ur.h */
  FLBO_DYN_BGK_WRITE = 9,      /* collision::BGK<moment::rhoU<CELL, true>, SecondOrder>: the moments
                                  go to the RHO / VELOCITY fields (moment.h:359-377) */
  FLBO_DYN_BGK_SCALAR_FORCE = 10, /* BGKForce<forcerhoU<CELL, ScalarForce, true>, SecondOrder, ScalarForce>
                                  collision.h:61-90, force.ur.h:332-353/584-616; F = sforce[id] */
  FLBO_DYN_SCALAR_BGK = 11,    /* D2Q5/D3Q7 BGK<MomentaTuple<rho<CELL, true>, useFieldU>, SecondOrder> */
  FLBO_DYN_SCALAR_BGK_FIELD = 12 /* D2Q5/D3Q7 BGK<useFieldrhoU, SecondOrder> */
};

enum { FLBO_HWBC_NORMAL = 0, FLBO_HWBC_MOVING = 1, FLBO_HWBC_ANTI = 2 };

typedef struct {
  uint8_t mask;      /* first task whose (flag & mask) != 0 runs (tmp.h:283-309) */
  uint8_t dynamics;  /* FLBO_DYN_* */
  uint8_t pad[2];
  double smagorinsky; /* C_s for the Smagorinsky kinds */
} flbo_task;

typedef struct {
  uint8_t cell_mask; /* boundary cells: (flag & cell_mask) != 0 (basic_boundary.h:241-262) */
  uint8_t void_mask; /* a neighbour with (flag & void_mask) != 0 is a wall link (basic_boundary.h:198-212) */
  uint8_t kind;      /* FLBO_HWBC_* */
  uint8_t pad;
} flbo_hwbc;

typedef struct {
  int q;              /* 9 (D2Q9), 19 (D3Q19), 5 (D2Q5) or 7 (D3Q7) */
  int nx, ny, nz;     /* storage extents including the overlap layer; nz == 1 for D2Q9 */
  int overlap;        /* overlap (ghost) width, 1 in every reference driver */
  const uint8_t* flags;     /* nx*ny*nz */
  int ntasks;
  flbo_task tasks[8];
  int nbc;
  flbo_hwbc bcs[4];
  double omega;       /* 1/tau; the FP32 path rounds it to float first */
  double* rho_field;  /* N values or NULL (then 1): RHO field of moving walls / useFieldrhoU;
                         FLBO_DYN_BGK_WRITE cells store their rho here */
  double* u_field;    /* d*N values (SoA) or NULL (then 0): VELOCITY field (likewise) */
  const double* mrt_rates;  /* q relaxation rates for the MRT kinds, or NULL for the defaults;
                               a NaN entry marks the shear rates that take omega */
  int nthreads;       /* OpenMP threads, <= 0 for the default */
  double force[3];    /* CONSTFORCE data field of BGKForce (force density, lattice units) */
  double* sforce;     /* SCALARFORCE field (N) of FLBO_DYN_BGK_SCALAR_FORCE cells, or NULL (0) */
  const double* u_ext; /* scalar lattices: VELOCITY field (d*N) advecting the scalar (the flow
                          lattice's), or NULL for u_field */
} flbo_lattice;

/* One force::Buoyancy coupling: SCALARFORCE += gbeta (scalar_rho - rho_init)
 * on flow cells whose flag matches mask (lbm/force.h:160-181). */
typedef struct {
  const double* scalar_rho;  /* the scalar lattice's RHO field (N), the TEMP / CONC of FreeLB */
  uint8_t mask;
  uint8_t pad[7];
  double gbeta, rho_init;
} flbo_buoyancy;

/* Run `steps` reference time steps (collide, PS stream, halfway BCs) in place
 * on pops[q][N] (logical, natural layout). scratch must hold q*N values. */
int flbo_run_f32(const flbo_lattice* L, float* pops, float* scratch, int steps);
int flbo_run_f64(const flbo_lattice* L, double* pops, double* scratch, int steps);

/* Coupled flow + scalar lattices (examples/cazsblock2d/cazsblock2d.cpp:385-409):
 * per step clear flow->sforce, add the buoyancy couplings, step the flow
 * lattice, then each of the ns scalar lattices. */
int flbo_run_coupled_f32(const flbo_lattice* flow, float* pops, float* scratch, const flbo_lattice* const* scalar,
                         float* const* spops, float* const* sscratch, const flbo_buoyancy* buoy, int nbuoy,
                         int nscalar, int steps);
int flbo_run_coupled_f64(const flbo_lattice* flow, double* pops, double* scratch,
                         const flbo_lattice* const* scalar, double* const* spops, double* const* sscratch,
                         const flbo_buoyancy* buoy, int nbuoy, int nscalar, int steps);

/* moment::rhoUImpl (moment.ur.h:461-517): rhou[(d+1)][N] = rho, u_x, u_y[, u_z]
 * on every cell whose flag matches `mask` (0 elsewhere). */
void flbo_rhou_f32(const flbo_lattice* L, const float* pops, uint8_t mask, float* rhou);
void flbo_rhou_f64(const flbo_lattice* L, const double* pops, uint8_t mask, double* rhou);

/* Equilibrium initialisation f_i = feq_i(rho, u) per cell with SecondOrder
 * (equilibrium.ur.h:65-94, 165-214); rho/u are (d+1)*N SoA. */
void flbo_init_feq_f32(const flbo_lattice* L, const float* rhou, float* pops);
void flbo_init_feq_f64(const flbo_lattice* L, const double* rhou, double* pops);

/* Two-level refinement, FreeLB's FullCommunicate(count) for one coarse and one
 * fine lattice (block_lattice.hh:1579-1591). The fine lattice covers the
 * coarse cells [origin, origin + m) per axis with two fine ghost layers
 * (extent 2m + 4); fine cell 2 + 2 (c - origin) + b lies in coarse cell c.
 * average: averCommunicate (block_lattice.hh:1724-1755) into the outer
 *   coarse layer of the patch, sources in getRefinedCellIdx order
 *   (basic_geometry.hh:1211-1239), getAverage (communicator.h:608-620),
 *   getPopC (unit_converter.h:657-659) with the fine omega;
 * interpolate: intpCommunicate / popIntpCommunicate (block_lattice.hh:1757-1811)
 *   into both fine ghost layers, stencils as AddtoSharedIntpComm
 *   (block_geometry2d.hh:988-1059), weights getIntpWeight (communicator.h:575-600),
 *   getInterpolation (communicator.h:622-645), getPopF (unit_converter.h:646-648)
 *   with the coarse omega.
 * feq is Equilibrium::SecondOrder (equilibrium.h:49-75) of the averaged or
 * interpolated RHO / VELOCITY fields. Populations are the natural post-stream
 * arrays [q][N] of flbo_run. */
void flbo_refine_average_f32(const flbo_lattice* C, float* popsC, const flbo_lattice* F, const float* popsF,
                             const int* origin);
void flbo_refine_average_f64(const flbo_lattice* C, double* popsC, const flbo_lattice* F, const double* popsF,
                             const int* origin);
void flbo_refine_interpolate_f32(const flbo_lattice* C, const float* popsC, const flbo_lattice* F, float* popsF,
                                 const int* origin);
void flbo_refine_interpolate_f64(const flbo_lattice* C, const double* popsC, const flbo_lattice* F,
                                 double* popsF, const int* origin);

/* Flag layout of the lid-driven cavity drivers: BlockGeometry3D/2D with one
 * block and overlap 1, FlagFM.forEach(cavity, AABB), SetupBoundary(cavity, BB),
 * forEach(toplid, BB -> BBMW) (benchmarks/cavity3d/cavity3d.cpp:121-142,
 * block_geometry3d.hh:76-102). flags has (Ni+2)(Nj+2)(Nk+2) entries (Nk = 0
 * selects 2-D with (Ni+2)(Nj+2)). */
void flbo_cavity_flags(int Ni, int Nj, int Nk, uint8_t* flags);

/* BaseConverter::ConvertFromRT with dx = 1 (unit_converter.h:209-217), in
 * float: omega = 1/RT and the lattice lid speed U_phys / (dx / dt). */
void flbo_convert_from_rt(float RT, float visc, float u_phys, float* omega, float* u_lattice);

#ifdef __cplusplus
}
#endif

#endif
