/*
 * CPU restatement of FreeLB's per-block lattice update -- TEST INFRASTRUCTURE.
 * See flb_oracle.h. Plain C11 + OpenMP; the precision-generic body lives in
 * flb_oracle_impl.h and is instantiated for float and double below.
 */
#include "flb_oracle.h"

#include <math.h>
#include <omp.h>
#include <string.h>

/* D3Q19 velocity set in the reference order (lattice_set.h:259-296). */
static const int C19[19][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                               {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                               {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
                               {1, 0, -1}, {-1, 0, 1}, {0, 1, -1}, {0, -1, 1}};
static const int OPP19[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
/* D2Q9 (lattice_set.h:168-194), z component zero. */
static const int C9[9][3] = {{0, 0, 0}, {1, 0, 0},   {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                             {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}};
static const int OPP9[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};

/* Advection-diffusion sets D2Q5 (lattice_set.h:144-155) and D3Q7 (196-207). */
static const int C5[5][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
static const int OPP5[5] = {0, 2, 1, 4, 3};
static const int C7[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
static const int OPP7[7] = {0, 2, 1, 4, 3, 6, 5};

static const int (*lat_c(int q))[3] { return q == 19 ? C19 : (q == 9 ? C9 : (q == 7 ? C7 : C5)); }
static const int* lat_opp(int q) { return q == 19 ? OPP19 : (q == 9 ? OPP9 : (q == 7 ? OPP7 : OPP5)); }
static int lat_d(int q) { return (q == 19 || q == 7) ? 3 : 2; }
static int lat_ok(int q) { return q == 19 || q == 9 || q == 7 || q == 5; }

/* D2Q9 MRT transform (collisionMRT.h:55-90), Lallemand & Luo 2000. */
static const int M9[9][9] = {{1, 1, 1, 1, 1, 1, 1, 1, 1},      {-4, -1, -1, -1, -1, 2, 2, 2, 2},
                             {4, -2, -2, -2, -2, 1, 1, 1, 1},  {0, 1, -1, 0, 0, 1, -1, 1, -1},
                             {0, -2, 2, 0, 0, 1, -1, 1, -1},   {0, 0, 0, 1, -1, 1, -1, -1, 1},
                             {0, 0, 0, -2, 2, 1, -1, -1, 1},   {0, 1, 1, -1, -1, 0, 0, 0, 0},
                             {0, 0, 0, 0, 0, 1, 1, -1, -1}};
/* Its inverse as numerator/denominator pairs (collisionMRT.h:67-77). */
static const int INVM9_N[9][9] = {{1, -1, 1, 0, 0, 0, 0, 0, 0},     {1, -1, -1, 1, -1, 0, 0, 1, 0},
                                  {1, -1, -1, -1, 1, 0, 0, 1, 0},   {1, -1, -1, 0, 0, 1, -1, -1, 0},
                                  {1, -1, -1, 0, 0, -1, 1, -1, 0},  {1, 1, 1, 1, 1, 1, 1, 0, 1},
                                  {1, 1, 1, -1, -1, -1, -1, 0, 1},  {1, 1, 1, 1, 1, -1, -1, 0, -1},
                                  {1, 1, 1, -1, -1, 1, 1, 0, -1}};
static const int INVM9_D[9][9] = {{9, 9, 9, 1, 1, 1, 1, 1, 1},     {9, 36, 18, 6, 6, 1, 1, 4, 1},
                                  {9, 36, 18, 6, 6, 1, 1, 4, 1},   {9, 36, 18, 1, 1, 6, 6, 4, 1},
                                  {9, 36, 18, 1, 1, 6, 6, 4, 1},   {9, 18, 36, 6, 12, 6, 12, 1, 4},
                                  {9, 18, 36, 6, 12, 6, 12, 1, 4}, {9, 18, 36, 6, 12, 6, 12, 1, 4},
                                  {9, 18, 36, 6, 12, 6, 12, 1, 4}};

/* D3Q19 MRT: no reference form exists (collisionMRT.h defines D2Q9 only).
 * This is the orthogonal moment basis of d'Humieres, Ginzburg, Krafczyk,
 * Lallemand & Luo, Phil. Trans. R. Soc. A 360 (2002) 437, rows in its order
 * (rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz, pxz,
 * mx, my, mz), evaluated on the reference velocity order. Parity unpinned by
 * the reference; pinned by MRT == BGK when every rate equals omega, by exact
 * orthogonality, and by mass/momentum conservation. */
static int M19[19][19];
static int M19_NORM[19];
static int M19_ODD[19]; /* row odd in c: M(j, opp i) = -M(j, i); else even */
static int m19_ready = 0;

static void m19_init(void) {
  if (m19_ready) return;
  for (int i = 0; i < 19; ++i) {
    const int x = C19[i][0], y = C19[i][1], z = C19[i][2];
    const int c2 = x * x + y * y + z * z;
    M19[0][i] = 1;
    M19[1][i] = 19 * c2 - 30;
    M19[2][i] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
    M19[3][i] = x;
    M19[4][i] = (5 * c2 - 9) * x;
    M19[5][i] = y;
    M19[6][i] = (5 * c2 - 9) * y;
    M19[7][i] = z;
    M19[8][i] = (5 * c2 - 9) * z;
    M19[9][i] = 3 * x * x - c2;
    M19[10][i] = (3 * c2 - 5) * (3 * x * x - c2);
    M19[11][i] = y * y - z * z;
    M19[12][i] = (3 * c2 - 5) * (y * y - z * z);
    M19[13][i] = x * y;
    M19[14][i] = y * z;
    M19[15][i] = x * z;
    M19[16][i] = (y * y - z * z) * x;
    M19[17][i] = (z * z - x * x) * y;
    M19[18][i] = (x * x - y * y) * z;
  }
  for (int j = 0; j < 19; ++j) {
    int n = 0;
    for (int i = 0; i < 19; ++i) n += M19[j][i] * M19[j][i];
    M19_NORM[j] = n;
    int odd = M19[j][0] == 0;
    for (int i = 1; i < 19; i += 2) odd = odd && M19[j][i + 1] == -M19[j][i];
    M19_ODD[j] = odd;
  }
  m19_ready = 1;
}

/* Default D3Q19 rates (d'Humieres et al. 2002, section 5); NaN marks the
 * shear-viscosity moments that relax with omega (or the Smagorinsky omega). */
static double mrt19_default_rate(int j) {
  static const double s[19] = {0.0,  1.19, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0,  1.2, NAN,
                               1.4,  NAN,  1.4, NAN, NAN, NAN, 1.98, 1.98, 1.98};
  return s[j];
}

#define REAL float
#define SFX f32
#define SQRT sqrtf
#include "flb_oracle_mrt.h"
#include "flb_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT

#define REAL double
#define SFX f64
#define SQRT sqrt
#include "flb_oracle_mrt.h"
#include "flb_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT

int flbo_run_f32(const flbo_lattice* L, float* pops, float* scratch, int steps) {
  m19_init();
  return run_impl_f32(L, pops, scratch, steps);
}
int flbo_run_f64(const flbo_lattice* L, double* pops, double* scratch, int steps) {
  m19_init();
  return run_impl_f64(L, pops, scratch, steps);
}
int flbo_run_coupled_f32(const flbo_lattice* flow, float* pops, float* scratch, const flbo_lattice* const* scalar,
                         float* const* spops, float* const* sscratch, const flbo_buoyancy* buoy, int nbuoy,
                         int nscalar, int steps) {
  return run_coupled_impl_f32(flow, pops, scratch, scalar, spops, sscratch, buoy, nbuoy, nscalar, steps);
}
int flbo_run_coupled_f64(const flbo_lattice* flow, double* pops, double* scratch,
                         const flbo_lattice* const* scalar, double* const* spops, double* const* sscratch,
                         const flbo_buoyancy* buoy, int nbuoy, int nscalar, int steps) {
  return run_coupled_impl_f64(flow, pops, scratch, scalar, spops, sscratch, buoy, nbuoy, nscalar, steps);
}
void flbo_rhou_f32(const flbo_lattice* L, const float* pops, uint8_t mask, float* rhou) {
  rhou_impl_f32(L, pops, mask, rhou);
}
void flbo_rhou_f64(const flbo_lattice* L, const double* pops, uint8_t mask, double* rhou) {
  rhou_impl_f64(L, pops, mask, rhou);
}
void flbo_init_feq_f32(const flbo_lattice* L, const float* rhou, float* pops) {
  init_feq_impl_f32(L, rhou, pops);
}
void flbo_init_feq_f64(const flbo_lattice* L, const double* rhou, double* pops) {
  init_feq_impl_f64(L, rhou, pops);
}

/* Cavity flags. Block extent (Ni+2)(Nj+2)(Nk+2) with cell centres at
 * (i - 1/2) dx (BasicBlock::getVoxel, basic_geometry.h:310-312; overlap 1 at
 * block_geometry3d.hh:156-171). AABB(0, N) with the epsilon test of
 * basic_geometry.hh:38-48 selects i in [1, N] -> AABB flag 2 over void 1.
 * SetupBoundary (block_geometry3d.hh:76-102) marks every inner cell with a
 * lattice neighbour outside the AABB -> BB flag 4, i.e. index 1 or N on any
 * axis. The lid AABB [1, N-1]^2 x [N-1, N] selects i, j in [2, N-1] on the
 * top layer k = Nk (y = Nj in 2-D), where BB cells become BBMW flag 8. */
void flbo_cavity_flags(int Ni, int Nj, int Nk, uint8_t* flags) {
  const int d3 = Nk > 0;
  const int nx = Ni + 2, ny = Nj + 2, nz = d3 ? Nk + 2 : 1;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t id = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
        const int in = x >= 1 && x <= Ni && y >= 1 && y <= Nj && (!d3 || (z >= 1 && z <= Nk));
        uint8_t f = 1;
        if (in) {
          const int edge = x == 1 || x == Ni || y == 1 || y == Nj || (d3 && (z == 1 || z == Nk));
          f = edge ? 4 : 2;
          if (edge) {
            const int lid = d3 ? (z == Nk && x >= 2 && x <= Ni - 1 && y >= 2 && y <= Nj - 1)
                               : (y == Nj && x >= 2 && x <= Ni - 1);
            if (lid) f = 8;
          }
        }
        flags[id] = f;
      }
}

/* BaseConverter::ConvertFromRT / Converter / getLatticeU with dx = 1
 * (unit_converter.h:180-217, 163): deltaT = (RT - 0.5) * cs2 * dx * dx / visc,
 * Conv_U = dx / deltaT, OMEGA = 1 / RT, u_lattice = U_phys / Conv_U. */
void flbo_convert_from_rt(float RT, float visc, float u_phys, float* omega, float* u_lattice) {
  volatile float cs2 = 1.0f / 3.0f;
  volatile float dx = 1.0f;
  volatile float dt = (RT - 0.5f) * cs2 * dx * dx / visc;
  volatile float conv_u = dx / dt;
  *omega = 1.0f / RT;
  *u_lattice = u_phys / conv_u;
}
