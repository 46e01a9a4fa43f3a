// Two-level refined lid-driven cavity on the GPU through the C++ adapter,
// written the way FreeLB's examples/cavref2d/cavref2d.cpp drives its refined
// lattice (in 3D): a coarse n^3 cavity (flags 1 void / 2 fluid / 4 wall /
// 8 moving lid, 16 covered by the patch), a fine block at twice the
// resolution over the coarse cells [o, o + m) per axis, fluid cells
// colliding with BGK<rhoU<CELL, true>> (moments written to the fields that
// the exchange reads), wall cells with BGK<useFieldrhoU>, halfway walls, and
//   ApplyCellDynamics(count) ; Stream(count) ; FullCommunicate(count)
// per count (the coarse level steps on even counts).
//
//   cavref3d n o m counts [out.bin]
//
// RT 0.6, nu 0.01, U 0.05 through ConvertFromRT (float), the fine omega by
// RefineConverter::getOmegaF. Prints MLUPS; with out.bin writes the coarse
// then the fine populations (float32, natural layout, q*N each).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "flb/block_lattice.hpp"

using T = float;
using LatSet = flb::D3Q19<T>;
using CELL = flb::Cell<T, LatSet>;

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s n o m counts [out.bin]\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[1]), o = std::atoi(argv[2]), m = std::atoi(argv[3]);
  const int counts = std::atoi(argv[4]);
  constexpr std::uint8_t VoidFlag = 1, AABBFlag = 2, BouncebackFlag = 4, BBMovingWallFlag = 8, Covered = 16;

  const T RT = 0.6f, visc = 0.01f, U = 0.05f;
  const T dt = (RT - T(0.5)) * (T(1) / T(3)) * T(1) * T(1) / visc;
  const T ulat = U / (T(1) / dt);
  const T omega_c = T(1) / RT;
  const T omega_f = T(1) / (T(2) / omega_c - T(0.5));

  const int cn = n + 2;
  const size_t NC = size_t(cn) * cn * cn;
  std::vector<std::uint8_t> cflags(NC, VoidFlag);
  for (int z = 1; z <= n; ++z)
    for (int y = 1; y <= n; ++y)
      for (int x = 1; x <= n; ++x) {
        const bool edge = x == 1 || x == n || y == 1 || y == n || z == 1 || z == n;
        const bool lid = z == n && x >= 2 && x <= n - 1 && y >= 2 && y <= n - 1;
        const bool covered = x > o && x < o + m - 1 && y > o && y < o + m - 1 && z > o && z < o + m - 1;
        cflags[size_t(x) + size_t(cn) * (size_t(y) + size_t(cn) * z)] =
            covered ? Covered : (lid ? BBMovingWallFlag : (edge ? BouncebackFlag : AABBFlag));
      }
  std::vector<double> uc(3 * NC, 0.0);
  for (size_t k = 0; k < NC; ++k)
    if (cflags[k] == BBMovingWallFlag) uc[k] = ulat;
  const int fn = 2 * m + 4;
  const size_t NF = size_t(fn) * fn * fn;
  std::vector<std::uint8_t> fflags(NF, Covered);
  for (int z = 1; z < fn - 1; ++z)
    for (int y = 1; y < fn - 1; ++y)
      for (int x = 1; x < fn - 1; ++x) fflags[size_t(x) + size_t(fn) * (size_t(y) + size_t(fn) * z)] = AABBFlag;

  using LAT = flb::BlockLattice<T, LatSet>;
  LAT Coarse(cn, cn, cn, omega_c), Fine(fn, fn, fn, omega_f);
  Coarse.setRhoField(std::vector<double>(NC, 1.0));
  Coarse.setVelocityField(uc);
  Coarse.initUniform(1.0);
  Fine.setRhoField(std::vector<double>(NF, 1.0));
  Fine.setVelocityField(std::vector<double>(3 * NF, 0.0));
  Fine.initUniform(1.0);
  flb::BBLikeFixedBlockBdManager<flb::bounceback::normal<CELL>, LAT> NS_BB("NS_BB", Coarse, BouncebackFlag,
                                                                           VoidFlag);
  flb::BBLikeFixedBlockBdManager<flb::bounceback::movingwall<CELL>, LAT> NS_BBMW("NS_BBMW", Coarse,
                                                                                 BBMovingWallFlag, VoidFlag);

  using BulkTask = flb::tmp::Key_TypePair<AABBFlag,
      flb::collision::BGK<flb::moment::rhoU<CELL, true>, flb::equilibrium::SecondOrder<CELL>>>;
  using WallTask = flb::tmp::Key_TypePair<std::uint8_t(BouncebackFlag | BBMovingWallFlag),
      flb::collision::BGK<flb::moment::useFieldrhoU<CELL>, flb::equilibrium::SecondOrder<CELL>>>;
  using NSTask = flb::TaskSelector<std::uint8_t, CELL, BulkTask, WallTask>;

  const int origin[3] = {o, o, o};
  flb::RefinedPair<T, LatSet> Pair(Coarse, Fine, origin);
  const auto t0 = std::chrono::steady_clock::now();
  for (int count = 0; count < counts; ++count) {
    Pair.ApplyCellDynamics<NSTask>(count, cflags, fflags);
    Pair.Stream(count);
    Pair.FullCommunicate(count);
  }
  flb::check(flb_synchronize(Coarse.handle()));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const double updates = double(n) * n * n * (counts / 2) + double(2 * m + 2) * (2 * m + 2) * (2 * m + 2) * counts;
  std::printf("cavref3d n=%d patch=%d counts=%d: %.1f MLUPS\n", n, 2 * m, counts, updates / ms / 1e3);
  if (argc > 5) {
    FILE* f = std::fopen(argv[5], "wb");
    if (!f) return 3;
    const std::vector<T> pc = Coarse.populations(), pf = Fine.populations();
    std::fwrite(pc.data(), sizeof(T), pc.size(), f);
    std::fwrite(pf.data(), sizeof(T), pf.size(), f);
    std::fclose(f);
  }
  return 0;
}
