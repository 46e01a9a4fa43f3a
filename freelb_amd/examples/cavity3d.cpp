// Lid-driven cavity on the GPU through the C++ adapter, written the way
// FreeLB's benchmarks/cavity3d/cavity3d.cpp drives its CPU lattice: flags
// 1 void / 2 fluid / 4 wall / 8 moving lid, a TaskSelector of Key_TypePairs,
// ApplyCellDynamics + Stream (+ BM.Apply in the halfway mode) per step.
//
//   cavity3d <fw|hw|force|batch> N steps [out.bin]
//
// fw: walls collide as BounceBack / BounceBackMovingWall (the published run);
// hw: every cell BGK, halfway bounce-back and moving wall after streaming;
// force: as hw with BGKForce<forcerhoU, SecondOrder, ConstForce> and the
//   constant force (1e-4, -5e-5, 2e-5);
// batch: three hw cavities stepped together by a BlockLatticeManager (one
//   launch per step); out.bin holds the first one.
// Prints MLUPS; with out.bin, writes rho and u (float32, 4*N values, SoA).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "flb/block_lattice.hpp"

using T = float;
using LatSet = flb::D3Q19<T>;
using CELL = flb::Cell<T, LatSet>;

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <fw|hw|force|batch> N steps [out.bin]\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  const int n = std::atoi(argv[2]);
  const int steps = std::atoi(argv[3]);
  constexpr std::uint8_t VoidFlag = 1, AABBFlag = 2, BouncebackFlag = 4, BBMovingWallFlag = 8;

  // BaseConverter::ConvertFromRT with dx = 1, RT 0.53, nu 0.01, U 0.1 (float)
  const T RT = 0.53f, visc = 0.01f, U = 0.1f;
  const T dt = (RT - T(0.5)) * (T(1) / T(3)) * T(1) * T(1) / visc;
  const T ulat = U / (T(1) / dt);
  const double omega = double(T(1) / RT);

  const int m = n + 2;
  const size_t N = size_t(m) * m * m;
  std::vector<std::uint8_t> flags(N, VoidFlag);
  for (int z = 1; z <= n; ++z)
    for (int y = 1; y <= n; ++y)
      for (int x = 1; x <= n; ++x) {
        const bool edge = x == 1 || x == n || y == 1 || y == n || z == 1 || z == n;
        const bool lid = z == n && x >= 2 && x <= n - 1 && y >= 2 && y <= n - 1;
        flags[size_t(x) + size_t(m) * (size_t(y) + size_t(m) * z)] =
            lid ? BBMovingWallFlag : (edge ? BouncebackFlag : AABBFlag);
      }
  std::vector<double> u(3 * N, 0.0);
  for (size_t k = 0; k < N; ++k)
    if (flags[k] == BBMovingWallFlag) u[k] = ulat;

  flb::BlockLattice<T, LatSet> NSLattice(m, m, m, omega);
  NSLattice.setRhoField(std::vector<double>(N, 1.0));
  NSLattice.setVelocityField(u);
  NSLattice.initUniform(1.0);

  using BGKrhoU = flb::collision::BGK<flb::moment::rhoU<CELL>, flb::equilibrium::SecondOrder<CELL>>;
  double ms = 0.0;
  if (mode == "fw") {
    using BulkTask = flb::tmp::Key_TypePair<AABBFlag, BGKrhoU>;
    using BBTask = flb::tmp::Key_TypePair<BouncebackFlag, flb::collision::BounceBack<CELL>>;
    using BBMVTask = flb::tmp::Key_TypePair<BBMovingWallFlag, flb::collision::BounceBackMovingWall<CELL>>;
    using NSTask = flb::TaskSelector<std::uint8_t, CELL, BulkTask, BBTask, BBMVTask>;
    NSLattice.ApplyCellDynamics<NSTask>(flags);  // first call uploads flags and tasks
    NSLattice.Stream();
    flb::check(flb_synchronize(NSLattice.handle()));
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 1; s < steps; ++s) {
      NSLattice.ApplyCellDynamics<NSTask>(flags);
      NSLattice.Stream();
    }
    flb::check(flb_synchronize(NSLattice.handle()));
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  } else if (mode == "batch") {
    using AllTask =
        flb::tmp::Key_TypePair<std::uint8_t(AABBFlag | BouncebackFlag | BBMovingWallFlag), BGKrhoU>;
    using NSTask = flb::TaskSelector<std::uint8_t, CELL, AllTask>;
    using LAT = flb::BlockLattice<T, LatSet>;
    flb::BlockLattice<T, LatSet> B1(m, m, m, omega), B2(m, m, m, omega);
    std::vector<LAT*> lats{&NSLattice, &B1, &B2};
    std::vector<std::unique_ptr<flb::BBLikeFixedBlockBdManager<flb::bounceback::normal<CELL>, LAT>>> bb;
    std::vector<std::unique_ptr<flb::BBLikeFixedBlockBdManager<flb::bounceback::movingwall<CELL>, LAT>>> mw;
    for (LAT* L : lats) {
      if (L != &NSLattice) {
        L->setRhoField(std::vector<double>(N, 1.0));
        L->setVelocityField(u);
        L->initUniform(1.0);
      }
      bb.emplace_back(new flb::BBLikeFixedBlockBdManager<flb::bounceback::normal<CELL>, LAT>(
          "NS_BB", *L, BouncebackFlag, VoidFlag));
      mw.emplace_back(new flb::BBLikeFixedBlockBdManager<flb::bounceback::movingwall<CELL>, LAT>(
          "NS_BBMW", *L, BBMovingWallFlag, VoidFlag));
      L->ApplyCellDynamics<NSTask>(flags);  // first step per block: uploads flags and tasks
      L->Stream();
    }
    flb::BlockLatticeManager<T, LatSet> Manager(lats);
    Manager.synchronize();
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 1; s < steps; ++s) {
      Manager.ApplyCellDynamics();
      Manager.Stream();
    }
    Manager.synchronize();
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / 3.0;
  } else if (mode == "force") {
    using ForceS = flb::force::ConstForce<CELL>;
    using BGKF = flb::collision::BGKForce<flb::moment::forcerhoU<CELL, ForceS>,
                                          flb::equilibrium::SecondOrder<CELL>, ForceS>;
    using AllTask = flb::tmp::Key_TypePair<std::uint8_t(AABBFlag | BouncebackFlag | BBMovingWallFlag), BGKF>;
    using NSTask = flb::TaskSelector<std::uint8_t, CELL, AllTask>;
    using LAT = flb::BlockLattice<T, LatSet>;
    NSLattice.setConstForce({double(std::strtof("1e-4", nullptr)), double(std::strtof("-5e-5", nullptr)),
                             double(std::strtof("2e-5", nullptr))});
    flb::BBLikeFixedBlockBdManager<flb::bounceback::normal<CELL>, LAT> NS_BB("NS_BB", NSLattice,
                                                                            BouncebackFlag, VoidFlag);
    flb::BBLikeFixedBlockBdManager<flb::bounceback::movingwall<CELL>, LAT> NS_BBMW(
        "NS_BBMW", NSLattice, BBMovingWallFlag, VoidFlag);
    flb::BlockBoundaryManager BM(&NS_BB, &NS_BBMW);
    flb::check(flb_synchronize(NSLattice.handle()));
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) {
      NSLattice.ApplyCellDynamics<NSTask>(flags);
      NSLattice.Stream();
      BM.Apply();
    }
    flb::check(flb_synchronize(NSLattice.handle()));
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  } else {
    using AllTask =
        flb::tmp::Key_TypePair<std::uint8_t(AABBFlag | BouncebackFlag | BBMovingWallFlag), BGKrhoU>;
    using NSTask = flb::TaskSelector<std::uint8_t, CELL, AllTask>;
    using LAT = flb::BlockLattice<T, LatSet>;
    flb::BBLikeFixedBlockBdManager<flb::bounceback::normal<CELL>, LAT> NS_BB("NS_BB", NSLattice,
                                                                            BouncebackFlag, VoidFlag);
    flb::BBLikeFixedBlockBdManager<flb::bounceback::movingwall<CELL>, LAT> NS_BBMW(
        "NS_BBMW", NSLattice, BBMovingWallFlag, VoidFlag);
    flb::BlockBoundaryManager BM(&NS_BB, &NS_BBMW);
    NSLattice.ApplyCellDynamics<NSTask>(flags);
    NSLattice.Stream();
    BM.Apply();
    flb::check(flb_synchronize(NSLattice.handle()));
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 1; s < steps; ++s) {
      NSLattice.ApplyCellDynamics<NSTask>(flags);
      NSLattice.Stream();
      BM.Apply();
    }
    flb::check(flb_synchronize(NSLattice.handle()));
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  const double mlups = steps > 1 ? double(n) * n * n * (steps - 1) / ms / 1e3 : 0.0;
  std::printf("cavity3d %s N=%d steps=%d: %.1f MLUPS (fluid cells)\n", mode.c_str(), n, steps, mlups);
  if (argc > 4) {
    const std::vector<T> rhou = NSLattice.moments(std::uint8_t(~VoidFlag));
    FILE* f = std::fopen(argv[4], "wb");
    if (!f) return 3;
    std::fwrite(rhou.data(), sizeof(T), rhou.size(), f);
    std::fclose(f);
  }
  return 0;
}
