// Differentially heated cavity on the GPU through the C++ adapter, written
// the way FreeLB's coupled examples drive their lattices
// (examples/cazsblock2d/cazsblock2d.cpp:225-409, without the cellular
// automaton): a D3Q19 flow lattice whose cells collide with
// BGKForce<forcerhoU<CELL, ScalarForce, true>, SecondOrder, ScalarForce>, a
// D3Q7 temperature lattice advected by the flow's VELOCITY field, and the
// buoyancy coupling force::Buoyancy run through BlockLatManagerCoupling.
// Flags: 1 void, 2 fluid, 4 adiabatic wall, 8 hot wall (x = 1), 16 cold wall
// (x = n). Per step:
//   Buoyancy.ApplyCellDynamics<BuoySel>(flags);
//   NS.ApplyCellDynamics<NSTask>(flags); NS.Stream(); NS_BM.Apply();
//   TH.ApplyCellDynamics<THTask>(flags); TH.Stream(); TH_BM.Apply();
//
//   natconv3d n steps [out.bin]
//
// tau 0.6 (flow) and 0.7 (temperature), gbeta 1e-4, T hot 1.5, cold 0.5,
// initial 1. Prints MLUPS (cell updates of both lattices per cell and step
// counted once); with out.bin writes the flow then the temperature
// populations (float32, natural layout, q*N each).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "flb/block_lattice.hpp"

using T = float;
using LatF = flb::D3Q19<T>;
using LatT = flb::D3Q7<T>;
using NSCELL = flb::Cell<T, LatF>;
using THCELL = flb::Cell<T, LatT>;

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s n steps [out.bin]\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[1]), steps = std::atoi(argv[2]);
  constexpr std::uint8_t VoidFlag = 1, AABBFlag = 2, BBFlag = 4, HotFlag = 8, ColdFlag = 16;
  constexpr std::uint8_t AllFlag = AABBFlag | BBFlag | HotFlag | ColdFlag;
  const int m = n + 2;
  const size_t N = size_t(m) * m * m;
  std::vector<std::uint8_t> flags(N, VoidFlag);
  std::vector<double> T0(N, 1.0);
  for (int z = 1; z <= n; ++z)
    for (int y = 1; y <= n; ++y)
      for (int x = 1; x <= n; ++x) {
        const size_t id = size_t(x) + size_t(m) * (size_t(y) + size_t(m) * z);
        const bool wall = x == 1 || x == n || y == 1 || y == n || z == 1 || z == n;
        flags[id] = x == 1 ? HotFlag : (x == n ? ColdFlag : (wall ? BBFlag : AABBFlag));
        if (x == 1) T0[id] = 1.5;
        if (x == n) T0[id] = 0.5;
      }

  using NSLAT = flb::BlockLattice<T, LatF>;
  using THLAT = flb::BlockLattice<T, LatT>;
  NSLAT NS(m, m, m, 1.0 / 0.6);
  THLAT TH(m, m, m, 1.0 / 0.7);
  NS.setRhoField(std::vector<double>(N, 1.0));
  NS.setVelocityField(std::vector<double>(3 * N, 0.0));
  NS.initUniform(1.0);
  TH.setRhoField(T0);  // the TEMP field; hot / cold cells keep their wall value
  TH.shareVelocity(NS);
  TH.setGBeta(1e-4);
  TH.setRhoInit(1.0);
  {
    std::vector<double> rhou(4 * N, 0.0);
    for (size_t k = 0; k < N; ++k) rhou[k] = T0[k];
    TH.initEquilibrium(rhou);
  }

  flb::BBLikeFixedBlockBdManager<flb::bounceback::normal<NSCELL>, NSLAT> NS_BB(
      "NS_BB", NS, std::uint8_t(BBFlag | HotFlag | ColdFlag), VoidFlag);
  flb::BlockBoundaryManager NS_BM(&NS_BB);
  flb::BBLikeFixedBlockBdManager<flb::bounceback::normal<THCELL>, THLAT> TH_BB("TH_BB", TH, BBFlag, VoidFlag);
  flb::BBLikeFixedBlockBdManager<flb::bounceback::anti_simple<THCELL>, THLAT> TH_D(
      "TH_D", TH, std::uint8_t(HotFlag | ColdFlag), VoidFlag);
  flb::BlockBoundaryManager TH_BM(&TH_BB, &TH_D);

  using SF = flb::force::ScalarForce<NSCELL>;
  using NSBulk = flb::collision::BGKForce<flb::moment::forcerhoU<NSCELL, SF, true>,
                                          flb::equilibrium::SecondOrder<NSCELL>, SF>;
  using NSTask = flb::TaskSelector<std::uint8_t, NSCELL, flb::tmp::Key_TypePair<AllFlag, NSBulk>>;
  using THBulk = flb::collision::BGK<
      flb::moment::MomentaTuple<flb::moment::rho<THCELL, true>, flb::moment::useFieldU<THCELL>>,
      flb::equilibrium::SecondOrder<THCELL>>;
  using THWall = flb::collision::BGK<flb::moment::useFieldrhoU<THCELL>, flb::equilibrium::SecondOrder<THCELL>>;
  using THTask = flb::TaskSelector<std::uint8_t, THCELL,
                                   flb::tmp::Key_TypePair<std::uint8_t(AABBFlag | BBFlag), THBulk>,
                                   flb::tmp::Key_TypePair<std::uint8_t(HotFlag | ColdFlag), THWall>>;
  using BuoySel = flb::CoupledTaskSelector<std::uint8_t, NSCELL, THCELL,
                                           flb::tmp::Key_TypePair<AllFlag, flb::force::Buoyancy<NSCELL, THCELL>>>;
  flb::BlockLatManagerCoupling<T, LatF, LatT> Buoyancy(NS, TH);

  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; ++s) {
    Buoyancy.ApplyCellDynamics<BuoySel>(flags);
    NS.ApplyCellDynamics<NSTask>(flags);
    NS.Stream();
    NS_BM.Apply();
    TH.ApplyCellDynamics<THTask>(flags);
    TH.Stream();
    TH_BM.Apply();
  }
  flb::check(flb_synchronize(NS.handle()));
  flb::check(flb_synchronize(TH.handle()));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::printf("natconv3d n=%d steps=%d: %.1f MLUPS\n", n, steps, double(n) * n * n * steps / ms / 1e3);
  if (argc > 3) {
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 3;
    const std::vector<T> pf = NS.populations(), pt = TH.populations();
    std::fwrite(pf.data(), sizeof(T), pf.size(), f);
    std::fwrite(pt.data(), sizeof(T), pt.size(), f);
    std::fclose(f);
  }
  return 0;
}
