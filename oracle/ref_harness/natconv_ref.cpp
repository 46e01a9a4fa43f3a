// Fixture writer for a coupled flow / advection-diffusion case, built against
// FreeLB's own headers (test infrastructure, compiled from /root/reference/src
// by oracle/Makefile).
//
// A differentially heated cavity assembled from FreeLB's parts the way
// examples/cazsblock2d/cazsblock2d.cpp:225-409 assembles its coupled lattices
// (without the cellular automaton):
//   * flow lattice (D2Q9 / D3Q19): fields RHO, VELOCITY, POP, SCALARFORCE;
//     every non-void cell collides with
//     BGKForce<forcerhoU<CELL, ScalarForce, true>, SecondOrder, ScalarForce>
//     (lbm/collision.h:61-90, lbm/moment.h:401-480, lbm/force.h:109-133);
//     halfway bounceback::normal on all wall cells;
//   * temperature lattice (D2Q5 / D3Q7): fields TEMP, POP, RHOINIT, GBETA and
//     the flow lattice's VELOCITY by reference; bulk and adiabatic wall cells
//     collide with BGK<MomentaTuple<rho<CELL, true>, useFieldU>, SecondOrder>,
//     the hot and cold wall cells with BGK<useFieldrhoU, SecondOrder> (their
//     TEMP field holds the wall temperature); halfway bounceback::normal on the
//     adiabatic walls and bounceback::anti_simple
//     (boundary/bounce_back_boundary.h:183-193) on the hot and cold walls;
//   * coupling: BlockLatManagerCoupling (data_struct/block_lattice.h:436-649)
//     with force::Buoyancy<NSCELL, THCELL> (lbm/force.h:160-181) on every
//     non-void cell, after the SCALARFORCE field is cleared, as the main loop
//     of cazsblock2d.cpp:385-409 does.
//
// Usage:
//   natconv_ref <d2q9|d3q19> Ni Nj Nk steps RT_flow RT_temp gbeta T_hot T_cold T_init out.bin
// (Nk ignored for d2q9; numbers are parsed with strtof.)
//
// Flags: Void 1, AABB 2, bounce-back wall 4, hot wall 8 (x = 1 layer), cold
// wall 16 (x = Ni layer); the hot / cold layers exclude the edges of the wall
// (cells on two walls stay adiabatic).
//
// Output layout (little endian):
//   char[8]  "FLBNC001"
//   int32    d, q_flow, q_temp, nx, ny, nz, steps
//   float32  omega_flow, omega_temp, gbeta, T_hot, T_cold, T_init
//   uint8    flags[N]
//   float32  initial flow pops [q_flow][N], temperature pops [q_temp][N], TEMP[N]
//   float32  final flow pops [q_flow][N], temperature pops [q_temp][N]
//   float32  final RHO[N], VELOCITY[d][N], TEMP[N], SCALARFORCE[N]

#include "freelb.h"
#include "freelb.hh"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using T = float;

int Thread_Num = 1;

namespace {

constexpr std::uint8_t VoidFlag = 1, AABBFlag = 2, BBFlag = 4, HotFlag = 8, ColdFlag = 16;
constexpr std::uint8_t AllFlag = AABBFlag | BBFlag | HotFlag | ColdFlag;

struct Params {
  int Ni, Nj, Nk, steps;
  T RTf, RTt, gbeta, Th, Tc, T0;
  std::string out;
};

template <typename Lat, typename Cell, typename BL>
void put_pops(FILE* f, BL& blat, std::size_t N) {
  std::vector<T> buf(N);
  for (unsigned i = 0; i < Lat::q; ++i) {
    for (std::size_t id = 0; id < N; ++id) {
      Cell cell(id, blat);
      buf[id] = cell[i];
    }
    std::fwrite(buf.data(), sizeof(T), N, f);
  }
}

template <typename F>
void put_field(FILE* f, std::size_t N, F&& get) {
  std::vector<T> buf(N);
  for (std::size_t id = 0; id < N; ++id) buf[id] = get(id);
  std::fwrite(buf.data(), sizeof(T), N, f);
}

template <typename LatF, typename LatT, int D, typename Geo, typename BoxT>
int run(const Params& p, Geo& geo, const BoxT& cavity, const BoxT& hot, const BoxT& cold, int nz) {
  BaseConverter<T> ConvF(LatF::cs2);
  ConvF.ConvertFromRT(T(1), p.RTf, T(1), T(p.Ni), T(0.01), T(0.01));
  BaseConverter<T> ConvT(LatT::cs2);
  ConvT.ConvertFromRT(T(1), p.RTt, T(1), T(p.Ni), T(0.01), T(0.01));

  BlockFieldManager<FLAG, T, D> FlagFM(geo, VoidFlag);
  FlagFM.forEach(cavity, [&](FLAG& field, std::size_t id) { field.SetField(id, AABBFlag); });
  FlagFM.template SetupBoundary<LatF>(cavity, BBFlag);
  FlagFM.forEach(hot, [&](FLAG& field, std::size_t id) {
    if (util::isFlag(field.get(id), BBFlag)) field.SetField(id, HotFlag);
  });
  FlagFM.forEach(cold, [&](FLAG& field, std::size_t id) {
    if (util::isFlag(field.get(id), BBFlag)) field.SetField(id, ColdFlag);
  });

  using NSFIELDS = TypePack<RHO<T>, VELOCITY<T, D>, POP<T, LatF::q>, SCALARFORCE<T>>;
  using NSCELL = Cell<T, LatF, NSFIELDS>;
  ValuePack NSInit(T(1), Vector<T, D>{}, T{}, T{});
  using NSLM = BlockLatticeManager<T, LatF, NSFIELDS>;
  NSLM NSLattice(geo, NSInit, ConvF);

  using THFIELDS = TypePack<TEMP<T>, POP<T, LatT::q>, RHOINIT<T>, GBETA<T>>;
  using THREFS = TypePack<VELOCITY<T, D>>;
  using THPACK = TypePack<THFIELDS, THREFS>;
  using THCELL = Cell<T, LatT, typename ExtractFieldPack<THPACK>::mergedpack>;
  ValuePack THInit(p.T0, T{}, p.T0, p.gbeta);
  using THLM = BlockLatticeManager<T, LatT, THPACK>;
  THLM THLattice(geo, THInit, ConvT, &NSLattice.template getField<VELOCITY<T, D>>());
  THLattice.template getField<TEMP<T>>().forEach(FlagFM, HotFlag,
                                                 [&](auto& field, std::size_t id) { field.SetField(id, p.Th); });
  THLattice.template getField<TEMP<T>>().forEach(FlagFM, ColdFlag,
                                                 [&](auto& field, std::size_t id) { field.SetField(id, p.Tc); });

  auto& nsb = NSLattice.getBlockLat(0);
  auto& thb = THLattice.getBlockLat(0);
  const std::size_t N = nsb.getN();
  // temperature populations at rest equilibrium of the cell's TEMP field
  for (std::size_t id = 0; id < N; ++id) {
    THCELL cell(id, thb);
    const T t = cell.template get<TEMP<T>>();
    for (unsigned i = 0; i < LatT::q; ++i) cell[i] = t * latset::w<LatT>(i);
  }

  using FM = BlockFieldManager<FLAG, T, D>;
  BBLikeFixedBlockBdManager<bounceback::normal<NSCELL>, NSLM, FM> NS_BB("NS_BB", NSLattice, FlagFM,
                                                                        std::uint8_t(BBFlag | HotFlag | ColdFlag),
                                                                        VoidFlag);
  BBLikeFixedBlockBdManager<bounceback::normal<THCELL>, THLM, FM> TH_BB("TH_BB", THLattice, FlagFM, BBFlag,
                                                                        VoidFlag);
  BBLikeFixedBlockBdManager<bounceback::anti_simple<THCELL>, THLM, FM> TH_D(
    "TH_D", THLattice, FlagFM, std::uint8_t(HotFlag | ColdFlag), VoidFlag);

  using SF = force::ScalarForce<NSCELL>;
  using NSBulk = collision::BGKForce<moment::forcerhoU<NSCELL, SF, true>, equilibrium::SecondOrder<NSCELL>, SF>;
  using NSTask = TaskSelector<std::uint8_t, NSCELL, tmp::Key_TypePair<AllFlag, NSBulk>>;
  using THBulk = collision::BGK<moment::MomentaTuple<moment::rho<THCELL, true>, moment::useFieldU<THCELL>>,
                                equilibrium::SecondOrder<THCELL>>;
  using THWall = collision::BGK<moment::useFieldrhoU<THCELL>, equilibrium::SecondOrder<THCELL>>;
  using THTask = TaskSelector<std::uint8_t, THCELL, tmp::Key_TypePair<std::uint8_t(AABBFlag | BBFlag), THBulk>,
                              tmp::Key_TypePair<std::uint8_t(HotFlag | ColdFlag), THWall>>;
  using BuoyTask = tmp::Key_TypePair<AllFlag, force::Buoyancy<NSCELL, THCELL>>;
  using BuoySel = CoupledTaskSelector<std::uint8_t, NSCELL, THCELL, BuoyTask>;
  BlockLatManagerCoupling Buoyancy(NSLattice, THLattice);

  FILE* f = std::fopen(p.out.c_str(), "wb");
  if (!f) {
    std::perror("fopen");
    return 2;
  }
  const auto& flags = FlagFM.getBlockField(0).getField(0);
  std::fwrite("FLBNC001", 1, 8, f);
  const int hdr[7] = {D, int(LatF::q), int(LatT::q), nsb.getNx(), nsb.getNy(), nz, p.steps};
  std::fwrite(hdr, sizeof(int), 7, f);
  const T prm[6] = {nsb.getOmega(), thb.getOmega(), p.gbeta, p.Th, p.Tc, p.T0};
  std::fwrite(prm, sizeof(T), 6, f);
  std::vector<std::uint8_t> fl(N);
  for (std::size_t id = 0; id < N; ++id) fl[id] = flags[id];
  std::fwrite(fl.data(), 1, N, f);
  put_pops<LatF, NSCELL>(f, nsb, N);
  put_pops<LatT, THCELL>(f, thb, N);
  put_field(f, N, [&](std::size_t id) { return THCELL(id, thb).template get<TEMP<T>>(); });

  for (int s = 0; s < p.steps; ++s) {
    NSLattice.template getField<SCALARFORCE<T>>().InitValue(T{});
    Buoyancy.template ApplyCellDynamics<BuoySel>(FlagFM);
    NSLattice.template ApplyCellDynamics<NSTask>(FlagFM);
    NSLattice.Stream();
    NS_BB.Apply();
    THLattice.template ApplyCellDynamics<THTask>(FlagFM);
    THLattice.Stream();
    TH_BB.Apply();
    TH_D.Apply();
  }

  put_pops<LatF, NSCELL>(f, nsb, N);
  put_pops<LatT, THCELL>(f, thb, N);
  put_field(f, N, [&](std::size_t id) { return NSCELL(id, nsb).template get<RHO<T>>(); });
  for (int c = 0; c < D; ++c)
    put_field(f, N, [&](std::size_t id) { return NSCELL(id, nsb).template get<VELOCITY<T, D>>()[c]; });
  put_field(f, N, [&](std::size_t id) { return THCELL(id, thb).template get<TEMP<T>>(); });
  put_field(f, N, [&](std::size_t id) { return NSCELL(id, nsb).template get<SCALARFORCE<T>>(); });
  std::fclose(f);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 13) {
    std::fprintf(stderr,
                 "usage: %s <d2q9|d3q19> Ni Nj Nk steps RT_flow RT_temp gbeta T_hot T_cold T_init out.bin\n",
                 argv[0]);
    return 2;
  }
  Params p;
  const std::string lat = argv[1];
  p.Ni = std::atoi(argv[2]);
  p.Nj = std::atoi(argv[3]);
  p.Nk = std::atoi(argv[4]);
  p.steps = std::atoi(argv[5]);
  p.RTf = std::strtof(argv[6], nullptr);
  p.RTt = std::strtof(argv[7], nullptr);
  p.gbeta = std::strtof(argv[8], nullptr);
  p.Th = std::strtof(argv[9], nullptr);
  p.Tc = std::strtof(argv[10], nullptr);
  p.T0 = std::strtof(argv[11], nullptr);
  p.out = argv[12];
  if (lat == "d3q19") {
    AABB<T, 3> cavity(Vector<T, 3>{}, Vector<T, 3>(T(p.Ni), T(p.Nj), T(p.Nk)));
    AABB<T, 3> hot(Vector<T, 3>(T(0), T(1), T(1)), Vector<T, 3>(T(1), T(p.Nj - 1), T(p.Nk - 1)));
    AABB<T, 3> cold(Vector<T, 3>(T(p.Ni - 1), T(1), T(1)), Vector<T, 3>(T(p.Ni), T(p.Nj - 1), T(p.Nk - 1)));
    BlockGeometry3D<T> geo(p.Ni, p.Nj, p.Nk, 1, cavity, T(1));
    return run<D3Q19<T>, D3Q7<T>, 3>(p, geo, cavity, hot, cold, p.Nk + 2);
  }
  if (lat == "d2q9") {
    AABB<T, 2> cavity(Vector<T, 2>(T(0), T(0)), Vector<T, 2>(T(p.Ni), T(p.Nj)));
    AABB<T, 2> hot(Vector<T, 2>(T(0), T(1)), Vector<T, 2>(T(1), T(p.Nj - 1)));
    AABB<T, 2> cold(Vector<T, 2>(T(p.Ni - 1), T(1)), Vector<T, 2>(T(p.Ni), T(p.Nj - 1)));
    BlockGeometry2D<T> geo(p.Ni, p.Nj, 1, cavity, T(1));
    return run<D2Q9<T>, D2Q5<T>, 2>(p, geo, cavity, hot, cold, 1);
  }
  std::fprintf(stderr, "unknown lattice %s\n", lat.c_str());
  return 2;
}
