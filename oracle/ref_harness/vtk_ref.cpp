// Fixture writer for FreeLB's VTK output of a lid-driven cavity, built against
// FreeLB's own headers (test infrastructure, compiled from /root/reference/src
// by oracle/Makefile).
//
// It runs the halfway cavity of cavity_ref.cpp ("hw": every non-void cell
// BGK<rhoU>, bounceback::normal / movingwall after Stream) for `steps` steps,
// applies the read-out moment::rhoU<CELL, true> on the non-void cells (as
// examples/cavity3d/cavity3d.cpp:192-194, 226-227 does before writing) and
// writes RHO, VELOCITY and the flag field with vtmo::vtmWriter
// (io/vtm_writer.h:288-430, io/vti_writer.h:424-987) into
// ./vtkoutput/vtidata/ of the working directory.
//
// Usage:
//   vtk_ref <d3q19|d2q9> Ni Nj Nk steps overlap_threshold name
// (Nk is ignored for d2q9; RT 0.53, nu 0.01, U 0.1 as cavity_ref.)

#include "freelb.h"
#include "freelb.hh"

#include <cstdio>
#include <cstdlib>
#include <string>

using T = float;

int Thread_Num = 1;

namespace {

constexpr std::uint8_t VoidFlag = std::uint8_t(1);
constexpr std::uint8_t AABBFlag = std::uint8_t(2);
constexpr std::uint8_t BouncebackFlag = std::uint8_t(4);
constexpr std::uint8_t BBMovingWallFlag = std::uint8_t(8);

template <typename LatSet, int D, typename Geo, typename BoxT>
int run(Geo& geo, const BoxT& cavity, const BoxT& toplid, int ni, int steps, int olth, const std::string& name) {
  BaseConverter<T> BaseConv(LatSet::cs2);
  BaseConv.ConvertFromRT(T(1), std::strtof("0.53", nullptr), T(0.001), T(ni), std::strtof("0.1", nullptr),
                         std::strtof("0.01", nullptr));
  using FIELDS = TypePack<RHO<T>, VELOCITY<T, LatSet::d>, POP<T, LatSet::q>>;
  using CELL = Cell<T, LatSet, FIELDS>;
  BlockFieldManager<FLAG, T, D> FlagFM(geo, VoidFlag);
  FlagFM.forEach(cavity, [&](FLAG& field, std::size_t id) { field.SetField(id, AABBFlag); });
  FlagFM.template SetupBoundary<LatSet>(cavity, BouncebackFlag);
  FlagFM.forEach(toplid, [&](FLAG& field, std::size_t id) {
    if (util::isFlag(field.get(id), BouncebackFlag)) field.SetField(id, BBMovingWallFlag);
  });
  ValuePack InitValues(BaseConv.getLatRhoInit(), Vector<T, D>{}, T{});
  using LM = BlockLatticeManager<T, LatSet, FIELDS>;
  using FM = BlockFieldManager<FLAG, T, D>;
  LM NSLattice(geo, InitValues, BaseConv);
  Vector<T, D> U_Wall{};
  U_Wall[0] = std::strtof("0.1", nullptr);
  const Vector<T, D> LatU_Wall = BaseConv.getLatticeU(U_Wall);
  NSLattice.template getField<VELOCITY<T, LatSet::d>>().forEach(
    toplid, FlagFM, BBMovingWallFlag, [&](auto& field, std::size_t id) { field.SetField(id, LatU_Wall); });
  BBLikeFixedBlockBdManager<bounceback::normal<CELL>, LM, FM> NS_BB("NS_BB", NSLattice, FlagFM, BouncebackFlag,
                                                                     VoidFlag);
  BBLikeFixedBlockBdManager<bounceback::movingwall<CELL>, LM, FM> NS_BBMW("NS_BBMW", NSLattice, FlagFM,
                                                                          BBMovingWallFlag, VoidFlag);
  BlockBoundaryManager BM(&NS_BB, &NS_BBMW);
  using BGKrhoU = collision::BGK<moment::rhoU<CELL>, equilibrium::SecondOrder<CELL>>;
  using AllTask = tmp::Key_TypePair<std::uint8_t(AABBFlag | BouncebackFlag | BBMovingWallFlag), BGKrhoU>;
  using NSTask = TaskSelector<std::uint8_t, CELL, AllTask>;
  for (int s = 0; s < steps; ++s) {
    NSLattice.template ApplyCellDynamics<NSTask>(FlagFM);
    NSLattice.Stream();
    BM.Apply();
  }
  using RhoUTask = tmp::Key_TypePair<std::uint8_t(AABBFlag | BouncebackFlag | BBMovingWallFlag),
                                     moment::rhoU<CELL, true>>;
  using RhoUSel = TaskSelector<std::uint8_t, CELL, RhoUTask>;
  NSLattice.template ApplyCellDynamics<RhoUSel>(FlagFM);

  vtmo::ScalarWriter RhoWriter("Rho", NSLattice.template getField<RHO<T>>());
  vtmo::VectorWriter VecWriter("Velocity", NSLattice.template getField<VELOCITY<T, LatSet::d>>());
  vtmo::ScalarWriter FlagWriter("flag", FlagFM);
  vtmo::vtmWriter<T, LatSet::d> Writer(name, geo, olth);
  Writer.addWriterSet(RhoWriter, VecWriter, FlagWriter);
  Writer.WriteBinary(steps);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s <d3q19|d2q9> Ni Nj Nk steps overlap_threshold name\n", argv[0]);
    return 2;
  }
  const std::string lat = argv[1];
  const int ni = std::atoi(argv[2]), nj = std::atoi(argv[3]), nk = std::atoi(argv[4]);
  const int steps = std::atoi(argv[5]), olth = std::atoi(argv[6]);
  const std::string name = argv[7];
  if (lat == "d3q19") {
    AABB<T, 3> cavity(Vector<T, 3>{}, Vector<T, 3>(T(ni), T(nj), T(nk)));
    AABB<T, 3> toplid(Vector<T, 3>(T(1), T(1), T(nk - 1)), Vector<T, 3>(T(ni - 1), T(nj - 1), T(nk)));
    BlockGeometry3D<T> geo(ni, nj, nk, 1, cavity, T(1));
    return run<D3Q19<T>, 3>(geo, cavity, toplid, ni, steps, olth, name);
  }
  if (lat == "d2q9") {
    AABB<T, 2> cavity(Vector<T, 2>(T(0), T(0)), Vector<T, 2>(T(ni), T(nj)));
    AABB<T, 2> toplid(Vector<T, 2>(T(1), T(nj - 1)), Vector<T, 2>(T(ni - 1), T(nj)));
    BlockGeometry2D<T> geo(ni, nj, 1, cavity, T(1));
    return run<D2Q9<T>, 2>(geo, cavity, toplid, ni, steps, olth, name);
  }
  std::fprintf(stderr, "unknown lattice %s\n", lat.c_str());
  return 2;
}
