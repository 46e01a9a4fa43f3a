// Fixture writer for the two-level refined lid-driven cavity, built against
// FreeLB's own headers (test infrastructure, compiled from /root/reference/src
// by oracle/Makefile).
//
// It follows examples/cavref2d/cavref2d.cpp:135-228: a BlockGeometryHelper
// divides the cavity into block cells of `blockcell` cells per axis, the
// block cells whose index lies in [p0, p1) on every axis are refined once,
// CreateBlocks merges same-level cells into blocks (level-0 blocks get a
// one-cell overlap, level-1 blocks two fine cells), and the loop
//   ApplyCellDynamics(count) ; Stream(count) ; BM.Apply(count) ; FullCommunicate(count)
// runs `steps` counts (the coarse level steps on even counts only;
// block_lattice.hh:736-764, 1462-1568). Fluid cells collide with
// BGK<rhoU<CELL, true>> (moments written to the RHO / VELOCITY fields, which
// averCommunicate / intpCommunicate read), wall cells with
// BGK<useFieldrhoU>, and the halfway managers bounceback::normal /
// movingwall rebuild the wall links.
//
// Usage:
//   refine_ref <d3q19|d2q9> Ni Nj Nk blockcell p0 p1 steps RT KineVisc Uwall out.bin [q0 q1]
// (Nk is ignored for d2q9.) With q0 q1, the block cells whose index lies in
// [q0, q1) on every axis are refined twice (a third level, BasicBlock::refine(2),
// geometry/basic_geometry.hh:293-307), the others of [p0, p1) once.
//
// Output layout (little endian):
//   char[8]  "FLBREF02"
//   int32    d, q, nblocks, steps
//   float32  uwall_lattice
//   per block:
//     int32    level, overlap, nx, ny, nz        (mesh including the overlap)
//     float32  omega, voxel size, min[3]         (min corner of the mesh, physical units)
//     uint8    flags[N]
//     float32  pops[q][N]                         (post-stream values, cell[i])
//     float32  rho[N], u[d][N]                    (the RHO / VELOCITY fields)

#include "freelb.h"
#include "freelb.hh"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using T = float;

int Thread_Num = 1;

namespace {

constexpr std::uint8_t VoidFlag = std::uint8_t(1);
constexpr std::uint8_t AABBFlag = std::uint8_t(2);
constexpr std::uint8_t BouncebackFlag = std::uint8_t(4);
constexpr std::uint8_t BBMovingWallFlag = std::uint8_t(8);

struct Params {
  std::string lattice;
  int Ni, Nj, Nk, cell, p0, p1, steps;
  int q0 = 0, q1 = 0;  // doubly refined block cells (none when q0 == q1)
  T RT, visc, uwall;
  std::string out;
};

template <typename LatSet, int D, typename Helper, typename Geo, typename BoxT>
int run(const Params& p, Helper& GeoHelper, const BoxT& cavity, const BoxT& toplid) {
  BaseConverter<T> BaseConv(LatSet::cs2);
  BaseConv.ConvertFromRT(T(1), p.RT, T(0.001), T(p.Ni), p.uwall, p.visc);

  GeoHelper.forEachBlockCell([&](BasicBlock<T, D>& block) {
    const auto idx = block.getIdxBlock().getMin();
    bool in = true, in2 = p.q0 < p.q1;
    for (int a = 0; a < D; ++a) {
      const int c = (idx[a] - 1) / p.cell;  // the helper's base block starts at index 1
      in = in && c >= p.p0 && c < p.p1;
      in2 = in2 && c >= p.q0 && c < p.q1;
    }
    if (in2) block.refine(std::uint8_t(2));
    else if (in) block.refine();
  });
  GeoHelper.CreateBlocks();
  GeoHelper.LoadBalancing(1);
  // useHelperOlap = false: level-1 blocks get the two-cell overlap that
  // intpCommunicate assumes (with the default single-cell overlap its
  // coarse stencils run past the sending block, an out-of-bounds read)
  Geo geo(GeoHelper, false);

  using FIELDS = TypePack<RHO<T>, VELOCITY<T, LatSet::d>, POP<T, LatSet::q>>;
  using FIELDREFS = TypePack<FLAG>;
  using FIELDSPACK = TypePack<FIELDS, FIELDREFS>;
  using ALLFIELDS = typename ExtractFieldPack<FIELDSPACK>::mergedpack;
  using CELL = Cell<T, LatSet, ALLFIELDS>;

  BlockFieldManager<FLAG, T, D> FlagFM(geo, VoidFlag);
  FlagFM.forEach(cavity, [&](auto& field, std::size_t id) { field.SetField(id, AABBFlag); });
  FlagFM.template SetupBoundary<LatSet>(cavity, BouncebackFlag);
  FlagFM.forEach(toplid, [&](auto& field, std::size_t id) {
    if (util::isFlag(field.get(id), BouncebackFlag)) field.SetField(id, BBMovingWallFlag);
  });

  ValuePack InitValues(BaseConv.getLatRhoInit(), Vector<T, D>{}, T{});
  using LM = BlockLatticeManager<T, LatSet, FIELDSPACK>;
  LM NSLattice(geo, InitValues, BaseConv);
  Vector<T, D> U_Wall{};
  U_Wall[0] = p.uwall;
  const Vector<T, D> LatU_Wall = BaseConv.getLatticeU(U_Wall);
  NSLattice.template getField<VELOCITY<T, LatSet::d>>().forEach(
    FlagFM, BBMovingWallFlag, [&](auto& field, std::size_t id) { field.SetField(id, LatU_Wall); });

  using FM = BlockFieldManager<FLAG, T, D>;
  BBLikeFixedBlockBdManager<bounceback::normal<CELL>, LM, FM> NS_BB("NS_BB", NSLattice, FlagFM,
                                                                     BouncebackFlag, VoidFlag);
  BBLikeFixedBlockBdManager<bounceback::movingwall<CELL>, LM, FM> NS_BBMW("NS_BBMW", NSLattice, FlagFM,
                                                                          BBMovingWallFlag, VoidFlag);
  BlockBoundaryManager BM(&NS_BB, &NS_BBMW);

  using BulkTask = tmp::Key_TypePair<AABBFlag,
    collision::BGK<moment::rhoU<CELL, true>, equilibrium::SecondOrder<CELL>>>;
  using WallTask = tmp::Key_TypePair<std::uint8_t(BouncebackFlag | BBMovingWallFlag),
    collision::BGK<moment::useFieldrhoU<CELL>, equilibrium::SecondOrder<CELL>>>;
  using TaskCollection = tmp::TupleWrapper<BulkTask, WallTask>;
  using TaskSel = tmp::TaskSelector<TaskCollection, std::uint8_t, CELL>;

  for (std::int64_t count = 0; count < p.steps; ++count) {
    NSLattice.template ApplyCellDynamics<TaskSel>(count, FlagFM);
    NSLattice.Stream(count);
    BM.Apply(count);
    NSLattice.FullCommunicate(count);
  }

  FILE* f = std::fopen(p.out.c_str(), "wb");
  if (!f) {
    std::perror("fopen");
    return 2;
  }
  std::fwrite("FLBREF02", 1, 8, f);
  const int nblocks = static_cast<int>(NSLattice.getBlockLats().size());
  int hdr[4] = {D, int(LatSet::q), nblocks, p.steps};
  std::fwrite(hdr, sizeof(int), 4, f);
  const T ul = LatU_Wall[0];
  std::fwrite(&ul, sizeof(T), 1, f);
  for (int b = 0; b < nblocks; ++b) {
    auto& blat = NSLattice.getBlockLat(b);
    const auto& block = geo.getBlock(b);
    const auto& flags = FlagFM.getBlockField(b).getField(0);
    const std::size_t N = blat.getN();
    int bh[5] = {int(block.getLevel()), block.getOverlap(), blat.getNx(), blat.getNy(),
                 D == 3 ? blat.getNz() : 1};
    std::fwrite(bh, sizeof(int), 5, f);
    float geomf[5] = {blat.getOmega(), block.getVoxelSize(), block.getMin()[0], block.getMin()[1],
                      D == 3 ? float(block.getMin()[D - 1]) : 0.0f};
    std::fwrite(geomf, sizeof(float), 5, f);
    std::vector<std::uint8_t> fl(N);
    for (std::size_t id = 0; id < N; ++id) fl[id] = flags[id];
    std::fwrite(fl.data(), 1, N, f);
    std::vector<T> buf(N);
    for (unsigned int i = 0; i < LatSet::q; ++i) {
      for (std::size_t id = 0; id < N; ++id) {
        CELL cell(id, blat);
        buf[id] = cell[i];
      }
      std::fwrite(buf.data(), sizeof(T), N, f);
    }
    for (int c = 0; c <= D; ++c) {
      for (std::size_t id = 0; id < N; ++id) {
        CELL cell(id, blat);
        buf[id] = c == 0 ? cell.template get<RHO<T>>() : cell.template get<VELOCITY<T, LatSet::d>>()[c - 1];
      }
      std::fwrite(buf.data(), sizeof(T), N, f);
    }
  }
  std::fclose(f);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 13 && argc != 15) {
    std::fprintf(stderr, "usage: %s <d3q19|d2q9> Ni Nj Nk blockcell p0 p1 steps RT KineVisc Uwall out [q0 q1]\n",
                 argv[0]);
    return 2;
  }
  Params p;
  p.lattice = argv[1];
  p.Ni = std::atoi(argv[2]);
  p.Nj = std::atoi(argv[3]);
  p.Nk = std::atoi(argv[4]);
  p.cell = std::atoi(argv[5]);
  p.p0 = std::atoi(argv[6]);
  p.p1 = std::atoi(argv[7]);
  p.steps = std::atoi(argv[8]);
  p.RT = std::strtof(argv[9], nullptr);
  p.visc = std::strtof(argv[10], nullptr);
  p.uwall = std::strtof(argv[11], nullptr);
  p.out = argv[12];
  if (argc == 15) {
    p.q0 = std::atoi(argv[13]);
    p.q1 = std::atoi(argv[14]);
  }
  if (p.lattice == "d3q19") {
    AABB<T, 3> cavity(Vector<T, 3>{}, Vector<T, 3>(T(p.Ni), T(p.Nj), T(p.Nk)));
    AABB<T, 3> toplid(Vector<T, 3>(T(1), T(1), T(p.Nk - 1)), Vector<T, 3>(T(p.Ni - 1), T(p.Nj - 1), T(p.Nk)));
    BlockGeometryHelper3D<T> helper(p.Ni, p.Nj, p.Nk, cavity, T(1), p.cell);
    return run<D3Q19<T>, 3, BlockGeometryHelper3D<T>, BlockGeometry3D<T>>(p, helper, cavity, toplid);
  }
  if (p.lattice == "d2q9") {
    AABB<T, 2> cavity(Vector<T, 2>(T(0), T(0)), Vector<T, 2>(T(p.Ni), T(p.Nj)));
    AABB<T, 2> toplid(Vector<T, 2>(T(1), T(p.Nj - 1)), Vector<T, 2>(T(p.Ni - 1), T(p.Nj)));
    BlockGeometryHelper2D<T> helper(p.Ni, p.Nj, cavity, T(1), p.cell);
    return run<D2Q9<T>, 2, BlockGeometryHelper2D<T>, BlockGeometry2D<T>>(p, helper, cavity, toplid);
  }
  std::fprintf(stderr, "unknown lattice %s\n", p.lattice.c_str());
  return 2;
}
