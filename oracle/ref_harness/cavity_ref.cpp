// Fixture writer for the lid-driven cavity, built against FreeLB's own headers.
//
// This program is test infrastructure: it drives the reference library
// (header-only, compiled from /root/reference/src by oracle/Makefile) through
// the same calls as benchmarks/cavity3d/cavity3d.cpp:111-222 and
// examples/couette2d/couette2d.cpp:178-216, then dumps the populations, the
// flag field and the reference's own rho/u read-out so the C restatement and
// the HIP kernels can be compared against it.
//
// Usage:
//   cavity_ref <d3q19|d2q9> <fw|hw|hwfield|mrt|smag|force> Ni Nj Nk steps RT KineVisc Uwall out.bin [Cs]
// (Nk is ignored for d2q9.) In the hw mode, out = "time:W" runs W untimed
// steps, then times `steps` steps and prints their MLUPS instead of dumping.
//
// Modes
//   fw      : wall cells collide as BounceBack / BounceBackMovingWall
//             (benchmarks/cavity3d/cavity3d.cpp:176-182, the published path).
//   hw      : every non-void cell collides with BGK<rhoU>; after Stream the
//             halfway managers bounceback::normal (BB flag) and
//             bounceback::movingwall (BBMW flag) rebuild the unknown pops
//             (boundary.h:75-129).
//   hwfield : like hw but wall cells collide with BGK<useFieldrhoU>
//             (examples/couette2d/couette2d.cpp:183-184).
//   mrt     : fw walls, fluid cells collide with MRT_Feq_RhoU (D2Q9 only:
//             lbm/collisionMRT.ur.h:45-129).
//   smag    : fw walls, fluid cells collide with SmagorinskyBGK<rhoU>
//             (lbm/collisionLES.h:29-67), SMAGORINSKY data field = Cs.
//   force   : like hw, every non-void cell collides with
//             BGKForce<forcerhoU, SecondOrder, ConstForce> (lbm/collision.h:61-90),
//             CONSTFORCE data field = (1e-4, -5e-5, 2e-5) (d components).
//
// Output layout (little endian):
//   char[8]  "FLBREF01"
//   int32    d, q, nx, ny, nz, steps
//   float32  omega, uwall_lattice
//   uint8    flags[N]
//   float32  pops[q][N]      (logical post-stream values, cell[i])
//   float32  rhou[d+1][N]    (rho, u_x, u_y[, u_z]) from moment::rhoU<CELL,true>
//                            on every non-void cell; 0 on void cells

#include "freelb.h"
#include "freelb.hh"

#include <chrono>
#include <cstdio>
#include <type_traits>
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
  std::string lattice, mode;
  int Ni, Nj, Nk, steps;
  T RT, visc, uwall, cs = T(0);
  std::string out;
  int time_warmup = -1;  // >= 0: out was "time:<warmup>", time the steps instead of dumping
};

template <typename BLM, typename FM, typename CELL, typename LatSet>
void dump(const Params& p, BLM& lat, FM& flagfm, T omega, T ulat, int nz) {
  using RhoUTask = tmp::Key_TypePair<std::uint8_t(AABBFlag | BouncebackFlag | BBMovingWallFlag),
                                     moment::rhoU<CELL, true>>;
  using RhoUSel = TaskSelector<std::uint8_t, CELL, RhoUTask>;
  lat.template ApplyCellDynamics<RhoUSel>(flagfm);

  auto& blat = lat.getBlockLat(0);
  const auto& flags = flagfm.getBlockField(0).getField(0);
  const int nx = blat.getNx();
  const int ny = blat.getNy();
  const std::size_t N = blat.getN();
  const int d = LatSet::d;
  const int q = LatSet::q;

  FILE* f = std::fopen(p.out.c_str(), "wb");
  if (!f) {
    std::perror("fopen");
    std::exit(2);
  }
  std::fwrite("FLBREF01", 1, 8, f);
  int hdr[6] = {d, q, nx, ny, nz, p.steps};
  std::fwrite(hdr, sizeof(int), 6, f);
  std::fwrite(&omega, sizeof(T), 1, f);
  std::fwrite(&ulat, sizeof(T), 1, f);
  std::vector<std::uint8_t> fl(N);
  for (std::size_t id = 0; id < N; ++id) fl[id] = flags[id];
  std::fwrite(fl.data(), 1, N, f);
  std::vector<T> buf(N);
  for (int i = 0; i < q; ++i) {
    for (std::size_t id = 0; id < N; ++id) {
      CELL cell(id, blat);
      buf[id] = cell[i];
    }
    std::fwrite(buf.data(), sizeof(T), N, f);
  }
  for (int c = 0; c <= d; ++c) {
    for (std::size_t id = 0; id < N; ++id) {
      CELL cell(id, blat);
      if (util::isFlag(fl[id], VoidFlag)) {
        buf[id] = T{};
      } else if (c == 0) {
        buf[id] = cell.template get<RHO<T>>();
      } else {
        buf[id] = cell.template get<VELOCITY<T, LatSet::d>>()[c - 1];
      }
    }
    std::fwrite(buf.data(), sizeof(T), N, f);
  }
  std::fclose(f);
}

template <typename LatSet, int D, bool SMAG, bool FORCE = false>
int run(const Params& p) {
  BaseConverter<T> BaseConv(LatSet::cs2);
  BaseConv.ConvertFromRT(T(1), p.RT, T(0.001), T(p.Ni), p.uwall, p.visc);

  using FIELDS = std::conditional_t<SMAG,
      TypePack<RHO<T>, VELOCITY<T, LatSet::d>, POP<T, LatSet::q>, SMAGORINSKY<T>>,
      std::conditional_t<FORCE,
        TypePack<RHO<T>, VELOCITY<T, LatSet::d>, POP<T, LatSet::q>, CONSTFORCE<T, LatSet::d>>,
        TypePack<RHO<T>, VELOCITY<T, LatSet::d>, POP<T, LatSet::q>>>>;
  using CELL = Cell<T, LatSet, FIELDS>;

  Vector<T, D> U_Wall{};
  U_Wall[0] = p.uwall;
  Vector<T, D> LatU_Wall = BaseConv.getLatticeU(U_Wall);

  auto body = [&](auto& Geo, auto& cavity, auto& toplid, int nz) {
    BlockFieldManager<FLAG, T, D> FlagFM(Geo, VoidFlag);
    FlagFM.forEach(cavity, [&](FLAG& field, std::size_t id) { field.SetField(id, AABBFlag); });
    FlagFM.template SetupBoundary<LatSet>(cavity, BouncebackFlag);
    FlagFM.forEach(toplid, [&](FLAG& field, std::size_t id) {
      if (util::isFlag(field.get(id), BouncebackFlag)) field.SetField(id, BBMovingWallFlag);
    });

    auto make_init = [&]() {
      if constexpr (SMAG) {
        return ValuePack(BaseConv.getLatRhoInit(), Vector<T, D>{}, T{}, p.cs);
      } else if constexpr (FORCE) {
        const T fv[3] = {std::strtof("1e-4", nullptr), std::strtof("-5e-5", nullptr),
                         std::strtof("2e-5", nullptr)};
        Vector<T, D> F{};
        for (int a = 0; a < D; ++a) F[a] = fv[a];
        return ValuePack(BaseConv.getLatRhoInit(), Vector<T, D>{}, T{}, F);
      } else {
        return ValuePack(BaseConv.getLatRhoInit(), Vector<T, D>{}, T{});
      }
    };
    auto InitValues = make_init();
    BlockLatticeManager<T, LatSet, FIELDS> NSLattice(Geo, InitValues, BaseConv);
    NSLattice.template getField<VELOCITY<T, LatSet::d>>().forEach(
      toplid, FlagFM, BBMovingWallFlag,
      [&](auto& field, std::size_t id) { field.SetField(id, LatU_Wall); });

    using LM = BlockLatticeManager<T, LatSet, FIELDS>;
    using FM = BlockFieldManager<FLAG, T, D>;
    BBLikeFixedBlockBdManager<bounceback::normal<CELL>, LM, FM> NS_BB("NS_BB", NSLattice, FlagFM,
                                                                       BouncebackFlag, VoidFlag);
    BBLikeFixedBlockBdManager<bounceback::movingwall<CELL>, LM, FM> NS_BBMW(
      "NS_BBMW", NSLattice, FlagFM, BBMovingWallFlag, VoidFlag);
    BlockBoundaryManager BM(&NS_BB, &NS_BBMW);

    using BGKrhoU = collision::BGK<moment::rhoU<CELL>, equilibrium::SecondOrder<CELL>>;
    using BGKfield = collision::BGK<moment::useFieldrhoU<CELL>, equilibrium::SecondOrder<CELL>>;

    if (p.mode == "mrt" || p.mode == "smag") {
      using BBTask = tmp::Key_TypePair<BouncebackFlag, collision::BounceBack<CELL>>;
      using BBMVTask = tmp::Key_TypePair<BBMovingWallFlag, collision::BounceBackMovingWall<CELL>>;
      if constexpr (SMAG) {
        using Bulk = tmp::Key_TypePair<AABBFlag,
            collision::SmagorinskyBGK<moment::rhoU<CELL>, equilibrium::SecondOrder<CELL>>>;
        using NSTask = TaskSelector<std::uint8_t, CELL, Bulk, BBTask, BBMVTask>;
        for (int s = 0; s < p.steps; ++s) {
          NSLattice.template ApplyCellDynamics<NSTask>(FlagFM);
          NSLattice.Stream();
        }
      } else if constexpr (D == 2) {
        using Bulk = tmp::Key_TypePair<AABBFlag, collision::MRT_Feq_RhoU<CELL>>;
        using NSTask = TaskSelector<std::uint8_t, CELL, Bulk, BBTask, BBMVTask>;
        for (int s = 0; s < p.steps; ++s) {
          NSLattice.template ApplyCellDynamics<NSTask>(FlagFM);
          NSLattice.Stream();
        }
      } else {
        std::fprintf(stderr, "FreeLB has no D3Q19 MRT\n");
        std::exit(2);
      }
    } else if (p.mode == "fw") {
      using BulkTask = tmp::Key_TypePair<AABBFlag, BGKrhoU>;
      using BBTask = tmp::Key_TypePair<BouncebackFlag, collision::BounceBack<CELL>>;
      using BBMVTask = tmp::Key_TypePair<BBMovingWallFlag, collision::BounceBackMovingWall<CELL>>;
      using NSTask = TaskSelector<std::uint8_t, CELL, BulkTask, BBTask, BBMVTask>;
      for (int s = 0; s < p.steps; ++s) {
        NSLattice.template ApplyCellDynamics<NSTask>(FlagFM);
        NSLattice.Stream();
      }
    } else if (p.mode == "hw") {
      using AllTask =
        tmp::Key_TypePair<std::uint8_t(AABBFlag | BouncebackFlag | BBMovingWallFlag), BGKrhoU>;
      using NSTask = TaskSelector<std::uint8_t, CELL, AllTask>;
      auto step = [&]() {
        NSLattice.template ApplyCellDynamics<NSTask>(FlagFM);
        NSLattice.Stream();
        BM.Apply();
      };
      if (p.time_warmup >= 0) {
        // timing mode: warm-up steps, then the timed steps, no dump
        for (int s = 0; s < p.time_warmup; ++s) step();
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < p.steps; ++s) step();
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const double cells = double(p.Ni) * p.Nj * (D == 3 ? p.Nk : 1);
        std::printf("MLUPS %.3f cells %.0f steps %d seconds %.6f threads %d\n", cells * p.steps / sec / 1e6, cells,
                    p.steps, sec, Thread_Num);
        std::exit(0);
      }
      for (int s = 0; s < p.steps; ++s) step();
    } else if (p.mode == "force") {
      if constexpr (FORCE) {
        using ForceS = force::ConstForce<CELL>;
        using BGKF = collision::BGKForce<moment::forcerhoU<CELL, ForceS>, equilibrium::SecondOrder<CELL>,
                                         ForceS>;
        using AllTask =
          tmp::Key_TypePair<std::uint8_t(AABBFlag | BouncebackFlag | BBMovingWallFlag), BGKF>;
        using NSTask = TaskSelector<std::uint8_t, CELL, AllTask>;
        for (int s = 0; s < p.steps; ++s) {
          NSLattice.template ApplyCellDynamics<NSTask>(FlagFM);
          NSLattice.Stream();
          BM.Apply();
        }
      }
    } else if (p.mode == "hwfield") {
      using BulkTask = tmp::Key_TypePair<AABBFlag, BGKrhoU>;
      using WallTask =
        tmp::Key_TypePair<std::uint8_t(BouncebackFlag | BBMovingWallFlag), BGKfield>;
      using NSTask = TaskSelector<std::uint8_t, CELL, BulkTask, WallTask>;
      for (int s = 0; s < p.steps; ++s) {
        NSLattice.template ApplyCellDynamics<NSTask>(FlagFM);
        NSLattice.Stream();
        BM.Apply();
      }
    } else {
      std::fprintf(stderr, "unknown mode %s\n", p.mode.c_str());
      std::exit(2);
    }
    dump<LM, FM, CELL, LatSet>(p, NSLattice, FlagFM, BaseConv.getOMEGA(), LatU_Wall[0], nz);
  };

  if constexpr (D == 3) {
    AABB<T, 3> cavity(Vector<T, 3>{}, Vector<T, 3>(T(p.Ni), T(p.Nj), T(p.Nk)));
    AABB<T, 3> toplid(Vector<T, 3>(T(1), T(1), T(p.Nk - 1)),
                      Vector<T, 3>(T(p.Ni - 1), T(p.Nj - 1), T(p.Nk)));
    BlockGeometry3D<T> Geo(p.Ni, p.Nj, p.Nk, 1, cavity, T(1));
    body(Geo, cavity, toplid, p.Nk + 2);
  } else {
    AABB<T, 2> cavity(Vector<T, 2>(T(0), T(0)), Vector<T, 2>(T(p.Ni), T(p.Nj)));
    AABB<T, 2> toplid(Vector<T, 2>(T(1), T(p.Nj - 1)), Vector<T, 2>(T(p.Ni - 1), T(p.Nj)));
    BlockGeometry2D<T> Geo(p.Ni, p.Nj, 1, cavity, T(1));
    body(Geo, cavity, toplid, 1);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 11 && argc != 12) {
    std::fprintf(stderr,
                 "usage: %s <d3q19|d2q9> <fw|hw|hwfield|mrt|smag> Ni Nj Nk steps RT KineVisc Uwall out [Cs]\n",
                 argv[0]);
    return 2;
  }
  Params p;
  p.lattice = argv[1];
  p.mode = argv[2];
  p.Ni = std::atoi(argv[3]);
  p.Nj = std::atoi(argv[4]);
  p.Nk = std::atoi(argv[5]);
  p.steps = std::atoi(argv[6]);
  // strtof rounds the decimal string straight to float, as the reference's
  // iniReader (stringstream >> float, io/ini_reader.h) does.
  p.RT = std::strtof(argv[7], nullptr);
  p.visc = std::strtof(argv[8], nullptr);
  p.uwall = std::strtof(argv[9], nullptr);
  p.out = argv[10];
  if (p.out.rfind("time:", 0) == 0) p.time_warmup = std::atoi(p.out.c_str() + 5);
  if (argc == 12) p.cs = std::strtof(argv[11], nullptr);
#ifdef _OPENMP
  Thread_Num = omp_get_max_threads();
#endif
  const bool smag = p.mode == "smag";
  const bool frc = p.mode == "force";
  if (p.lattice == "d3q19")
    return smag ? run<D3Q19<T>, 3, true>(p) : (frc ? run<D3Q19<T>, 3, false, true>(p) : run<D3Q19<T>, 3, false>(p));
  if (p.lattice == "d2q9")
    return smag ? run<D2Q9<T>, 2, true>(p) : (frc ? run<D2Q9<T>, 2, false, true>(p) : run<D2Q9<T>, 2, false>(p));
  std::fprintf(stderr, "unknown lattice %s\n", p.lattice.c_str());
  return 2;
}
