// flb/block_lattice.hpp -- C++ host adapter over the C ABI (include/flb_abi.h).
//
// It keeps the call shape of FreeLB's block-lattice interface so that a
// FreeLB driver's main loop ports by changing types, not structure:
//
//   FreeLB (benchmarks/cavity3d/cavity3d.cpp:155-222)     this adapter
//   -----------------------------------------------       ------------------------------------
//   using BulkTask = tmp::Key_TypePair<AABB,               using BulkTask = tmp::Key_TypePair<AABB,
//     collision::BGK<moment::rhoU<CELL>,                     collision::BGK<moment::rhoU<CELL>,
//                    equilibrium::SecondOrder<CELL>>>;                      equilibrium::SecondOrder<CELL>>>;
//   using NSTask = TaskSelector<uint8_t, CELL, ...>;       using NSTask = TaskSelector<uint8_t, CELL, ...>;
//   NSLattice.ApplyCellDynamics<NSTask>(FlagFM);           NSLattice.ApplyCellDynamics<NSTask>(flags);
//   NSLattice.Stream();                                    NSLattice.Stream();
//   BM.Apply();                                            BM.Apply();
//
// Types mirror FreeLB's names (lbm/collision.h, lbm/moment.h,
// lbm/equilibrium.h, boundary/bounce_back_boundary.h, utils/tmp.h,
// lbm/lattice_set.h); each dynamics type maps to one FLB_DYN_* kind, and a
// TaskSelector to the first-match task table the GPU kernel evaluates per
// cell. Everything runs on the GPU: the adapter only builds tables and calls
// the C ABI. Errors throw flb::Error carrying the ABI's code and message.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../flb_abi.h"

namespace flb {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != FLB_OK) throw Error(rc, std::string("flb: ") + flb_last_error());
}

// ---- lattice sets (lbm/lattice_set.h:168-194, 259-296) -----------------------
template <typename T>
struct D3Q19 {
  using FloatType = T;
  static constexpr unsigned int d = 3;
  static constexpr unsigned int q = 19;
};
template <typename T>
struct D2Q9 {
  using FloatType = T;
  static constexpr unsigned int d = 2;
  static constexpr unsigned int q = 9;
};

// advection-diffusion sets (lbm/lattice_set.h:144-155, 196-207)
template <typename T>
struct D2Q5 {
  using FloatType = T;
  static constexpr unsigned int d = 2;
  static constexpr unsigned int q = 5;
};
template <typename T>
struct D3Q7 {
  using FloatType = T;
  static constexpr unsigned int d = 3;
  static constexpr unsigned int q = 7;
};
template <typename LatSet>
constexpr bool is_scalar_lattice = LatSet::q == 5 || LatSet::q == 7;

// The CELL of FreeLB's dynamics templates; only its lattice set matters here.
template <typename T, typename LatSet>
struct Cell {
  using FloatType = T;
  using LatticeSet = LatSet;
};

// ---- moments / equilibria / collisions (tags) ----------------------------------
namespace moment {
template <typename CELL, bool WriteToField = false>
struct rhoU {};  // lbm/moment.h:381-398
template <typename CELL>
struct useFieldrhoU {};  // lbm/moment.h:71-90
template <typename CELL, typename ForceScheme, bool WriteToField = false>
struct forcerhoU {};  // lbm/moment.h:441-480
template <typename CELL, bool WriteToField = false>
struct rho {};  // lbm/moment.h:159-182
template <typename CELL>
struct useFieldU {};  // lbm/moment.h:56-69
template <typename... Momenta>
struct MomentaTuple {};  // lbm/moment.h:666-679
}  // namespace moment

namespace force {
template <typename CELL>
struct ConstForce {};  // lbm/force.h:86-107 (the CONSTFORCE data field)
template <typename CELL, unsigned int dir = 2>
struct ScalarForce {};  // lbm/force.h:109-133 (the SCALARFORCE field, on the last axis)
// lbm/force.h:160-181: SCALARFORCE of the flow cell += GBETA (rho1 - RHOINIT)
template <typename CELL0, typename CELL1, bool Scalar = true>
struct Buoyancy {
  static_assert(Scalar, "the vector-force form of Buoyancy is not provided");
};
}  // namespace force

namespace equilibrium {
template <typename CELL>
struct SecondOrder {};  // lbm/equilibrium.h:79-117
}  // namespace equilibrium

namespace collision {
template <typename MomentaScheme, typename EquilibriumScheme>
struct BGK;  // lbm/collision.h:32-59
template <typename CELL, typename EQ>
struct BGK<moment::rhoU<CELL>, EQ> {
  static constexpr uint8_t kind = FLB_DYN_BGK;
};
template <typename CELL, typename EQ>
struct BGK<moment::rhoU<CELL, true>, EQ> {  // the moments also go to the RHO / VELOCITY fields
  static constexpr uint8_t kind = FLB_DYN_BGK_WRITE;
};
template <typename CELL, typename EQ>
struct BGK<moment::useFieldrhoU<CELL>, EQ> {
  // on a scalar lattice: relax toward the field's scalar (a fixed-value wall)
  static constexpr uint8_t kind =
      is_scalar_lattice<typename CELL::LatticeSet> ? FLB_DYN_SCALAR_BGK_FIELD : FLB_DYN_BGK_FIELD;
};
// advection-diffusion: the scalar to the field, u from the (shared) VELOCITY field
template <typename CELL, typename EQ>
struct BGK<moment::MomentaTuple<moment::rho<CELL, true>, moment::useFieldU<CELL>>, EQ> {
  static_assert(is_scalar_lattice<typename CELL::LatticeSet>, "a D2Q5 / D3Q7 dynamics");
  static constexpr uint8_t kind = FLB_DYN_SCALAR_BGK;
};
template <typename CELL>
struct BounceBack {  // lbm/collision.ur.h:62-93
  static constexpr uint8_t kind = FLB_DYN_BOUNCEBACK;
};
template <typename CELL>
struct BounceBackMovingWall {  // lbm/collision.ur.h:130-187
  static constexpr uint8_t kind = FLB_DYN_BOUNCEBACK_MW;
};
template <typename CELL, bool WriteToField = false>
struct MRT_Feq_RhoU {  // lbm/collisionMRT.h:132-190
  static constexpr uint8_t kind = FLB_DYN_MRT;
};
template <typename MomentaScheme, typename EquilibriumScheme, bool WriteToField = false>
struct SmagorinskyBGK {  // lbm/collisionLES.h:29-67
  static constexpr uint8_t kind = FLB_DYN_SMAG_BGK;
};
template <typename MomentaScheme, typename EquilibriumScheme, typename ForceScheme>
struct BGKForce;  // lbm/collision.h:61-90
template <typename CELL, typename EQ>
struct BGKForce<moment::forcerhoU<CELL, force::ScalarForce<CELL>, true>, EQ, force::ScalarForce<CELL>> {
  static constexpr uint8_t kind = FLB_DYN_BGK_SCALAR_FORCE;
};
template <typename CELL, typename EQ>
struct BGKForce<moment::forcerhoU<CELL, force::ConstForce<CELL>>, EQ, force::ConstForce<CELL>> {
  static constexpr uint8_t kind = FLB_DYN_BGK_FORCE;
};
// MRT relaxation whose shear rates take the Smagorinsky omega (no FreeLB form).
template <typename CELL>
struct SmagorinskyMRT {
  static constexpr uint8_t kind = FLB_DYN_SMAG_MRT;
};
}  // namespace collision

// ---- task selection (utils/tmp.h:194-202, 283-309) -------------------------------
namespace tmp {
template <auto KeyValue, typename Type>
struct Key_TypePair {
  static constexpr auto key = KeyValue;
  using type = Type;
};
}  // namespace tmp

template <typename FlagType, typename CELL, typename... Tasks>
struct TaskSelector {
  static std::vector<flb_task> table(float smagorinsky) {
    std::vector<flb_task> t;
    (t.push_back(flb_task{static_cast<uint8_t>(Tasks::key), Tasks::type::kind, {0, 0}, smagorinsky}),
     ...);
    return t;
  }
};

// ---- halfway bounce-back kinds (boundary/bounce_back_boundary.h:169-256) -----------
namespace bounceback {
template <typename CELL>
struct normal {
  static constexpr uint8_t kind = FLB_HWBC_NORMAL;
};
template <typename CELL>
struct movingwall {
  static constexpr uint8_t kind = FLB_HWBC_MOVING;
};
template <typename CELL>
struct anti_simple {  // bounce_back_boundary.h:183-193 (scalar lattices)
  static constexpr uint8_t kind = FLB_HWBC_ANTI;
};
}  // namespace bounceback

// CoupledTaskSelector<FlagType, CELL0, CELL1, Key_TypePair<mask, force::Buoyancy<...>>...>
// (utils/tmp.h:333-337): the masks of the buoyancy tasks.
template <typename FlagType, typename CELL0, typename CELL1, typename... Tasks>
struct CoupledTaskSelector {
  static std::vector<uint8_t> masks() { return {static_cast<uint8_t>(Tasks::key)...}; }
};

enum class Storage {
  F32 = FLB_STORE_F32,
  F16S = FLB_STORE_F16S,
  BF16S = FLB_STORE_BF16S,
  F64 = FLB_STORE_F64,
  F16C = FLB_STORE_F16C
};

// ---- one block on the GPU ----------------------------------------------------------
// T is the host type of populations and fields (float for F32 / F16S / BF16S
// storage, double for F64).
template <typename T, typename LatSet>
class BlockLattice {
 public:
  using FloatType = T;
  using LatticeSet = LatSet;

  // nx, ny, nz include the overlap layer (FreeLB's block extent); omega = 1/tau.
  BlockLattice(int nx, int ny, int nz, double omega, Storage storage = Storage::F32,
               int overlap = 1, void* stream = nullptr, const int periodic[3] = nullptr) {
    flb_block_desc d{};
    d.q = int(LatSet::q);
    d.storage = int(storage);
    d.nx = nx;
    d.ny = ny;
    d.nz = LatSet::d == 3 ? nz : 1;
    d.overlap = overlap;
    for (int a = 0; a < 3; ++a) d.periodic[a] = periodic ? periodic[a] : 0;
    d.omega = omega;
    d.stream = stream;
    check(flb_block_create(&d, &h_));
    n_ = size_t(nx) * size_t(ny) * size_t(d.nz);
  }
  ~BlockLattice() { flb_block_destroy(h_); }
  BlockLattice(const BlockLattice&) = delete;
  BlockLattice& operator=(const BlockLattice&) = delete;

  flb_block handle() const { return h_; }
  size_t size() const { return n_; }

  // The flag field (FreeLB's BlockFieldManager<FLAG>). Like FreeLB's
  // FlagFM.copyToDevice(), an upload is explicit: ApplyCellDynamics uploads a
  // flag field the first time it sees that object and reuses the device copy
  // afterwards; call setFlags again after changing its contents.
  void setFlags(const std::vector<uint8_t>& flags) {
    if (flags.size() != n_) throw Error(FLB_ERR_ARG, "flb: flag field size mismatch");
    check(flb_upload_flags(h_, flags.data()));
    flags_src_ = &flags;
  }
  // RHO / VELOCITY fields (VELOCITY as d*N structure of arrays).
  void setRhoField(const std::vector<double>& rho) { check(flb_set_field_rho(h_, rho.data())); }
  void setVelocityField(const std::vector<double>& u) { check(flb_set_field_u(h_, u.data())); }
  void setSmagorinsky(float cs) { smagorinsky_ = cs; }
  // CONSTFORCE data field of BGKForce<..., ConstForce> cells (d components).
  void setConstForce(const std::vector<double>& F) { check(flb_set_force(h_, F.data(), int(F.size()))); }

  // f_i = rho0 * w_i everywhere (BlockLatticeManager ctor, block_lattice.hh:676-681).
  void initUniform(double rho0 = 1.0) {
    std::vector<double> rhou(n_ * (LatSet::d + 1), 0.0);
    for (size_t k = 0; k < n_; ++k) rhou[k] = rho0;
    check(flb_init_equilibrium(h_, rhou.data()));
  }
  void initEquilibrium(const std::vector<double>& rhou) { check(flb_init_equilibrium(h_, rhou.data())); }

  std::vector<T> populations() const {
    std::vector<T> p(n_ * LatSet::q);
    if constexpr (sizeof(T) == 8) check(flb_download_pops_f64(h_, p.data()));
    else check(flb_download_pops_f32(h_, p.data()));
    return p;
  }
  void setPopulations(const std::vector<T>& p) {
    if constexpr (sizeof(T) == 8) check(flb_upload_pops_f64(h_, p.data()));
    else check(flb_upload_pops_f32(h_, p.data()));
  }

  // BlockLattice::ApplyCellDynamics<TaskSelector>(flagarr) (block_lattice.hh:72-92).
  template <typename Selector>
  void ApplyCellDynamics(const std::vector<uint8_t>& flags) {
    if (flags_src_ != &flags) setFlags(flags);
    const std::vector<flb_task> t = Selector::table(smagorinsky_);
    if (t.size() != tasks_.size() || !same_tasks(t)) {
      check(flb_set_tasks(h_, t.data(), int(t.size())));
      tasks_ = t;
    }
    check(flb_apply_cell_dynamics(h_, count_));
  }
  // BlockLatticeManager::ApplyCellDynamics(std::int64_t count, BFM)
  // (block_lattice.hh:751-764): steps the block when its level is due at count.
  template <typename Selector>
  void ApplyCellDynamics(std::int64_t count, const std::vector<uint8_t>& flags) {
    count_ = count;
    ApplyCellDynamics<Selector>(flags);
    count_ = 0;
  }
  // BlockLattice::Stream (block_lattice.hh:64-70): completes the fused step;
  // Stream(count) is the level-gated form (block_lattice.hh:737-742).
  void Stream() { check(flb_stream(h_, 0)); }
  void Stream(std::int64_t count) { check(flb_stream(h_, count)); }
  // BasicBlock level and BlockLatticeManager::getMaxLevel of the hierarchy.
  void setLevel(int level, int max_level) { check(flb_set_level(h_, level, max_level)); }

  // rho, u of cells with flag & mask (moment::rhoU<CELL, true>): (d+1)*N SoA.
  std::vector<T> moments(uint8_t mask) const {
    std::vector<T> out(n_ * (LatSet::d + 1));
    check(flb_moments(h_, mask, out.data()));
    return out;
  }

  // RHO / VELOCITY fields as the cells left them ((d+1)*N SoA; rho 1, u 0
  // where none exist): what moment::rhoU<CELL, true> cells wrote.
  std::vector<double> fields() const {
    std::vector<double> out(n_ * (LatSet::d + 1));
    check(flb_download_fields(h_, out.data()));
    return out;
  }

  // BlockLattice::EnableToleranceU / getToleranceU (getTolU(shift)) and the RHO forms
  void EnableTolerance() { check(flb_enable_tolerance(h_)); }
  double getToleranceU(int shift = 0) {
    double r = 0.0;
    check(flb_tolerance(h_, 1, shift, &r));
    return r;
  }
  double getToleranceRho(int shift = 0) {
    double r = 0.0;
    check(flb_tolerance(h_, 0, shift, &r));
    return r;
  }

  // A scalar lattice's VELOCITY field reference: FreeLB passes
  // &NSLattice.getField<VELOCITY>() to the scalar lattice's constructor
  // (examples/cazsblock2d/cazsblock2d.cpp:240-251).
  template <typename FlowLat>
  void shareVelocity(BlockLattice<T, FlowLat>& flow) {
    check(flb_share_velocity(h_, flow.handle()));
  }
  // The GBETA / RHOINIT data of a scalar lattice (constant fields in FreeLB).
  void setGBeta(double g) { gbeta_ = g; }
  void setRhoInit(double r) { rho_init_ = r; }
  double gbeta() const { return gbeta_; }
  double rhoInit() const { return rho_init_; }

  // called by the halfway managers
  void addHalfway(const flb_hwbc& bc) {
    bcs_.push_back(bc);
    check(flb_set_halfway_bc(h_, bcs_.data(), int(bcs_.size())));
  }

 private:
  bool same_tasks(const std::vector<flb_task>& t) const {
    for (size_t k = 0; k < t.size(); ++k)
      if (t[k].mask != tasks_[k].mask || t[k].dynamics != tasks_[k].dynamics ||
          t[k].smagorinsky != tasks_[k].smagorinsky)
        return false;
    return true;
  }

  flb_block h_ = nullptr;
  size_t n_ = 0;
  std::int64_t count_ = 0;  // count of the level-gated call in progress (0: always due)
  const std::vector<uint8_t>* flags_src_ = nullptr;
  std::vector<flb_task> tasks_;
  std::vector<flb_hwbc> bcs_;
  float smagorinsky_ = 0.0f;
  double gbeta_ = 0.0, rho_init_ = 1.0;
};

// ---- coupled lattices (data_struct/block_lattice.h:436-649) --------------------------
// BlockLatManagerCoupling(flow, scalar).ApplyCellDynamics<CoupledTaskSelector<...
// force::Buoyancy ...>>(flags) registers the buoyancy of each task's mask
// once; the force itself is formed inside the flow lattice's next
// collision from the scalar field the scalar lattice's last step wrote, which
// is the value FreeLB's coupling pass computes in its loop
// (SCALARFORCE.InitValue(0); coupling pass; flow step; scalar step,
// examples/cazsblock2d/cazsblock2d.cpp:385-409).
template <typename T, typename FlowLat, typename ScalarLat>
class BlockLatManagerCoupling {
 public:
  BlockLatManagerCoupling(BlockLattice<T, FlowLat>& flow, BlockLattice<T, ScalarLat>& scalar)
      : flow_(flow), scalar_(scalar) {}
  BlockLattice<T, FlowLat>& getLat0() { return flow_; }
  BlockLattice<T, ScalarLat>& getLat1() { return scalar_; }

  template <typename Selector>
  void ApplyCellDynamics(const std::vector<uint8_t>&) {
    if (registered_) return;
    for (uint8_t m : Selector::masks())
      check(flb_add_buoyancy(flow_.handle(), scalar_.handle(), m, scalar_.gbeta(), scalar_.rhoInit()));
    registered_ = true;
  }

 private:
  BlockLattice<T, FlowLat>& flow_;
  BlockLattice<T, ScalarLat>& scalar_;
  bool registered_ = false;
};

// ---- two refinement levels (flb_refine_*) ---------------------------------------------
// A coarse block and a fine block at twice its resolution covering the coarse
// cells [origin, origin + m) (fine extent 2m + 4, overlap 1). FreeLB's refined
// loop (examples/cavref2d/cavref2d.cpp:221-228) maps onto it as
//   pair.ApplyCellDynamics<Sel>(count, coarse_flags, fine_flags);
//   pair.Stream(count);                 // BM.Apply(count) is folded into the step
//   pair.FullCommunicate(count);        // averCommunicate / intpCommunicate
// The fine level steps every count, the coarse level on even counts
// (block_lattice.hh:736-764). Run(count0, n) does n counts in one call.
template <typename T, typename LatSet>
class RefinedPair {
 public:
  RefinedPair(BlockLattice<T, LatSet>& coarse, BlockLattice<T, LatSet>& fine, const int origin[3],
              void* stream = nullptr)
      : coarse_(coarse), fine_(fine), stream_(stream) {
    for (int a = 0; a < 3; ++a) origin_[a] = origin[a];
  }
  ~RefinedPair() { flb_refine_destroy(r_); }
  RefinedPair(const RefinedPair&) = delete;
  RefinedPair& operator=(const RefinedPair&) = delete;

  template <typename Selector>
  void ApplyCellDynamics(int64_t count, const std::vector<uint8_t>& coarse_flags,
                         const std::vector<uint8_t>& fine_flags) {
    fine_.template ApplyCellDynamics<Selector>(fine_flags);
    if (count % 2 == 0) coarse_.template ApplyCellDynamics<Selector>(coarse_flags);
  }
  void Stream(int64_t count) {
    fine_.Stream();
    if (count % 2 == 0) coarse_.Stream();
  }
  void FullCommunicate(int64_t count) { check(flb_refine_communicate(handle(), count)); }
  // the pair's AverageCommunicate(count) / InterpolateCommunicate(count), for
  // hierarchies that run every pair's average before any interpolation
  void AverageCommunicate(int64_t count) { check(flb_refine_average(handle(), count)); }
  void InterpolateCommunicate(int64_t count) { check(flb_refine_interpolate(handle(), count)); }
  // levels of the pair in a deeper hierarchy (coarse level, finest level)
  void setLevels(int coarse_level, int max_level) { check(flb_refine_set_levels(handle(), coarse_level, max_level)); }
  void Run(int64_t count0, int n) { check(flb_refine_run(handle(), count0, n)); }

  // created on first use: both blocks need their flags and tasks first
  flb_refine handle() {
    if (!r_) check(flb_refine_create(coarse_.handle(), fine_.handle(), origin_, stream_, &r_));
    return r_;
  }

 private:
  BlockLattice<T, LatSet>& coarse_;
  BlockLattice<T, LatSet>& fine_;
  void* stream_ = nullptr;
  int origin_[3] = {0, 0, 0};
  flb_refine r_ = nullptr;
};

// ---- many blocks: BlockLatticeManager (data_struct/block_lattice.hh:744-778) -------
// Steps all of its blocks with one launch per step (flb_batch_*). The blocks
// keep their own set-up (flags, tasks, fields) and read-out; FreeLB's
// per-block ApplyCellDynamics / Stream pair becomes ApplyCellDynamics +
// Stream on the manager, and many steps can be recorded once and replayed.
template <typename T, typename LatSet>
class BlockLatticeManager {
 public:
  explicit BlockLatticeManager(std::vector<BlockLattice<T, LatSet>*> blocks, void* stream = nullptr)
      : blocks_(std::move(blocks)) {
    std::vector<flb_block> h;
    for (auto* b : blocks_) h.push_back(b->handle());
    check(flb_batch_create(h.data(), int(h.size()), stream, &t_));
  }
  ~BlockLatticeManager() { flb_batch_destroy(t_); }
  BlockLatticeManager(const BlockLatticeManager&) = delete;
  BlockLatticeManager& operator=(const BlockLatticeManager&) = delete;

  // One step of every block; the blocks' task tables must already be set
  // (by a first per-block ApplyCellDynamics<Selector> + Stream, or set-up).
  void ApplyCellDynamics() { pending_ = true; }
  void Stream() {
    if (!pending_) throw Error(FLB_ERR_STATE, "flb: Stream without ApplyCellDynamics");
    pending_ = false;
    check(flb_batch_step(t_, 1));
  }
  void step(int n) { check(flb_batch_step(t_, n)); }
  void capture(int n) { check(flb_batch_capture(t_, n)); }
  void replay() { check(flb_batch_replay(t_)); }
  flb_launch_cfg autotune(int steps = 10) {
    flb_launch_cfg best{};
    double ms = 0.0;
    check(flb_batch_autotune(t_, steps, &best, &ms));
    return best;
  }
  void synchronize() { check(flb_batch_synchronize(t_)); }
  BlockLattice<T, LatSet>& getBlockLat(size_t k) { return *blocks_[k]; }

 private:
  std::vector<BlockLattice<T, LatSet>*> blocks_;
  flb_batch t_ = nullptr;
  bool pending_ = false;
};

// ---- halfway managers (boundary/boundary.h:75-129, basic_boundary.h:331-348) -------
class AbstractBlockBoundary {
 public:
  virtual ~AbstractBlockBoundary() = default;
  virtual void Apply() = 0;
};

// BBLikeFixedBlockBdManager<bounceback::normal|movingwall<CELL>, ...>: registers the
// manager's boundary cells (flag & cellflag) and wall flag (voidflag) with the
// lattice; the fix-up itself runs inside the fused update kernel.
template <typename CELLDYNAMICS, typename LATTICE>
class BBLikeFixedBlockBdManager final : public AbstractBlockBoundary {
 public:
  BBLikeFixedBlockBdManager(std::string name, LATTICE& lat, std::uint8_t cellflag,
                            std::uint8_t voidflag = std::uint8_t(1))
      : name_(std::move(name)), lat_(lat) {
    lat_.addHalfway(flb_hwbc{cellflag, voidflag, CELLDYNAMICS::kind, 0});
  }
  void Apply() override { check(flb_apply_halfway_bc(lat_.handle())); }
  const std::string& name() const { return name_; }

 private:
  std::string name_;
  LATTICE& lat_;
};

class BlockBoundaryManager {
 public:
  template <typename... Args>
  explicit BlockBoundaryManager(Args*... args) : bds_{args...} {}
  void Apply() {
    for (AbstractBlockBoundary* b : bds_) b->Apply();
  }

 private:
  std::vector<AbstractBlockBoundary*> bds_;
};

}  // namespace flb
