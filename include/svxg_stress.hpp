// svxg_stress.hpp — a C++20 StressProblem over the device library's C ABI (svxg.h).
//
// The same public surface as the reference's svx::StressProblem
// (core/include/svx/stress/stress.hpp:74-160): public `nodes`, `bonds`, `mats` and materials(); assemble(),
// assembled(), invalidate(), preconditioner_current(), block_jacobi(), assembly_work(), release(),
// remove_bond(), retire_nodes(), reassemble(), extend_warm_start(), appended(), free_nodes(),
// matrix_blocks(), memory_bytes(), solver_bytes(), work_per_iteration(), solve(),
// begin() / iterate() / running() / stop() / current(), bond_load(), add_force(); and the free
// functions bond_strength() / bond_utilization() (stress.hpp:53-66). Callers mutate nodes / bonds and
// then assemble(), as World does; the device copy follows on the next assemble().
//
// The material table is a template parameter (BasicStressProblem<Table>): any table whose
// operator[](u8) gives a material with the fields E, G, rho, ft, fb, fc, cohesion and friction - the
// reference's svx::MaterialTable (core/include/svx/material/material.hpp:91-113) as it is, or
// svxg::MaterialTable below. memory_bytes() / solver_bytes() take any enum with Held / Used (the
// reference's svx::Bytes, core/include/svx/base/mem.hpp:14). So World's call sites (`P.mats =
// mats_.get()`, world.cpp:455; `memory_bytes(k)`, world.cpp:2493; `bond_utilization(B, L,
// fragility, mats())`) compile unchanged against `using StressProblem =
// svxg::BasicStressProblem<svx::MaterialTable>`.
//
// Header-only; link with libsvxg.so. No exceptions (the reference's libraries build without them).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "svxg.h"

namespace svxg {

using f32 = float;
using f64 = double;
using i32 = int32_t;
using i64 = int64_t;
using u8 = uint8_t;

struct V3 {
  f64 x = 0, y = 0, z = 0;
};

// SNode / SBond (stress.hpp:21-51): the same fields
struct SNode {
  V3 c;
  f64 mass = 0.0;
  bool fixed = false;
  bool gone = false;
};

struct SBond {
  i32 a = -1, b = -1;
  bool broken = false;
  bool in_k = false;
  u8 ma = 1, mb = 1;  // (material ids; the reference's default is Concrete = 1)
  V3 n{0, 0, 1}, t1{1, 0, 0}, t2{0, 1, 0};
  f64 area = 0.0;
  V3 p;
  f64 s1 = 0.0, s2 = 0.0, c1 = 0.0, c2 = 0.0, rmax = 0.0, la = 0.0, lb = 0.0;
  f64 strength = 1.0;
  bool sectioned = false;
  f32 ft = 0, fb = 0, fc = 0, coh = 0, mu = 0;
  i32 faces = 0, tag = -1;
};

using Material = svxg_material;
using BondLoad = svxg_bond_load;
enum class FailMode : u8 { None = 0, Tension, Crush, Shear };
// what a memory measure counts (the reference's Bytes, mem.hpp:10-14): Held what containers hold
// (their capacities), Used what they use (their elements)
enum class Bytes : u8 { Held, Used };

// A material table shaped like the reference's (material.hpp:91-113): 127 ids, operator[] of an id
// never set gives id 0's material. Default: id 0 reinforced concrete, id 1 concrete (the reference's
// presets' strengths; a world sets its own).
constexpr int kMaxMaterials = 127;
class MaterialTable {
 public:
  MaterialTable() {
    m_[0] = Material{30e9, 12.5e9, 2400, 1.2e6, 3.5e6, 30e6, 1.6e6, 0.7};
    m_[1] = Material{28e9, 11.5e9, 2300, 0.8e6, 1.2e6, 25e6, 1.2e6, 0.7};
    used_[0] = used_[1] = true;
  }
  // ids 0 .. size() - 1 from a list
  explicit MaterialTable(const std::vector<Material>& list) {
    for (size_t i = 0; i < list.size() && i < size_t(kMaxMaterials); ++i) set(u8(i), list[i]);
  }
  const Material& operator[](u8 id) const { return id < kMaxMaterials && used_[id] ? m_[id] : m_[0]; }
  const Material& operator[](int id) const { return (*this)[u8(id)]; }
  void set(u8 id, const Material& m) {
    if (id >= kMaxMaterials) return;
    m_[id] = m;
    used_[id] = true;
  }
  bool registered(u8 id) const { return id < kMaxMaterials && used_[id]; }

 private:
  Material m_[kMaxMaterials]{};
  bool used_[kMaxMaterials]{};
};

// the C ABI's material record of a table entry
template <class M>
inline svxg_material to_c_material(const M& m) {
  return svxg_material{f64(m.E), f64(m.G), f64(m.rho), f64(m.ft), f64(m.fb), f64(m.fc), f64(m.cohesion), f64(m.friction)};
}

struct StressOptions {
  f64 rtol = 2e-3;
  svxg_amg_options amg{0.1, 0.6, 48, 1, 16, 1, 2, 1, 1, 0, 1, 16384, 0.05};
  i32 amg_min_nodes = 160;
  i32 precision = SVXG_FP64;
  i32 mixed_fp64_rows = 1 << 30;  // (device: see svxg_stress_options)
  i32 vcycle_fp16_rows = 16384;   // (device: see svxg_stress_options)
};

using PcgResult = svxg_pcg_result;

inline svxg_node to_c(const SNode& s) {
  svxg_node n{};
  n.c[0] = s.c.x;
  n.c[1] = s.c.y;
  n.c[2] = s.c.z;
  n.mass = s.mass;
  n.fixed = s.fixed;
  n.gone = s.gone;
  return n;
}

inline svxg_bond to_c(const SBond& s) {
  svxg_bond b{};
  b.a = s.a;
  b.b = s.b;
  b.broken = s.broken;
  b.in_k = s.in_k;
  b.ma = s.ma;
  b.mb = s.mb;
  const V3* v[3] = {&s.n, &s.t1, &s.t2};
  double* d[3] = {b.n, b.t1, b.t2};
  for (int k = 0; k < 3; ++k) {
    d[k][0] = v[k]->x;
    d[k][1] = v[k]->y;
    d[k][2] = v[k]->z;
  }
  b.area = s.area;
  b.p[0] = s.p.x;
  b.p[1] = s.p.y;
  b.p[2] = s.p.z;
  b.s1 = s.s1;
  b.s2 = s.s2;
  b.c1 = s.c1;
  b.c2 = s.c2;
  b.rmax = s.rmax;
  b.la = s.la;
  b.lb = s.lb;
  b.strength = s.strength;
  b.sectioned = s.sectioned;
  b.ft = s.ft;
  b.fb = s.fb;
  b.fc = s.fc;
  b.coh = s.coh;
  b.mu = s.mu;
  b.faces = s.faces;
  b.tag = s.tag;
  return b;
}

// One context (device + HBM budget) shared by the process's problems.
inline svxg_ctx* default_context() {
  static svxg_ctx* c = [] {
    svxg_ctx* h = nullptr;
    return svxg_ctx_create(0, 0, &h) == SVXG_OK ? h : nullptr;
  }();
  return c;
}

template <class Table = MaterialTable>
class BasicStressProblem {
 public:
  std::vector<SNode> nodes;
  std::vector<SBond> bonds;
  const Table* mats = nullptr;  // the bonds' materials (a world's; nullptr: the default table)
  const Table& materials() const { return mats ? *mats : default_table(); }

  BasicStressProblem() = default;
  BasicStressProblem(const BasicStressProblem&) = delete;
  BasicStressProblem& operator=(const BasicStressProblem&) = delete;
  ~BasicStressProblem() { drop(); }

  bool assemble(const StressOptions& opt = {}) {
    opt_ = opt;
    drop();
    std::vector<svxg_node> n(nodes.size());
    std::vector<svxg_bond> b(bonds.size());
    for (size_t i = 0; i < nodes.size(); ++i) n[i] = to_c(nodes[i]);
    for (size_t i = 0; i < bonds.size(); ++i) b[i] = to_c(bonds[i]);
    // (every id of the table: an id a bond names resolves as the table resolves it)
    std::vector<svxg_material> m(static_cast<size_t>(kMaxMaterials));
    for (int i = 0; i < kMaxMaterials; ++i) m[size_t(i)] = to_c_material(materials()[u8(i)]);
    if (svxg_problem_create(default_context(), n.data(), i32(n.size()), b.data(), i32(b.size()), m.data(),
                            i32(m.size()), &p_) != SVXG_OK)
      return false;
    svxg_stress_options o{opt.rtol, opt.amg_min_nodes, opt.precision, opt.amg, opt.mixed_fp64_rows, opt.vcycle_fp16_rows};
    const int st = svxg_problem_assemble(p_, &o);
    pull_flags();
    return st == SVXG_OK;
  }
  bool assembled() const { return info().assembled != 0; }
  void invalidate() {
    if (p_) svxg_problem_invalidate(p_);
  }
  bool preconditioner_current() const { return info().preconditioner_current != 0; }
  bool block_jacobi() const { return info().block_jacobi != 0; }
  i64 assembly_work() const { return info().assembly_work; }
  void release() {
    if (p_) svxg_problem_release(p_);
    for (SBond& b : bonds) b.in_k = false;
  }
  // (the bond's own flags follow the library's: broken, and out of an assembled K - the other bonds'
  // do not change, so a judge's picks cost no copy of the bond list each)
  void remove_bond(i32 b) {
    bonds[size_t(b)].broken = true;
    if (p_) {
      svxg_problem_remove_bonds(p_, &b, 1);
      if (assembled()) bonds[size_t(b)].in_k = false;
    }
  }
  void retire_nodes(const std::vector<i32>& list) {
    for (i32 i : list) nodes[size_t(i)].gone = true;
    if (p_) {
      // (the caller's topology flags go down first: a bond restored to intact stays out of K)
      for (size_t k = 0; k < bonds.size(); ++k) svxg_problem_set_bond_broken(p_, i32(k), bonds[k].broken);
      svxg_problem_retire_nodes(p_, list.data(), i32(list.size()));
      pull_flags();
    }
  }
  // K again from nodes and bonds as they are (stress.cpp:178-211). Here the whole problem is set up
  // again, appended nodes included in the new multigrid, so nothing is left for appended() to count.
  bool reassemble() { return assemble(opt_); }
  // stress.cpp:213-250: u of the nodes from first_new on, carried from their older neighbours
  void extend_warm_start(std::vector<f64>& u, i32 first_new) const {
    if (u.size() != 6 * nodes.size()) return;
    std::vector<svxg_node> n(nodes.size());
    std::vector<svxg_bond> b(bonds.size());
    for (size_t i = 0; i < nodes.size(); ++i) n[i] = to_c(nodes[i]);
    for (size_t i = 0; i < bonds.size(); ++i) b[i] = to_c(bonds[i]);
    svxg_extend_warm_start(n.data(), i32(n.size()), b.data(), i32(b.size()), u.data(), first_new);
  }
  i32 appended() const { return 0; }
  i32 free_nodes() const { return info().nfree; }
  i64 matrix_blocks() const { return info().blocks; }
  i64 work_per_iteration() const { return info().work_per_iteration; }
  // (stress.cpp:341-347: the graph's vectors - capacities for Held, sizes for Used - and the solver's
  // device bytes: the operator, hierarchy and solve vectors that release() frees, the same either way)
  template <class K = Bytes>
  i64 solver_bytes(K kind = K::Held) const {
    (void)kind;
    return info().device_bytes;
  }
  template <class K = Bytes>
  i64 memory_bytes(K kind = K::Held) const {
    const bool held = kind == K::Held;
    return i64((held ? nodes.capacity() : nodes.size()) * sizeof(SNode) +
               (held ? bonds.capacity() : bonds.size()) * sizeof(SBond)) +
           solver_bytes(kind);
  }

  PcgResult solve(const std::vector<f64>& f, std::vector<f64>& u, f64 rtol, int maxit, bool warm) {
    PcgResult r{};
    if (!p_ && !assemble(opt_)) {
      r.rel_res = INFINITY;
      r.breakdown = 1;
      return r;
    }
    u.resize(6 * nodes.size(), 0.0);
    sync_flags();
    svxg_problem_solve(p_, f.data(), u.data(), rtol, maxit, warm ? 1 : 0, &r);
    return r;
  }
  void begin(const std::vector<f64>& f, const std::vector<f64>& u0) {
    if (!p_ && !assemble(opt_)) return;
    sync_flags();
    svxg_problem_begin(p_, f.data(), u0.size() == 6 * nodes.size() ? u0.data() : nullptr);
    running_ = true;
  }
  PcgResult iterate(int maxit, f64 rtol) {
    PcgResult r{};
    if (p_ && running_) svxg_problem_iterate(p_, maxit, rtol, &r);
    return r;
  }
  bool running() const { return running_; }
  void stop() { running_ = false; }
  void current(std::vector<f64>& u) const {
    u.assign(6 * nodes.size(), 0.0);
    if (p_) svxg_problem_current(p_, u.data());
  }
  // Bond b's load from u (stress.cpp:477-502). The device computes every bond's load from u in one
  // pass and this keeps them with a copy of u: a later call whose bond reads the same values of its
  // two nodes in u gets its load from there (a bond's load depends on those twelve values alone), so
  // World's loop over every bond (world.cpp:1171-1193) costs one device pass, not one a bond.
  BondLoad bond_load(i32 b, const std::vector<f64>& u) const {
    if (!p_ || b < 0 || size_t(b) >= bonds.size() || u.size() != 6 * nodes.size()) return BondLoad{};
    const SBond& B = bonds[size_t(b)];
    if (!loads_.empty() && same_nodes(u, B.a) && (B.b < 0 || same_nodes(u, B.b))) return loads_[size_t(b)];
    loads_.assign(bonds.size(), BondLoad{});
    if (svxg_problem_bond_loads(p_, u.data(), loads_.data()) != SVXG_OK) {
      loads_.clear();
      return BondLoad{};
    }
    loads_u_ = u;
    return loads_[size_t(b)];
  }
  // every bond's load at once (the judge loop's form: one device pass)
  std::vector<BondLoad> bond_loads(const std::vector<f64>* u = nullptr) const {
    std::vector<BondLoad> all(bonds.size());
    if (p_) svxg_problem_bond_loads(p_, u ? u->data() : nullptr, all.data());
    return all;
  }
  static void add_force(std::vector<f64>& f, i32 node, const V3& c, const V3& F, const V3& at) {
    const double cc[3] = {c.x, c.y, c.z}, FF[3] = {F.x, F.y, F.z}, aa[3] = {at.x, at.y, at.z};
    svxg_add_force(f.data(), node, cc, FF, aa);
  }
  svxg_problem* handle() const { return p_; }

 private:
  // node i's six values in u are those the kept loads were made from (bitwise)
  bool same_nodes(const std::vector<f64>& u, i32 i) const {
    return loads_u_.size() == u.size() && i >= 0 &&
           std::memcmp(u.data() + 6 * size_t(i), loads_u_.data() + 6 * size_t(i), 6 * sizeof(f64)) == 0;
  }
  svxg_problem_info info() const {
    svxg_problem_info i{};
    if (p_) svxg_problem_info_get(p_, &i);
    return i;
  }
  void pull_flags() {
    std::vector<svxg_bond> b(bonds.size());
    if (!p_ || b.empty()) return;
    svxg_problem_get_bonds(p_, b.data());
    for (size_t k = 0; k < bonds.size(); ++k) {
      bonds[k].broken = b[k].broken != 0;
      bonds[k].in_k = b[k].in_k != 0;
    }
  }
  void sync_flags() {
    for (size_t k = 0; k < bonds.size(); ++k) svxg_problem_set_bond_broken(p_, i32(k), bonds[k].broken);
  }
  void drop() {
    if (p_) svxg_problem_destroy(p_);
    p_ = nullptr;
    running_ = false;
    loads_.clear();
    loads_u_.clear();
  }
  static const Table& default_table() {
    static const Table t{};
    return t;
  }
  svxg_problem* p_ = nullptr;
  mutable std::vector<BondLoad> loads_;  // (bond_load: every bond's load from loads_u_)
  mutable std::vector<f64> loads_u_;
  StressOptions opt_{};
  bool running_ = false;
};

// bond_strength / bond_utilization (stress.hpp:53-66; stress.cpp:14-57) of one bond on the host
// (the device runs them for every bond: svxg_problem_bond_checks).
struct BondStrength {
  f64 ft, fb, fc, coh, mu;
};

template <class Table>
inline BondStrength bond_strength(const SBond& b, f64 fragility, const Table& mats) {
  const f64 s = b.strength / std::max(1e-6, fragility);
  if (b.sectioned) return {s * b.ft, s * b.fb, s * b.fc, s * b.coh, f64(b.mu)};
  const auto& A = mats[b.ma];
  const auto& B = mats[b.mb];
  return {s * std::min(A.ft, B.ft), s * std::min(A.fb, B.fb), s * std::min(A.fc, B.fc),
          s * std::min(A.cohesion, B.cohesion), std::min(A.friction, B.friction)};
}

template <class Table>
inline f64 bond_utilization(const SBond& b, const BondLoad& L, f64 fragility, const Table& mats,
                            FailMode* mode = nullptr) {
  BondStrength S = bond_strength(b, fragility, mats);
  const f64 weak = 1e-9;
  S.ft = std::max(S.ft, weak);
  S.fb = std::max(S.fb, weak);
  S.fc = std::max(S.fc, weak);
  const f64 A = std::max(b.area, 1e-12);
  const f64 sN = L.N / A;
  const f64 sb = (b.s2 > 0 ? std::abs(L.M1) * b.c2 / b.s2 : 0.0) + (b.s1 > 0 ? std::abs(L.M2) * b.c1 / b.s1 : 0.0);
  const f64 phi_t = sN >= 0.0 ? sN / S.ft + sb / S.fb : std::max(0.0, sb + sN) / S.fb;
  const f64 phi_c = std::max(0.0, sb - sN) / S.fc;
  const f64 J = std::max(b.s1 + b.s2, 1e-18);
  const f64 tau = 1.5 * std::sqrt(L.V1 * L.V1 + L.V2 * L.V2) / A + std::abs(L.T) * b.rmax / J;
  const f64 cap = std::max(S.coh + S.mu * std::max(0.0, -sN), weak);
  const f64 phi_s = tau / cap;
  f64 phi = phi_t;
  FailMode m = FailMode::Tension;
  if (phi_c > phi) {
    phi = phi_c;
    m = FailMode::Crush;
  }
  if (phi_s > phi) {
    phi = phi_s;
    m = FailMode::Shear;
  }
  if (mode) *mode = phi > 0 ? m : FailMode::None;
  return phi;
}

using StressProblem = BasicStressProblem<>;

}  // namespace svxg
