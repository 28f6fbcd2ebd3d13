// World's own loops over a structure's bonds, written as World writes them, compiled against the C++
// StressProblem of include/svxg_stress.hpp with a material table of World's shape (not the
// library's): the judge loop (core/src/world/world.cpp:1171-1193: a bond_load and a
// bond_utilization a bond on the structure's u), probe_utilization (world.cpp:1925-1938: the same,
// then the worst bond's load again), `P.mats = table` (world.cpp:455) and memory_bytes(Bytes::Used)
// (world.cpp:2493). Run on the 64 x 64 x 48 block (422,375 bonds): the loop's phi must equal the
// device's bond checks bit for bit, and a loop over every bond must take one device pass (timed).
// Prints one line a check; exit status = the number of failed checks.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "svxg_stress.hpp"

extern "C" {
typedef struct svxgen_graph svxgen_graph;
int32_t svxgen_materials(svxg_material* out);
svxgen_graph* svxgen_block(int32_t nx, int32_t ny, int32_t nz, double L, double voids, uint64_t seed);
void svxgen_counts(const svxgen_graph* h, int32_t* n, int32_t* m);
void svxgen_copy(const svxgen_graph* h, svxg_node* nodes, svxg_bond* bonds, double* f);
void svxgen_free(svxgen_graph* h);
}

static int g_fail = 0;
#define CHECK(cond)                                                                  \
  do {                                                                               \
    const bool ok_ = (cond);                                                         \
    std::printf("%s %s:%d %s\n", ok_ ? "ok  " : "FAIL", __FILE__, __LINE__, #cond); \
    if (!ok_) ++g_fail;                                                              \
  } while (0)

// A world's material table: its own record type and id lookup, as material.hpp:47-113 has them
// (the fields the stress path reads, plus some it does not)
namespace world {
enum class MaterialId : uint8_t { Air = 0, Concrete = 1 };
enum class Bytes : uint8_t { Held, Used };
struct Material {
  const char* name = "m";
  double E = 30e9, G = 12.5e9, rho = 2400, ft = 1e6, fb = 1e6, fc = 30e6, cohesion = 1e6, friction = 0.7;
  double Gf = 400;
  bool ductile = false;
};
class MaterialTable {
 public:
  const Material& operator[](MaterialId id) const { return (*this)[static_cast<uint8_t>(id)]; }
  const Material& operator[](uint8_t id) const { return id < n_ ? m_[id] : m_[0]; }
  void add(const Material& m) { m_[n_++] = m; }

 private:
  Material m_[127];
  uint8_t n_ = 0;
};
}  // namespace world

using StressProblem = svxg::BasicStressProblem<world::MaterialTable>;
using svxg::BondLoad;
using svxg::f64;
using svxg::i32;

struct Structure {
  StressProblem P;
  std::vector<f64> u;
};

int main() {
  world::MaterialTable table;
  {
    svxg_material m[32];
    const int k = svxgen_materials(m);
    for (int i = 0; i < k; ++i) {
      world::Material w;
      w.E = m[i].E;
      w.G = m[i].G;
      w.rho = m[i].rho;
      w.ft = m[i].ft;
      w.fb = m[i].fb;
      w.fc = m[i].fc;
      w.cohesion = m[i].cohesion;
      w.friction = m[i].friction;
      table.add(w);
    }
  }
  auto mats = [&]() -> const world::MaterialTable& { return table; };
  svxgen_graph* h = svxgen_block(64, 64, 48, 0.5, 0.15, 1);
  int32_t n = 0, m = 0;
  svxgen_counts(h, &n, &m);
  std::vector<svxg_node> cn(static_cast<size_t>(n));
  std::vector<svxg_bond> cb(static_cast<size_t>(m));
  std::vector<f64> F(6 * size_t(n));
  svxgen_copy(h, cn.data(), cb.data(), F.data());
  svxgen_free(h);
  Structure s;
  s.P.mats = &table;  // (world.cpp:455)
  for (const svxg_node& c : cn) {
    svxg::SNode nd;
    nd.c = {c.c[0], c.c[1], c.c[2]};
    nd.mass = c.mass;
    nd.fixed = c.fixed;
    nd.gone = c.gone;
    s.P.nodes.push_back(nd);
  }
  for (const svxg_bond& c : cb) {
    svxg::SBond b;
    b.a = c.a;
    b.b = c.b;
    b.ma = c.ma;
    b.mb = c.mb;
    b.n = {c.n[0], c.n[1], c.n[2]};
    b.t1 = {c.t1[0], c.t1[1], c.t1[2]};
    b.t2 = {c.t2[0], c.t2[1], c.t2[2]};
    b.area = c.area;
    b.p = {c.p[0], c.p[1], c.p[2]};
    b.s1 = c.s1;
    b.s2 = c.s2;
    b.c1 = c.c1;
    b.c2 = c.c2;
    b.rmax = c.rmax;
    b.la = c.la;
    b.lb = c.lb;
    b.strength = c.strength;
    b.sectioned = c.sectioned;
    b.faces = c.faces;
    s.P.bonds.push_back(b);
  }
  std::printf("block: %d nodes, %d bonds\n", n, m);
  svxg::StressOptions opt;
  opt.precision = SVXG_MIXED;
  CHECK(s.P.assemble(opt));
  const svxg_pcg_result r = s.P.solve(F, s.u, 1e-8, 5000, false);
  CHECK(r.converged);
  // the device's checks of the same u, for the bits
  std::vector<f64> phi_dev(static_cast<size_t>(m));
  std::vector<uint8_t> mode_dev(static_cast<size_t>(m));
  CHECK(svxg_problem_bond_checks(s.P.handle(), s.u.data(), 1.0, phi_dev.data(), mode_dev.data()) == SVXG_OK);
  const f64 fragility = 1.0;
  for (int rep = 0; rep < 3; ++rep) {
    // world.cpp:1171-1193, the judge's loop
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::pair<f64, i32>> over;
    std::vector<float> phi_of(static_cast<size_t>(m), 0.0f);
    f64 maxphi = 0.0;
    size_t same = 0;
    const i32 nb = static_cast<i32>(s.P.bonds.size());
    for (i32 b = 0; b < nb; ++b) {
      const svxg::SBond& B = s.P.bonds[size_t(b)];
      if (B.broken) continue;
      const BondLoad L = s.P.bond_load(b, s.u);
      const f64 phi = bond_utilization(B, L, fragility, mats());
      phi_of[size_t(b)] = static_cast<float>(phi);
      same += std::memcmp(&phi, &phi_dev[size_t(b)], sizeof phi) == 0 ? 1 : 0;
      maxphi = std::max(maxphi, phi);
      if (phi >= 1.0) over.push_back({phi, b});
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("judge loop over %d bonds: %.2f ms (max phi %.4f, %zu over)\n", nb, ms, maxphi, over.size());
    CHECK(same == size_t(nb));
    if (rep > 0) CHECK(ms < 100.0);  // (one device pass: a pass a bond would take minutes)
    // the u of the next judgement changes in place (a new solve into the same vector)
    if (rep == 0) s.P.solve(F, s.u, 1e-10, 5000, true);
    if (rep == 0) svxg_problem_bond_checks(s.P.handle(), s.u.data(), 1.0, phi_dev.data(), mode_dev.data());
  }
  {
    // world.cpp:1925-1938, probe_utilization
    const auto t0 = std::chrono::steady_clock::now();
    f64 mx = 0.0;
    i32 cnt = 0, worst = -1;
    for (i32 b = 0; b < static_cast<i32>(s.P.bonds.size()); ++b) {
      if (s.P.bonds[size_t(b)].broken) continue;
      const f64 phi = bond_utilization(s.P.bonds[size_t(b)], s.P.bond_load(b, s.u), fragility, mats());
      if (phi > mx) worst = b;
      mx = std::max(mx, phi);
      cnt += phi >= 1.0 ? 1 : 0;
    }
    svxg::FailMode mode;
    const f64 pw = bond_utilization(s.P.bonds[size_t(worst)], s.P.bond_load(worst, s.u), fragility, mats(), &mode);
    const BondLoad Lw = s.P.bond_load(worst, s.u);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("probe loop: %.2f ms, worst bond %d phi %.4f mode %d N %.1f, %d over\n", ms, worst, pw,
                static_cast<int>(mode), Lw.N, cnt);
    CHECK(pw == mx && std::memcmp(&pw, &phi_dev[size_t(worst)], sizeof pw) == 0);
    CHECK(static_cast<uint8_t>(mode) == mode_dev[size_t(worst)]);
  }
  // a change to u seen by the bonds that read it: the loads are made again
  {
    std::vector<f64> u2 = s.u;
    const svxg::SBond& B0 = s.P.bonds[0];
    u2[6 * size_t(B0.a)] += 1e-3;
    const BondLoad L2 = s.P.bond_load(0, u2);
    std::vector<BondLoad> all(static_cast<size_t>(m));
    svxg_problem_bond_loads(s.P.handle(), u2.data(), all.data());
    CHECK(std::memcmp(&L2, &all[0], sizeof L2) == 0);
  }
  const svxg::i64 used = s.P.memory_bytes(world::Bytes::Used), held = s.P.memory_bytes(world::Bytes::Held);
  std::printf("memory_bytes: used %lld held %lld solver %lld\n", (long long)used, (long long)held,
              (long long)s.P.solver_bytes(world::Bytes::Used));
  CHECK(used > 0 && held >= used && s.P.solver_bytes() > 0);
  CHECK(&s.P.materials() == &table);
  std::printf("%d failed\n", g_fail);
  return g_fail;
}
