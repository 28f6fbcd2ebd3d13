// The reference's stress tests (tests/core/test_stress.cpp) restated over the C++ StressProblem of
// include/svxg_stress.hpp, so the device library is exercised exactly as a C++ caller (World) would
// use it. Prints one line a check; exit status = the number of failed checks.
#include <cmath>
#include <cstdio>
#include <vector>

#include "svxg_stress.hpp"

using namespace svxg;

static int g_fail = 0;
#define CHECK(cond)                                                          \
  do {                                                                       \
    const bool ok_ = (cond);                                                 \
    std::printf("%s %s:%d %s\n", ok_ ? "ok  " : "FAIL", __FILE__, __LINE__, #cond); \
    if (!ok_) ++g_fail;                                                      \
  } while (0)

static bool approx(double a, double b, double rel) { return std::abs(a - b) <= rel * std::max(std::abs(a), std::abs(b)); }

static V3 cross(const V3& a, const V3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static V3 add(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

static std::vector<Material> presets() {
  std::vector<Material> m(12);
  svxg_material rc{30e9, 12.5e9, 2400, 1.2e6, 3.5e6, 30e6, 1.6e6, 0.7};
  svxg_material concrete{28e9, 11.5e9, 2300, 0.8e6, 1.2e6, 25e6, 1.2e6, 0.7};
  for (auto& x : m) x = rc;
  m[1] = concrete;
  return m;
}
static const MaterialTable kMats(presets());

// test_stress.cpp:14-32
static SBond bond(const StressProblem& P, i32 A, i32 B, const V3& n, const V3& p, f64 a) {
  SBond s;
  s.a = A;
  s.b = B;
  s.ma = s.mb = 1;
  s.n = n;
  s.t1 = std::abs(n.x) > 0.5 ? V3{0, 1, 0} : V3{1, 0, 0};
  s.t2 = cross(s.n, s.t1);
  s.area = a * a;
  s.p = p;
  s.s1 = s.s2 = a * a * a * a / 12.0;
  s.c1 = s.c2 = 0.5 * a;
  s.rmax = std::sqrt(0.5) * a;
  const V3 ca = P.nodes[size_t(A)].c;
  s.la = std::abs(dot(sub(p, ca), n));
  s.lb = B >= 0 ? std::abs(dot(sub(P.nodes[size_t(B)].c, p), n)) : s.la;
  s.faces = 16;
  return s;
}

// test_stress.cpp:35-46
static void cantilever(StressProblem& P, int n, f64 L, f64 mass) {
  P.mats = &kMats;
  for (int i = 0; i < n; ++i) {
    SNode nd;
    nd.c = V3{(i + 0.5) * L, 0, 0};
    nd.mass = mass;
    P.nodes.push_back(nd);
  }
  P.bonds.push_back(bond(P, 0, -1, V3{-1, 0, 0}, V3{0, 0, 0}, L));
  for (int i = 0; i + 1 < n; ++i) P.bonds.push_back(bond(P, i, i + 1, V3{1, 0, 0}, V3{(i + 1) * L, 0, 0}, L));
}

static std::vector<f64> gravity(const StressProblem& P) {
  std::vector<f64> f(6 * P.nodes.size(), 0.0);
  for (size_t i = 0; i < P.nodes.size(); ++i)
    if (!P.nodes[i].gone) f[6 * i + 2] = -9.81 * P.nodes[i].mass;
  return f;
}

static f64 shear(const BondLoad& L) { return std::sqrt(L.V1 * L.V1 + L.V2 * L.V2); }
static f64 moment(const BondLoad& L) { return std::sqrt(L.M1 * L.M1 + L.M2 * L.M2); }

static void cube(StressProblem& P, int N, f64 L) {
  P.mats = &kMats;
  auto id = [&](int x, int y, int z) { return (x * N + y) * N + z; };
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y)
      for (int z = 0; z < N; ++z) {
        SNode nd;
        nd.c = V3{(x + 0.5) * L, (y + 0.5) * L, (z + 0.5) * L};
        nd.mass = 400.0;
        P.nodes.push_back(nd);
      }
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y)
      for (int z = 0; z < N; ++z) {
        const V3 c = P.nodes[size_t(id(x, y, z))].c;
        if (z == 0) P.bonds.push_back(bond(P, id(x, y, z), -1, V3{0, 0, -1}, sub(c, V3{0, 0, 0.5 * L}), L));
        if (x + 1 < N) P.bonds.push_back(bond(P, id(x, y, z), id(x + 1, y, z), V3{1, 0, 0}, add(c, V3{0.5 * L, 0, 0}), L));
        if (y + 1 < N) P.bonds.push_back(bond(P, id(x, y, z), id(x, y + 1, z), V3{0, 1, 0}, add(c, V3{0, 0.5 * L, 0}), L));
        if (z + 1 < N) P.bonds.push_back(bond(P, id(x, y, z), id(x, y, z + 1), V3{0, 0, 1}, add(c, V3{0, 0, 0.5 * L}), L));
      }
}

// test_stress.cpp:200-227: n^3 cubes on supports; z >= keep_z retired (gone) or left out
static void block(StressProblem& P, int n, f64 L, int keep_z, bool retire_top) {
  P.mats = &kMats;
  const int nz = retire_top ? n : keep_z;
  auto id = [&](int x, int y, int z) { return (x * n + y) * nz + z; };
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y)
      for (int z = 0; z < nz; ++z) {
        SNode nd;
        nd.c = V3{(x + 0.5) * L, (y + 0.5) * L, (z + 0.5) * L};
        nd.mass = 400.0;
        nd.gone = z >= keep_z;
        P.nodes.push_back(nd);
      }
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y)
      for (int z = 0; z < nz; ++z) {
        const V3 c = P.nodes[size_t(id(x, y, z))].c;
        if (z == 0) P.bonds.push_back(bond(P, id(x, y, z), -1, V3{0, 0, -1}, sub(c, V3{0, 0, 0.5 * L}), L));
        if (x + 1 < n) P.bonds.push_back(bond(P, id(x, y, z), id(x + 1, y, z), V3{1, 0, 0}, add(c, V3{0.5 * L, 0, 0}), L));
        if (y + 1 < n) P.bonds.push_back(bond(P, id(x, y, z), id(x, y + 1, z), V3{0, 1, 0}, add(c, V3{0, 0.5 * L, 0}), L));
        if (z + 1 < nz) P.bonds.push_back(bond(P, id(x, y, z), id(x, y, z + 1), V3{0, 0, 1}, add(c, V3{0, 0, 0.5 * L}), L));
      }
}

int main() {
  {  // test_stress.cpp:59-87
    const int n = 12;
    const f64 L = 0.5, m = 300.0;
    StressProblem P;
    cantilever(P, n, L, m);
    StressOptions so;
    so.amg_min_nodes = 4;
    CHECK(P.assemble(so));
    std::vector<f64> u;
    const PcgResult r = P.solve(gravity(P), u, 1e-10, 500, false);
    CHECK(r.converged);
    const std::vector<BondLoad> loads = P.bond_loads(&u);
    f64 W = 0.0, M = 0.0;
    for (int i = 0; i < n; ++i) {
      W += 9.81 * m;
      M += 9.81 * m * (i + 0.5) * L;
    }
    CHECK(approx(shear(loads[0]), W, 1e-6));
    CHECK(approx(moment(loads[0]), M, 1e-6));
    CHECK(std::abs(loads[0].N) < 1e-6 * W);
    CHECK(approx(shear(loads[n - 1]), 9.81 * m, 1e-6));
    CHECK(approx(moment(loads[n - 1]), 9.81 * m * 0.5 * L, 1e-6));
    const f64 phi_root = bond_utilization(P.bonds[0], loads[0], 1.0, kMats);
    const f64 phi_tip = bond_utilization(P.bonds[size_t(n - 1)], loads[size_t(n - 1)], 1.0, kMats);
    CHECK(phi_root > 50.0 * phi_tip);
  }
  {  // test_stress.cpp:89-138
    StressProblem A, B;
    cube(A, 8, 0.4);
    cube(B, 8, 0.4);
    StressOptions sa, sb;
    sa.amg_min_nodes = 16;
    sb.amg_min_nodes = 1 << 30;
    CHECK(A.assemble(sa));
    CHECK(B.assemble(sb));
    std::vector<f64> ua, ub;
    const PcgResult ra = A.solve(gravity(A), ua, 1e-9, 2000, false);
    const PcgResult rb = B.solve(gravity(B), ub, 1e-9, 5000, false);
    CHECK(ra.converged);
    CHECK(rb.converged);
    CHECK(ra.iters < rb.iters);
    f64 err = 0.0, mag = 0.0;
    for (size_t k = 0; k < ua.size(); ++k) {
      err = std::max(err, std::abs(ua[k] - ub[k]));
      mag = std::max(mag, std::abs(ua[k]));
    }
    CHECK(err <= 1e-6 * mag);
    const std::vector<BondLoad> loads = A.bond_loads(&ua);
    f64 R = 0.0;
    for (size_t b = 0; b < A.bonds.size(); ++b)
      if (A.bonds[b].b < 0) R += loads[b].N;
    CHECK(approx(-R, 9.81 * 400.0 * 512, 1e-5));
  }
  {  // test_stress.cpp:140-152
    StressProblem P;
    cantilever(P, 9, 0.5, 100.0);
    P.bonds.erase(P.bonds.begin());
    P.nodes[4].fixed = true;
    CHECK(P.assemble());
    std::vector<f64> f(6 * 9, 0.0), u;
    f[0] = -1000.0;
    f[6 * 8] = 1000.0;
    const PcgResult r = P.solve(f, u, 1e-10, 500, false);
    CHECK(r.converged);
    const std::vector<BondLoad> loads = P.bond_loads(&u);
    bool all = true;
    for (const BondLoad& L : loads) all = all && approx(L.N, 1000.0, 1e-6);
    CHECK(all);
  }
  {  // test_stress.cpp:154-174
    StressProblem A, B;
    cantilever(A, 10, 0.5, 200.0);
    cantilever(B, 10, 0.5, 200.0);
    A.bonds.push_back(bond(A, 6, -1, V3{0, 0, -1}, sub(A.nodes[6].c, V3{0, 0, 0.25}), 0.5));
    B.bonds.push_back(bond(B, 6, -1, V3{0, 0, -1}, sub(B.nodes[6].c, V3{0, 0, 0.25}), 0.5));
    StressOptions so;
    so.amg_min_nodes = 1 << 30;
    CHECK(A.assemble(so));
    A.remove_bond(3);
    B.bonds[3].broken = true;
    CHECK(B.assemble(so));
    std::vector<f64> ua, ub;
    A.solve(gravity(A), ua, 1e-11, 2000, false);
    B.solve(gravity(B), ub, 1e-11, 2000, false);
    f64 err = 0.0, mag = 0.0;
    for (size_t k = 0; k < ua.size(); ++k) {
      err = std::max(err, std::abs(ua[k] - ub[k]));
      mag = std::max(mag, std::abs(ub[k]));
    }
    CHECK(err <= 1e-6 * mag);
  }
  {  // test_stress.cpp:176-198
    StressProblem P;
    cantilever(P, 2, 0.5, 1.0);
    const SBond& b = P.bonds[1];
    const BondStrength S = bond_strength(b, 1.0, kMats);
    BondLoad L{};
    FailMode mode;
    L.N = 1.01 * S.ft * b.area;
    CHECK(approx(bond_utilization(b, L, 1.0, kMats, &mode), 1.01, 1e-12) && mode == FailMode::Tension);
    L = BondLoad{};
    L.N = -1.2 * S.fc * b.area;
    CHECK(approx(bond_utilization(b, L, 1.0, kMats, &mode), 1.2, 1e-12) && mode == FailMode::Crush);
    L = BondLoad{};
    L.V1 = S.coh * b.area;
    CHECK(approx(bond_utilization(b, L, 1.0, kMats, &mode), 1.5, 1e-12) && mode == FailMode::Shear);
    L.N = -b.area * S.coh / S.mu;
    CHECK(approx(bond_utilization(b, L, 1.0, kMats), 0.75, 1e-12));
    CHECK(approx(bond_utilization(b, L, 2.0, kMats), 1.0, 1e-12));
  }
  {  // test_stress.cpp:238-276
    StressProblem A, B;
    cantilever(A, 12, 0.5, 200.0);
    cantilever(B, 12, 0.5, 200.0);
    for (StressProblem* P : {&A, &B})
      for (int i : {3, 6, 9}) P->bonds.push_back(bond(*P, i, -1, V3{0, 0, -1}, sub(P->nodes[size_t(i)].c, V3{0, 0, 0.25}), 0.5));
    StressOptions so;
    so.amg_min_nodes = 1 << 30;
    CHECK(A.assemble(so));
    A.retire_nodes({10, 11});
    for (SBond& b : A.bonds)
      if (b.a == 10 && b.b == 11) b.broken = false;
    A.retire_nodes({1});
    for (int i : {1, 10, 11}) B.nodes[size_t(i)].gone = true;
    CHECK(B.assemble(so));
    std::vector<f64> ua, ub;
    const PcgResult ra = A.solve(gravity(A), ua, 1e-11, 2000, false);
    const PcgResult rb = B.solve(gravity(B), ub, 1e-11, 2000, false);
    CHECK(ra.converged);
    CHECK(!ra.breakdown);
    CHECK(rb.converged);
    f64 err = 0.0, mag = 0.0;
    for (size_t i = 0; i < A.nodes.size(); ++i) {
      if (A.nodes[i].gone) continue;
      for (int q = 0; q < 6; ++q) {
        err = std::max(err, std::abs(ua[6 * i + size_t(q)] - ub[6 * i + size_t(q)]));
        mag = std::max(mag, std::abs(ub[6 * i + size_t(q)]));
      }
    }
    CHECK(err <= 1e-6 * mag);
    bool inner = false;
    for (const SBond& b : A.bonds) inner = inner || (b.a == 10 && b.b == 11 && !b.broken);
    CHECK(inner);
    CHECK(A.assemble(so));
    bool intact = true;
    for (const SBond& b : A.bonds)
      if (b.a == 10 && b.b == 11) intact = intact && !b.broken;
    CHECK(intact);
  }
  {  // test_stress.cpp:278-291
    StressOptions so;
    so.amg_min_nodes = 16;
    StressProblem with, without;
    block(with, 14, 0.4, 9, true);
    block(without, 14, 0.4, 9, false);
    CHECK(with.assemble(so));
    CHECK(without.assemble(so));
    std::vector<f64> uw, uo;
    const PcgResult rw = with.solve(gravity(with), uw, 1e-8, 500, false);
    const PcgResult ro = without.solve(gravity(without), uo, 1e-8, 500, false);
    std::printf("pcg iterations: %d without the retired rows, %d with them\n", ro.iters, rw.iters);
    CHECK(rw.converged);
    CHECK(ro.converged);
    CHECK(rw.iters <= ro.iters + 4);
  }
  {  // test_stress.cpp:293-305
    StressProblem P;
    cantilever(P, 6, 0.5, 100.0);
    CHECK(P.assemble());
    std::vector<f64> u(6 * P.nodes.size(), 1.0);
    std::vector<f64> f(6 * P.nodes.size(), 0.0);
    const PcgResult r0 = P.solve(f, u, 1e-9, 100, true);
    CHECK(r0.converged);
    bool zero = true;
    for (f64 v : u) zero = zero && v == 0.0;
    CHECK(zero);
    f[2] = NAN;
    const PcgResult r1 = P.solve(f, u, 1e-9, 100, false);
    CHECK(!r1.converged);
    CHECK(r1.breakdown);
  }
  std::printf("%d failed\n", g_fail);
  return g_fail;
}
