// Synthetic fragment graphs for the stress solve's workloads (libsvxgen.so, CPU only).
//
// Deterministic on every machine: a splitmix64 integer generator decides everything, and every
// value is made by correctly rounded IEEE operations (+ - * / sqrt) only - no libm function
// (exp, log, sin) decides a topology or a value, so a fixture written on one machine matches the
// graph generated on another (the reference keeps the same rule with its own dmath,
// tests/core/test_dmath.cpp:69-84).
//
// Conventions follow the reference's test fixtures (tests/core/test_stress.cpp:14-52): a bond
// through a square section of side a has area a^2, s1 = s2 = a^4/12, c1 = c2 = a/2,
// rmax = a sqrt(1/2), t1 = y if |n.x| > 0.5 else x, t2 = n x t1, la / lb the distances of the
// section from the node centres along n (a support's lb = la); gravity f_z = -9.81 m.
// Node mass = material density x L^3. Material properties: the reference's preset table
// (core/src/material/material.cpp:71-98).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/svxg.h"

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
  int below(int n) { return static_cast<int>(next() % static_cast<uint64_t>(n)); }
};

// e^x from basic operations only: x = k ln2 + r, |r| <= ln2/2, a 24-term Taylor sum, scaled by 2^k
double det_exp(double x) {
  const double ln2 = 0.6931471805599453;
  double t = x / ln2;
  long k = static_cast<long>(t >= 0 ? t + 0.5 : t - 0.5);
  const double r = x - static_cast<double>(k) * ln2;
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 24; ++i) {
    term = term * r / static_cast<double>(i);
    sum += term;
  }
  while (k > 0) {
    sum *= 2.0;
    --k;
  }
  while (k < 0) {
    sum *= 0.5;
    ++k;
  }
  return sum;
}

// approximately standard normal: the sum of twelve uniforms minus six (Irwin-Hall)
double normalish(Rng& g) {
  double s = 0.0;
  for (int i = 0; i < 12; ++i) s += g.uniform();
  return s - 6.0;
}

// The reference's material presets, by id (material.cpp:72-98): E, G, rho, ft, fb, fc, cohesion,
// friction.
const svxg_material kPresets[12] = {
    {30e9, 12.5e9, 2400, 1.2e6, 3.5e6, 30e6, 1.6e6, 0.7},      // rc
    {28e9, 11.5e9, 2300, 0.8e6, 1.2e6, 25e6, 1.2e6, 0.7},      // concrete
    {200e9, 80e9, 7850, 180e6, 180e6, 200e6, 100e6, 0.35},     // steel
    {10e9, 4e9, 1900, 0.15e6, 0.3e6, 8e6, 0.35e6, 0.75},       // masonry
    {0.1e9, 0.04e9, 1700, 0.01e6, 0.01e6, 0.4e6, 0.03e6, 0.55},// soil
    {40e9, 16e9, 2600, 2.5e6, 3.0e6, 80e6, 4e6, 0.8},          // rock
    {60e9, 25e9, 2800, 10e6, 10e6, 200e6, 15e6, 0.8},          // bedrock
    {11e9, 0.7e9, 550, 1.2e6, 8e6, 20e6, 2e6, 0.5},            // wood
    {30e9, 12e9, 2600, 0.25e6, 0.4e6, 40e6, 0.5e6, 0.7},       // stone
    {70e9, 29e9, 2500, 3e6, 4e6, 100e6, 3e6, 0.4},             // glass
    {34e9, 13e9, 2550, 9e6, 10e6, 35e6, 4e6, 0.7},             // rebar
    {26e9, 10e9, 980, 44e6, 44e6, 40e6, 25e6, 0.35},           // steel_section
};

// structural materials a workload draws from: rc (30 GPa), steel (200), masonry (10), wood (11),
// glass (70): stiffness ratio up to 20:1
const uint8_t kStructural[5] = {0, 2, 3, 7, 9};

struct Graph {
  std::vector<svxg_node> nodes;
  std::vector<svxg_bond> bonds;
  std::vector<double> f;
  std::vector<int32_t> graph_of;  // (batches: the graph of each node; unused for one graph)
};

svxg_bond base_bond(int32_t a, int32_t b, const double n[3], const double p[3], double side, uint8_t ma, uint8_t mb,
                    const double ca[3], const double* cb) {
  svxg_bond s;
  std::memset(&s, 0, sizeof s);
  s.a = a;
  s.b = b;
  s.ma = ma;
  s.mb = mb;
  for (int q = 0; q < 3; ++q) {
    s.n[q] = n[q];
    s.p[q] = p[q];
  }
  const bool xn = n[0] > 0.5 || n[0] < -0.5;
  s.t1[0] = xn ? 0.0 : 1.0;
  s.t1[1] = xn ? 1.0 : 0.0;
  s.t1[2] = 0.0;
  s.t2[0] = s.n[1] * s.t1[2] - s.n[2] * s.t1[1];
  s.t2[1] = s.n[2] * s.t1[0] - s.n[0] * s.t1[2];
  s.t2[2] = s.n[0] * s.t1[1] - s.n[1] * s.t1[0];
  s.area = side * side;
  s.s1 = s.s2 = side * side * side * side / 12.0;
  s.c1 = s.c2 = 0.5 * side;
  s.rmax = 0.7071067811865476 * side;
  auto adot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  const double pa[3] = {p[0] - ca[0], p[1] - ca[1], p[2] - ca[2]};
  const double la = adot(pa, n);
  s.la = la < 0 ? -la : la;
  if (cb) {
    const double bp[3] = {cb[0] - p[0], cb[1] - p[1], cb[2] - p[2]};
    const double lb = adot(bp, n);
    s.lb = lb < 0 ? -lb : lb;
  } else {
    s.lb = s.la;
  }
  s.strength = 1.0;
  s.faces = 16;
  s.tag = -1;
  return s;
}

// A lattice of cells (cell -> node index, -1: empty) with face bonds, supports under z == zmin
// cells (when supported) and section factors U[0.5, 1]. Node order: the cells' order in `cells`.
void lattice(Graph& G, const std::vector<std::array<int, 3>>& cells, const std::vector<uint8_t>& mat, double L,
             bool supported, int zmin, Rng& g) {
  std::unordered_map<int64_t, int32_t> at;
  auto key = [](int x, int y, int z) {
    return (static_cast<int64_t>(x + 1048576) << 42) | (static_cast<int64_t>(y + 1048576) << 21) | static_cast<int64_t>(z + 1048576);
  };
  const int32_t base = static_cast<int32_t>(G.nodes.size());
  at.reserve(cells.size() * 2);
  for (size_t i = 0; i < cells.size(); ++i) at[key(cells[i][0], cells[i][1], cells[i][2])] = base + static_cast<int32_t>(i);
  for (size_t i = 0; i < cells.size(); ++i) {
    svxg_node nd;
    std::memset(&nd, 0, sizeof nd);
    nd.c[0] = (cells[i][0] + 0.5) * L;
    nd.c[1] = (cells[i][1] + 0.5) * L;
    nd.c[2] = (cells[i][2] + 0.5) * L;
    nd.mass = kPresets[mat[i]].rho * (L * L * L);
    G.nodes.push_back(nd);
  }
  for (size_t i = 0; i < cells.size(); ++i) {
    const int x = cells[i][0], y = cells[i][1], z = cells[i][2];
    const int32_t a = base + static_cast<int32_t>(i);
    const double* ca = G.nodes[size_t(a)].c;
    if (supported && z == zmin) {
      const double n[3] = {0, 0, -1};
      const double p[3] = {ca[0], ca[1], ca[2] - 0.5 * L};
      const double side = std::sqrt((0.5 + 0.5 * g.uniform()) * (L * L));
      G.bonds.push_back(base_bond(a, -1, n, p, side, mat[i], mat[i], ca, nullptr));
    }
    const int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int k = 0; k < 3; ++k) {
      auto it = at.find(key(x + d[k][0], y + d[k][1], z + d[k][2]));
      if (it == at.end()) continue;
      const int32_t b = it->second;
      const double n[3] = {double(d[k][0]), double(d[k][1]), double(d[k][2])};
      const double p[3] = {ca[0] + 0.5 * L * n[0], ca[1] + 0.5 * L * n[1], ca[2] + 0.5 * L * n[2]};
      const double side = std::sqrt((0.5 + 0.5 * g.uniform()) * (L * L));
      G.bonds.push_back(base_bond(a, b, n, p, side, mat[i], mat[size_t(b - base)], ca, G.nodes[size_t(b)].c));
    }
  }
}

void gravity(Graph& G) {
  G.f.assign(6 * G.nodes.size(), 0.0);
  for (size_t i = 0; i < G.nodes.size(); ++i) G.f[6 * i + 2] = -9.81 * G.nodes[i].mass;
}

// A connected random lattice subset of `count` cells grown from the origin (a falling piece or a
// small building), x-major ordered.
std::vector<std::array<int, 3>> grow(int count, Rng& g) {
  std::unordered_set<int64_t> in;
  std::vector<std::array<int, 3>> cells, front;
  auto key = [](int x, int y, int z) {
    return (static_cast<int64_t>(x + 1048576) << 42) | (static_cast<int64_t>(y + 1048576) << 21) | static_cast<int64_t>(z + 1048576);
  };
  front.push_back({0, 0, 0});
  while (static_cast<int>(cells.size()) < count && !front.empty()) {
    const int pick = g.below(static_cast<int>(front.size()));
    const std::array<int, 3> c = front[size_t(pick)];
    front[size_t(pick)] = front.back();
    front.pop_back();
    if (!in.insert(key(c[0], c[1], c[2])).second) continue;
    cells.push_back(c);
    const int d[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (const auto& e : d) {
      // (taller than wide: vertical growth twice as likely, as buildings and columns are)
      const std::array<int, 3> nb{c[0] + e[0], c[1] + e[1], c[2] + e[2]};
      if (in.count(key(nb[0], nb[1], nb[2]))) continue;
      front.push_back(nb);
      if (e[2] != 0) front.push_back(nb);
    }
  }
  std::sort(cells.begin(), cells.end());
  return cells;
}

}  // namespace

extern "C" {

typedef struct svxgen_graph svxgen_graph;

// The reference's 12 material presets; returns the count.
int32_t svxgen_materials(svxg_material* out) {
  if (out) std::memcpy(out, kPresets, sizeof kPresets);
  return 12;
}

// C1: a tower of nx x ny x nz cubes of side L on supports, one structural material per storey.
svxgen_graph* svxgen_tower(int32_t nx, int32_t ny, int32_t nz, double L, uint64_t seed) {
  Rng g{seed * 0x2545F4914F6CDD1Dull + 1};
  std::vector<uint8_t> storey(static_cast<size_t>(nz));
  for (int z = 0; z < nz; ++z) storey[size_t(z)] = kStructural[g.below(5)];
  std::vector<std::array<int, 3>> cells;
  std::vector<uint8_t> mat;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        cells.push_back({x, y, z});
        mat.push_back(storey[size_t(z)]);
      }
  Graph* G = new Graph;
  lattice(*G, cells, mat, L, true, 0, g);
  gravity(*G);
  return reinterpret_cast<svxgen_graph*>(G);
}

// C2: an nx x ny x nz block with a fraction `voids` of its cells missing (seeded), only the cells
// connected to the ground kept, materials drawn per (x, y) column.
svxgen_graph* svxgen_block(int32_t nx, int32_t ny, int32_t nz, double L, double voids, uint64_t seed) {
  Rng g{seed * 0x2545F4914F6CDD1Dull + 7};
  const size_t N = size_t(nx) * size_t(ny) * size_t(nz);
  auto id = [&](int x, int y, int z) { return (size_t(x) * size_t(ny) + size_t(y)) * size_t(nz) + size_t(z); };
  std::vector<uint8_t> present(N, 0);
  for (size_t k = 0; k < N; ++k) present[k] = g.uniform() >= voids ? 1 : 0;
  std::vector<uint8_t> column(size_t(nx) * size_t(ny));
  for (auto& c : column) c = kStructural[g.below(5)];
  // keep what reaches the ground (breadth-first from the z = 0 cells)
  std::vector<uint8_t> keep(N, 0);
  std::vector<size_t> queue;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      if (present[id(x, y, 0)]) {
        keep[id(x, y, 0)] = 1;
        queue.push_back(id(x, y, 0));
      }
  for (size_t h = 0; h < queue.size(); ++h) {
    const size_t k = queue[h];
    const int z = int(k % size_t(nz)), y = int((k / size_t(nz)) % size_t(ny)), x = int(k / (size_t(nz) * size_t(ny)));
    const int d[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (const auto& e : d) {
      const int a = x + e[0], b = y + e[1], c = z + e[2];
      if (a < 0 || b < 0 || c < 0 || a >= nx || b >= ny || c >= nz) continue;
      const size_t j = id(a, b, c);
      if (present[j] && !keep[j]) {
        keep[j] = 1;
        queue.push_back(j);
      }
    }
  }
  std::vector<std::array<int, 3>> cells;
  std::vector<uint8_t> mat;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z)
        if (keep[id(x, y, z)]) {
          cells.push_back({x, y, z});
          mat.push_back(column[size_t(x) * size_t(ny) + size_t(y)]);
        }
  Graph* G = new Graph;
  lattice(*G, cells, mat, L, true, 0, g);
  gravity(*G);
  return reinterpret_cast<svxgen_graph*>(G);
}

// A piece of `count` cells: a connected random lattice subset, either standing on supports (its
// lowest cells) or a free body pinned at the node nearest its centre of mass (the reference's
// pin_reference, world_pieces.cpp:313-326). One structural material.
svxgen_graph* svxgen_piece(int32_t count, int32_t pinned, double L, uint64_t seed) {
  Rng g{seed * 0x2545F4914F6CDD1Dull + 13};
  const uint8_t m = kStructural[g.below(5)];
  std::vector<std::array<int, 3>> cells = grow(std::max(1, count), g);
  int zmin = 1 << 30;
  for (const auto& c : cells) zmin = std::min(zmin, c[2]);
  std::vector<uint8_t> mat(cells.size(), m);
  Graph* G = new Graph;
  lattice(*G, cells, mat, L, !pinned, zmin, g);
  if (pinned) {
    double s[3] = {0, 0, 0}, M = 0.0;
    for (const auto& nd : G->nodes) {
      for (int q = 0; q < 3; ++q) s[q] += nd.mass * nd.c[q];
      M += nd.mass;
    }
    for (double& v : s) v /= M;
    size_t best = 0;
    double bd = -1.0;
    for (size_t i = 0; i < G->nodes.size(); ++i) {
      const double* c = G->nodes[i].c;
      const double d = (c[0] - s[0]) * (c[0] - s[0]) + (c[1] - s[1]) * (c[1] - s[1]) + (c[2] - s[2]) * (c[2] - s[2]);
      if (bd < 0 || d < bd) {
        bd = d;
        best = i;
      }
    }
    G->nodes[best].fixed = 1;
  }
  gravity(*G);
  return reinterpret_cast<svxgen_graph*>(G);
}

// Node counts of a batch of `count` pieces: clamp(round(median e^{sigma z}), lo, hi), z
// approximately standard normal.
void svxgen_piece_sizes(int32_t count, double median, double sigma, int32_t lo, int32_t hi, uint64_t seed, int32_t* out) {
  Rng g{seed * 0x2545F4914F6CDD1Dull + 29};
  for (int32_t k = 0; k < count; ++k) {
    const double v = median * det_exp(sigma * normalish(g));
    long r = static_cast<long>(v + 0.5);
    out[k] = static_cast<int32_t>(std::min<long>(hi, std::max<long>(lo, r)));
  }
}

// Up to `count` intact bonds whose removal, all together, leaves every free node of the graph
// connected to a support bond or a pinned node (no mechanism): a random spanning forest of the
// intact bonds is grown from the supported and pinned nodes, and the cuts are drawn from the intact
// bonds outside it. broken[k] != 0 marks bond k already broken. Returns the number written to out.
int32_t svxgen_safe_cuts(int32_t n, const svxg_node* nodes, int32_t m, const svxg_bond* bonds, const uint8_t* broken,
                         int32_t count, uint64_t seed, int32_t* out) {
  Rng g{seed * 0x2545F4914F6CDD1Dull + 41};
  std::vector<int32_t> deg(static_cast<size_t>(n) + 1, 0);
  for (int32_t k = 0; k < m; ++k)
    if (!broken[k] && bonds[k].b >= 0) {
      ++deg[size_t(bonds[k].a) + 1];
      ++deg[size_t(bonds[k].b) + 1];
    }
  for (int32_t i = 0; i < n; ++i) deg[size_t(i) + 1] += deg[size_t(i)];
  std::vector<int32_t> adj(static_cast<size_t>(deg[size_t(n)])), put(deg.begin(), deg.end() - 1);
  for (int32_t k = 0; k < m; ++k)
    if (!broken[k] && bonds[k].b >= 0) {
      adj[size_t(put[size_t(bonds[k].a)]++)] = k;
      adj[size_t(put[size_t(bonds[k].b)]++)] = k;
    }
  std::vector<uint8_t> tree(static_cast<size_t>(m), 0), seen(static_cast<size_t>(n), 0);
  std::vector<int32_t> front;
  for (int32_t i = 0; i < n; ++i)
    if (nodes[i].fixed) {
      seen[size_t(i)] = 1;
      front.push_back(i);
    }
  for (int32_t k = 0; k < m; ++k)
    if (!broken[k] && bonds[k].b < 0 && !seen[size_t(bonds[k].a)]) {
      seen[size_t(bonds[k].a)] = 1;
      tree[size_t(k)] = 1;
      front.push_back(bonds[k].a);
    }
  while (!front.empty()) {
    const size_t pick = size_t(g.below(static_cast<int>(front.size())));
    const int32_t i = front[pick];
    front[pick] = front.back();
    front.pop_back();
    for (int32_t e = deg[size_t(i)]; e < deg[size_t(i) + 1]; ++e) {
      const int32_t k = adj[size_t(e)];
      const int32_t j = bonds[k].a == i ? bonds[k].b : bonds[k].a;
      if (seen[size_t(j)]) continue;
      seen[size_t(j)] = 1;
      tree[size_t(k)] = 1;
      front.push_back(j);
    }
  }
  std::vector<int32_t> cand;
  for (int32_t k = 0; k < m; ++k)
    if (!broken[k] && !tree[size_t(k)]) cand.push_back(k);
  const int32_t c = std::min<int32_t>(count, int32_t(cand.size()));
  for (int32_t t = 0; t < c; ++t) {
    const size_t j = size_t(t) + size_t(g.below(static_cast<int>(cand.size()) - t));
    std::swap(cand[size_t(t)], cand[j]);
    out[t] = cand[size_t(t)];
  }
  return c;
}

// A building of nx x ny x nz cells on supports (the city chunk), one material a storey.
svxgen_graph* svxgen_building(int32_t nx, int32_t ny, int32_t nz, double L, uint64_t seed) {
  return svxgen_tower(nx, ny, nz, L, seed);
}

void svxgen_counts(const svxgen_graph* h, int32_t* n, int32_t* m) {
  const Graph* G = reinterpret_cast<const Graph*>(h);
  *n = static_cast<int32_t>(G->nodes.size());
  *m = static_cast<int32_t>(G->bonds.size());
}

void svxgen_copy(const svxgen_graph* h, svxg_node* nodes, svxg_bond* bonds, double* f) {
  const Graph* G = reinterpret_cast<const Graph*>(h);
  if (nodes) std::memcpy(nodes, G->nodes.data(), G->nodes.size() * sizeof(svxg_node));
  if (bonds) std::memcpy(bonds, G->bonds.data(), G->bonds.size() * sizeof(svxg_bond));
  if (f) std::memcpy(f, G->f.data(), G->f.size() * sizeof(double));
}

void svxgen_free(svxgen_graph* h) { delete reinterpret_cast<Graph*>(h); }

// Uniform doubles in [0, 1) from the generator (tests: reproducible random vectors).
void svxgen_uniform(uint64_t seed, int64_t count, double* out) {
  Rng g{seed * 0x2545F4914F6CDD1Dull + 41};
  for (int64_t k = 0; k < count; ++k) out[k] = g.uniform();
}

}  // extern "C"
