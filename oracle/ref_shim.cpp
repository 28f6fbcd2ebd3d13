// ref_shim — the C entry points of oracle.h (prefixed ref_) over the reference's own
// svx::StressProblem, compiled with the reference's sources where they lie (oracle/Makefile).
//
// TEST INFRASTRUCTURE ONLY: it pins the oracle's restatement bit for bit, writes the golden
// fixtures under tests/golden/, and is the "reference" CPU baseline of bench.py. The product path
// never loads it.
#include <cmath>
#include <vector>

#include "oracle.h"
#include "svx/base/parallel.hpp"
#include "svx/stress/stress.hpp"

using namespace svx;

struct ref_problem {
  StressProblem P;
  MaterialTable table;
  StressOptions opt;
};

namespace {

SBond to_ref(const svxg_bond& s) {
  SBond b;
  b.a = s.a;
  b.b = s.b;
  b.broken = s.broken != 0;
  b.in_k = s.in_k != 0;
  b.ma = static_cast<MaterialId>(s.ma);
  b.mb = static_cast<MaterialId>(s.mb);
  b.n = V3{s.n[0], s.n[1], s.n[2]};
  b.t1 = V3{s.t1[0], s.t1[1], s.t1[2]};
  b.t2 = V3{s.t2[0], s.t2[1], s.t2[2]};
  b.area = s.area;
  b.p = V3{s.p[0], s.p[1], s.p[2]};
  b.s1 = s.s1;
  b.s2 = s.s2;
  b.c1 = s.c1;
  b.c2 = s.c2;
  b.rmax = s.rmax;
  b.la = s.la;
  b.lb = s.lb;
  b.strength = s.strength;
  b.sectioned = s.sectioned != 0;
  b.ft = s.ft;
  b.fb = s.fb;
  b.fc = s.fc;
  b.coh = s.coh;
  b.mu = s.mu;
  b.faces = s.faces;
  b.tag = s.tag;
  return b;
}

void from_ref(const SBond& b, svxg_bond* s) {
  s->broken = b.broken ? 1 : 0;
  s->in_k = b.in_k ? 1 : 0;
}

StressOptions to_ref(const svxg_stress_options& o) {
  StressOptions so;
  so.rtol = o.rtol;
  so.amg_min_nodes = o.amg_min_nodes;
  so.amg.coarse_max = o.amg.coarse_max;
  so.amg.strength = o.amg.strength;
  so.amg.smoothed = o.amg.smoothed != 0;
  so.amg.coarse_scale = o.amg.coarse_scale;
  so.amg.max_levels = o.amg.max_levels;
  so.amg.sweeps = o.amg.sweeps;
  so.amg.coarsen_dense = o.amg.coarsen_dense != 0;
  so.amg.reaggregate = o.amg.reaggregate != 0;
  return so;
}

void fill_table(MaterialTable& t, const svxg_material* mats, int32_t nmats) {
  for (int32_t i = 0; i < nmats; ++i) {
    Material m;
    m.name = "m" + std::to_string(i);
    m.E = mats[i].E;
    m.G = mats[i].G;
    m.rho = mats[i].rho;
    m.ft = mats[i].ft;
    m.fb = mats[i].fb;
    m.fc = mats[i].fc;
    m.cohesion = mats[i].cohesion;
    m.friction = mats[i].friction;
    t.set(static_cast<MaterialId>(i), m);
  }
}

void to_c(const PcgResult& r, svxg_pcg_result* out) {
  out->iters = r.iters;
  out->converged = r.converged ? 1 : 0;
  out->breakdown = r.breakdown ? 1 : 0;
  out->refinements = 0;
  out->rel_res = r.rel_res;
}

}  // namespace

extern "C" {

ref_problem* ref_create(const svxg_node* nodes, int32_t n, const svxg_bond* bonds, int32_t m, const svxg_material* mats,
                        int32_t nmats) {
  ref_problem* h = new ref_problem;
  fill_table(h->table, mats, nmats);
  h->P.mats = &h->table;
  h->P.nodes.resize(size_t(n));
  for (int32_t i = 0; i < n; ++i) {
    SNode& s = h->P.nodes[size_t(i)];
    s.c = V3{nodes[i].c[0], nodes[i].c[1], nodes[i].c[2]};
    s.mass = nodes[i].mass;
    s.fixed = nodes[i].fixed != 0;
    s.gone = nodes[i].gone != 0;
  }
  h->P.bonds.reserve(size_t(m));
  for (int32_t k = 0; k < m; ++k) h->P.bonds.push_back(to_ref(bonds[k]));
  return h;
}

void ref_destroy(ref_problem* h) { delete h; }

int ref_assemble(ref_problem* h, const svxg_stress_options* opt) {
  h->opt = to_ref(*opt);
  return h->P.assemble(h->opt) ? 1 : 0;
}

void ref_release(ref_problem* h) { h->P.release(); }

int ref_solve(ref_problem* h, const double* f, double* u, double rtol, int32_t maxit, int32_t warm, svxg_pcg_result* res) {
  const size_t n6 = 6 * h->P.nodes.size();
  std::vector<f64> fv(f, f + n6), uv(u, u + n6);
  const PcgResult r = h->P.solve(fv, uv, rtol, maxit, warm != 0);
  to_c(r, res);
  for (size_t k = 0; k < n6 && k < uv.size(); ++k) u[k] = uv[k];
  return 1;
}

void ref_begin(ref_problem* h, const double* f, const double* u0) {
  const size_t n6 = 6 * h->P.nodes.size();
  std::vector<f64> fv(f, f + n6), uv;
  if (u0) uv.assign(u0, u0 + n6);
  h->P.begin(fv, uv);
}

void ref_iterate(ref_problem* h, int32_t maxit, double rtol, svxg_pcg_result* res) { to_c(h->P.iterate(maxit, rtol), res); }

void ref_current(ref_problem* h, double* u) {
  std::vector<f64> uv;
  h->P.current(uv);
  for (size_t k = 0; k < uv.size(); ++k) u[k] = uv[k];
}

void ref_remove_bond(ref_problem* h, int32_t b) { h->P.remove_bond(b); }

void ref_retire_nodes(ref_problem* h, const int32_t* ids, int32_t count) {
  h->P.retire_nodes(std::vector<i32>(ids, ids + count));
}

void ref_set_bond_broken(ref_problem* h, int32_t b, int32_t broken) { h->P.bonds[size_t(b)].broken = broken != 0; }

void ref_extend_warm_start(const svxg_node* nodes, int32_t n, const svxg_bond* bonds, int32_t m, double* u,
                           int32_t first_new) {
  ref_problem* h = ref_create(nodes, n, bonds, m, nullptr, 0);
  std::vector<f64> uv(u, u + 6 * size_t(n));
  h->P.extend_warm_start(uv, first_new);
  std::copy(uv.begin(), uv.end(), u);
  ref_destroy(h);
}

void ref_bond_load(ref_problem* h, int32_t b, const double* u, svxg_bond_load* out) {
  const std::vector<f64> uv(u, u + 6 * h->P.nodes.size());
  const BondLoad L = h->P.bond_load(b, uv);
  *out = svxg_bond_load{L.N, L.V1, L.V2, L.T, L.M1, L.M2};
}

void ref_bond_strength(const svxg_bond* b, double fragility, const svxg_material* mats, int32_t nmats, double out[5]) {
  MaterialTable t;
  fill_table(t, mats, nmats);
  const BondStrength S = bond_strength(to_ref(*b), fragility, t);
  out[0] = S.ft;
  out[1] = S.fb;
  out[2] = S.fc;
  out[3] = S.coh;
  out[4] = S.mu;
}

double ref_bond_utilization(const svxg_bond* b, const svxg_bond_load* L, double fragility, const svxg_material* mats,
                            int32_t nmats, uint8_t* mode) {
  MaterialTable t;
  fill_table(t, mats, nmats);
  BondLoad l;
  l.N = L->N;
  l.V1 = L->V1;
  l.V2 = L->V2;
  l.T = L->T;
  l.M1 = L->M1;
  l.M2 = L->M2;
  FailMode fm = FailMode::None;
  const f64 phi = bond_utilization(to_ref(*b), l, fragility, t, &fm);
  if (mode) *mode = static_cast<uint8_t>(fm);
  return phi;
}

// every bond's loads and checks from one u (bond_load + bond_utilization of each bond in turn)
void ref_bond_checks(ref_problem* h, const double* u, double fragility, svxg_bond_load* loads, double* phi,
                     uint8_t* mode) {
  const std::vector<f64> uv(u, u + 6 * h->P.nodes.size());
  for (size_t b = 0; b < h->P.bonds.size(); ++b) {
    const BondLoad L = h->P.bond_load(static_cast<i32>(b), uv);
    if (loads) loads[b] = svxg_bond_load{L.N, L.V1, L.V2, L.T, L.M1, L.M2};
    if (phi) {
      FailMode fm = FailMode::None;
      phi[b] = bond_utilization(h->P.bonds[b], L, fragility, h->table, &fm);
      if (mode) mode[b] = static_cast<uint8_t>(fm);
    }
  }
}

void ref_info(ref_problem* h, svxg_problem_info* out) {
  *out = svxg_problem_info{};
  out->nodes = static_cast<int32_t>(h->P.nodes.size());
  out->bonds = static_cast<int32_t>(h->P.bonds.size());
  out->nfree = h->P.free_nodes();
  out->block_jacobi = h->P.block_jacobi();
  out->assembled = h->P.assembled();
  out->preconditioner_current = h->P.preconditioner_current();
  out->blocks = h->P.matrix_blocks();
  out->work_per_iteration = h->P.work_per_iteration();
  out->assembly_work = h->P.assembly_work();
}

void ref_get_bonds(ref_problem* h, svxg_bond* out) {
  for (size_t k = 0; k < h->P.bonds.size(); ++k) from_ref(h->P.bonds[k], &out[k]);
}

int32_t ref_inv6(const double* A, double* out) { return blk6::inv6(A, out) ? 1 : 0; }

void ref_set_threads(int32_t n) { set_num_threads(n); }
int32_t ref_threads(void) { return num_threads(); }

}  // extern "C"
