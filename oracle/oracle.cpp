// oracle — CPU restatement of the reference's fragment-graph stress solve.
//
// TEST INFRASTRUCTURE ONLY: the checker of the device library (tests/, __graft_entry__.smoke(),
// bench.py's cpu_baseline leg). Never loaded by the product path.
//
// It restates the reference's algorithm step for step, in the same floating-point operation order,
// so that on the same input it produces the same bits as the reference compiled with
// -ffp-contract=off -fno-fast-math (CMakeLists.txt:16-17). Each part cites what it follows:
//   block helpers            core/src/solve/amg.cpp:32-172 (inv6, fro6, rigid_block, atbd_add,
//                            mul6 / mul6_add row sums, the FP32 bsub / badd lane sums)
//   Bsr6 / Bsr6Builder       amg.cpp:174-224, amg.hpp:18-44
//   Amg::build / coarsen     amg.cpp:228-533 (finalize, estimate_lmax, aggregation, P, RAP)
//   coarse Cholesky          amg.cpp:535-592
//   FP32 copies and V-cycle  amg.cpp:594-731 (make_fast, smooth_f, cycle_f, apply)
//   StressProblem            core/src/stress/stress.cpp:59-475 (bond_matrices, build_matrix,
//                            assemble, remove_bond, retire_nodes, precondition, begin, iterate)
//   bond_load / checks       stress.cpp:14-57, 477-502
//   ordered chunk sums       core/src/base/parallel.cpp:208-227 (dots in chunks of 4096 entries)
//
// Pinning (tests/test_oracle_*.py): every function here is compared bit for bit with the
// reference compiled from its own sources (oracle/_ref) on the reference's test problems and the
// synthetic configs, through committed fixtures written by tests/golden/make_golden.py, and
// against the closed forms of tests/core/test_stress.cpp (cantilever statics, chain tension,
// support reactions, the failure-mode known answers) and brute-force dense solves.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

namespace {

using f32 = float;
using f64 = double;
using i32 = int32_t;
using i64 = int64_t;

struct V3 {
  f64 x = 0, y = 0, z = 0;
};
inline V3 v3(const double* p) { return V3{p[0], p[1], p[2]}; }
inline V3 vsub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f64 vdot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 vcross(const V3& a, const V3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline f64 vget(const V3& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

constexpr i64 kGrain = 512;  // rows per chunk (amg.hpp:29-31)

// ---- 6x6 block arithmetic (row-major f64) --------------------------------------------------

bool inv6(const f64* A, f64* out) {
  f64 W[6][12];
  f64 big = 0.0;
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 12; ++c) {
      W[r][c] = c < 6 ? A[r * 6 + c] : (c - 6 == r ? 1.0 : 0.0);
      if (c < 6) big = std::max(big, std::abs(W[r][c]));
    }
  if (big == 0.0) return false;
  for (int c = 0; c < 6; ++c) {
    int best = c;
    for (int r = c + 1; r < 6; ++r)
      if (std::abs(W[r][c]) > std::abs(W[best][c])) best = r;
    if (std::abs(W[best][c]) <= 1e-15 * big) return false;
    if (best != c)
      for (int k = 0; k < 12; ++k) std::swap(W[c][k], W[best][k]);
    const f64 d = 1.0 / W[c][c];
    for (int k = 0; k < 12; ++k) W[c][k] *= d;
    for (int r = 0; r < 6; ++r) {
      if (r == c) continue;
      const f64 m = W[r][c];
      if (m == 0.0) continue;
      for (int k = 0; k < 12; ++k) W[r][k] -= m * W[c][k];
    }
  }
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) out[r * 6 + c] = W[r][6 + c];
  return true;
}

f64 fro6(const f64* M) {
  f64 s = 0.0;
  for (int i = 0; i < 36; ++i) s += M[i] * M[i];
  return std::sqrt(s);
}

// the map of a rigid motion (U, Theta) about X to a point at X + r: [[I, -[r]x], [0, I]]
void rigid(const V3& r, f64* P) {
  for (int i = 0; i < 36; ++i) P[i] = (i % 7 == 0) ? 1.0 : 0.0;
  P[4] = r.z;
  P[5] = -r.y;
  P[9] = -r.z;
  P[11] = r.x;
  P[15] = r.y;
  P[16] = -r.x;
}

// C += s * A^T (B D)
void atbd_add(const f64* A, const f64* B, const f64* D, f64 s, f64* C) {
  f64 T[36];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      f64 acc = 0.0;
      for (int k = 0; k < 6; ++k) acc += B[r * 6 + k] * D[k * 6 + c];
      T[r * 6 + c] = acc;
    }
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      f64 acc = 0.0;
      for (int k = 0; k < 6; ++k) acc += A[k * 6 + r] * T[k * 6 + c];
      C[r * 6 + c] += s * acc;
    }
}

// y = M x (each row summed left to right)
void mv6(const f64* M, const f64* x, f64* y) {
  for (int r = 0; r < 6; ++r) {
    const f64* m = M + 6 * r;
    y[r] = m[0] * x[0] + m[1] * x[1] + m[2] * x[2] + m[3] * x[3] + m[4] * x[4] + m[5] * x[5];
  }
}
// y += (M x): the row's sum first, then added
void mv6_add(const f64* M, const f64* x, f64* y) {
  for (int r = 0; r < 6; ++r) {
    const f64* m = M + 6 * r;
    y[r] += m[0] * x[0] + m[1] * x[1] + m[2] * x[2] + m[3] * x[3] + m[4] * x[4] + m[5] * x[5];
  }
}
// C = A B, every element summed over q from zero
void mul6(const f64* A, const f64* B, f64* C) {
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      f64 s = 0.0;
      for (int q = 0; q < 6; ++q) s += A[r * 6 + q] * B[q * 6 + c];
      C[r * 6 + c] = s;
    }
}
// C += (A B) or C += (A^T B)
void mul6_add(const f64* A, const f64* B, f64* C, bool trans) {
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      f64 s = 0.0;
      for (int q = 0; q < 6; ++q) s += (trans ? A[q * 6 + r] : A[r * 6 + q]) * B[q * 6 + c];
      C[r * 6 + c] = C[r * 6 + c] + s;
    }
}
// FP32 lane sums: y -= B x or y += B x accumulated column by column into y itself
inline void fsub(const f32* B, const f32* x, f32* y) {
  for (int j = 0; j < 6; ++j)
    for (int r = 0; r < 6; ++r) y[r] = y[r] - B[r * 6 + j] * x[j];
}
inline void fadd(const f32* B, const f32* x, f32* y) {
  for (int j = 0; j < 6; ++j)
    for (int r = 0; r < 6; ++r) y[r] = y[r] + B[r * 6 + j] * x[j];
}

// ---- block-sparse matrices -------------------------------------------------------------------

struct Bsr {
  i32 n = 0;
  std::vector<i32> rowptr, col;
  std::vector<f64> val;
  i64 blocks() const { return static_cast<i64>(col.size()); }
  void apply(const f64* x, f64* y) const {
    for (i32 i = 0; i < n; ++i) {
      f64* yi = y + 6 * size_t(i);
      for (int q = 0; q < 6; ++q) yi[q] = 0.0;
      for (i32 k = rowptr[size_t(i)]; k < rowptr[size_t(i) + 1]; ++k) mv6_add(&val[36 * size_t(k)], x + 6 * size_t(col[size_t(k)]), yi);
    }
  }
  const f64* diag(i32 i) const { return &val[36 * size_t(rowptr[size_t(i)])]; }
  f64* diag(i32 i) { return &val[36 * size_t(rowptr[size_t(i)])]; }
  f64* find(i32 r, i32 c) {
    for (i32 k = rowptr[size_t(r)]; k < rowptr[size_t(r) + 1]; ++k)
      if (col[size_t(k)] == c) return &val[36 * size_t(k)];
    return nullptr;
  }
};

// rows of (column, storage slot); blocks zero on first touch; finish(): diagonal first, then
// ascending columns (a missing diagonal is added as a zero block)
class Builder {
 public:
  explicit Builder(i32 n) : n_(n), rows_(size_t(n)) {}
  f64* at(i32 r, i32 c) {
    for (const auto& e : rows_[size_t(r)])
      if (e.first == c) return &store_[36 * size_t(e.second)];
    const i32 s = static_cast<i32>(store_.size() / 36);
    store_.resize(store_.size() + 36, 0.0);
    rows_[size_t(r)].push_back({c, s});
    return &store_[36 * size_t(s)];
  }
  Bsr finish() {
    Bsr A;
    A.n = n_;
    A.rowptr.assign(size_t(n_) + 1, 0);
    for (i32 r = 0; r < n_; ++r) {
      auto& row = rows_[size_t(r)];
      std::sort(row.begin(), row.end(), [r](const std::pair<i32, i32>& a, const std::pair<i32, i32>& b) {
        if ((a.first == r) != (b.first == r)) return a.first == r;
        return a.first < b.first;
      });
      if (row.empty() || row.front().first != r) row.insert(row.begin(), {r, -1});
      A.rowptr[size_t(r) + 1] = A.rowptr[size_t(r)] + static_cast<i32>(row.size());
    }
    const size_t nb = size_t(A.rowptr[size_t(n_)]);
    A.col.resize(nb);
    A.val.assign(36 * nb, 0.0);
    size_t k = 0;
    for (i32 r = 0; r < n_; ++r)
      for (const auto& e : rows_[size_t(r)]) {
        A.col[k] = e.first;
        if (e.second >= 0) std::memcpy(&A.val[36 * k], &store_[36 * size_t(e.second)], sizeof(f64) * 36);
        ++k;
      }
    return A;
  }

 private:
  i32 n_;
  std::vector<std::vector<std::pair<i32, i32>>> rows_;
  std::vector<f64> store_;
};

// ---- smoothed-aggregation multigrid ------------------------------------------------------------

struct AmgOpt {
  i32 coarse_max = 48;
  f64 strength = 0.1;
  bool smoothed = true;
  f64 coarse_scale = 0.6;
  int max_levels = 16;
  int sweeps = 1;
  bool coarsen_dense = true;
  bool reaggregate = true;
};

struct Level {
  Bsr A;
  std::vector<f64> Dinv;
  std::vector<V3> pos, off;
  std::vector<i32> agg;
  std::vector<i32> Prow, Pcol;
  std::vector<f64> Pval;
  std::vector<i32> Rrow;
  std::vector<std::pair<i32, i32>> Rent;
  // FP32 copies (row-major 36-float blocks; the arithmetic order is the lane order of the FP32
  // sweeps, whatever the storage)
  std::vector<f32> Af, Df, Pf, Rf;
  mutable std::vector<f32> xf, bf, rf, of, cf;
};

class Multigrid {
 public:
  bool build(const Bsr& A, const std::vector<V3>& pos, const AmgOpt& o) {
    opt_ = o;
    lv_.clear();
    chol_.clear();
    chol_n_ = 0;
    work_ = 0;
    build_work_ = 0;
    if (A.n == 0) return true;
    lv_.emplace_back();
    lv_[0].A = A;
    lv_[0].pos = pos;
    if (!finalize(lv_[0])) {
      lv_.clear();
      return false;
    }
    while (static_cast<int>(lv_.size()) < opt_.max_levels && lv_.back().A.n > opt_.coarse_max) {
      const Bsr& L = lv_.back().A;
      const bool dense = lv_.size() > 1 && L.blocks() > 80 * static_cast<i64>(L.n);
      if (dense && !opt_.coarsen_dense) break;
      if (!coarsen(lv_.size() - 1, !dense && opt_.smoothed)) break;
      if (dense) break;
    }
    factor();
    to_single();
    for (size_t l = 0; l < lv_.size(); ++l) {
      const bool last = l + 1 == lv_.size();
      const i64 nd = chol_n_ / 6;
      const i64 coarsest = chol_n_ > 0 ? std::max<i64>(lv_[l].A.blocks(), 2 * nd * nd) : 8 * 2 * lv_[l].A.blocks();
      work_ += last ? coarsest : lv_[l].A.blocks() * (4 * opt_.sweeps + 2);
    }
    return true;
  }
  bool built() const { return !lv_.empty(); }
  i64 work_per_apply() const { return work_; }
  i64 build_work() const { return build_work_; }
  size_t levels() const { return lv_.size(); }
  const Level& level(size_t l) const { return lv_[l]; }

  void apply(const f64* r, f64* z) const {
    if (lv_.empty()) return;
    const Level& L = lv_[0];
    const i32 n = L.A.n;
    for (size_t k = 0; k < 6 * size_t(n); ++k) L.bf[k] = static_cast<f32>(r[k]);
    cycle(0, L.bf.data(), L.xf.data());
    for (size_t k = 0; k < 6 * size_t(n); ++k) z[k] = L.xf[k];
  }

 private:
  bool finalize(Level& L) {
    const i32 n = L.A.n;
    L.Dinv.assign(36 * size_t(n), 0.0);
    for (i32 i = 0; i < n; ++i) {
      const f64* D = L.A.diag(i);
      if (!inv6(D, &L.Dinv[36 * size_t(i)])) {
        f64 R[36];
        std::memcpy(R, D, sizeof R);
        f64 tr = 0.0;
        for (int q = 0; q < 6; ++q) tr += std::abs(R[q * 7]);
        const f64 eps = tr > 0 ? 1e-9 * tr : 1.0;
        for (int q = 0; q < 6; ++q) R[q * 7] += eps;
        if (!inv6(R, &L.Dinv[36 * size_t(i)])) return false;
      }
    }
    build_work_ += 12 * static_cast<i64>(n);
    return true;
  }

  f64 lmax(const Level& L) {
    const i32 n = L.A.n;
    build_work_ += 15 * (L.A.blocks() + n);
    const size_t m = 6 * size_t(n);
    std::vector<f64> v(m), w(m);
    for (size_t k = 0; k < m; ++k) v[k] = 1.0 + 0.001 * static_cast<f64>(k % 7);
    f64 lam = 1.0;
    for (int it = 0; it < 15; ++it) {
      L.A.apply(v.data(), w.data());
      for (i32 i = 0; i < n; ++i) mv6(&L.Dinv[36 * size_t(i)], &w[6 * size_t(i)], &v[6 * size_t(i)]);
      f64 s = 0.0;
      for (f64 q : v) s += q * q;
      s = std::sqrt(s);
      if (!(s > 0)) break;
      lam = s;
      for (f64& q : v) q /= s;
    }
    return 1.1 * lam;
  }

  bool coarsen(size_t l, bool smoothed) {
    Level& F = lv_[l];
    const Bsr& A = F.A;
    const i32 n = A.n;
    std::vector<f64> dn(static_cast<size_t>(n));
    for (i32 i = 0; i < n; ++i) dn[size_t(i)] = fro6(A.diag(i));
    f64 theta = opt_.strength;
    auto strong = [&](i32 i, i32 k) {
      const i32 j = A.col[size_t(k)];
      return j != i && fro6(&A.val[36 * size_t(k)]) >= theta * std::sqrt(dn[size_t(i)] * dn[size_t(j)]);
    };
    auto lonely = [&](i32 i) { return A.rowptr[size_t(i) + 1] - A.rowptr[size_t(i)] <= 1; };
    std::vector<i32> agg(size_t(n), -1);
    i32 na = 0;
    for (int round = 0;; ++round) {
      std::fill(agg.begin(), agg.end(), -1);
      na = 0;
      // roots: every strong neighbour still unaggregated
      for (i32 i = 0; i < n; ++i) {
        if (agg[size_t(i)] >= 0 || lonely(i)) continue;
        bool ok = true;
        for (i32 k = A.rowptr[size_t(i)]; k < A.rowptr[size_t(i) + 1] && ok; ++k)
          if (strong(i, k) && agg[size_t(A.col[size_t(k)])] >= 0) ok = false;
        if (!ok) continue;
        agg[size_t(i)] = na;
        for (i32 k = A.rowptr[size_t(i)]; k < A.rowptr[size_t(i) + 1]; ++k)
          if (strong(i, k)) agg[size_t(A.col[size_t(k)])] = na;
        ++na;
      }
      // the rest join the aggregate of their strongest aggregated neighbour (as of the roots)
      const std::vector<i32> snap = agg;
      for (i32 i = 0; i < n; ++i) {
        if (snap[size_t(i)] >= 0) continue;
        f64 best = -1.0;
        i32 pick = -1;
        for (i32 k = A.rowptr[size_t(i)]; k < A.rowptr[size_t(i) + 1]; ++k) {
          const i32 j = A.col[size_t(k)];
          if (j == i || snap[size_t(j)] < 0) continue;
          const f64 s = fro6(&A.val[36 * size_t(k)]);
          if (s > best) {
            best = s;
            pick = snap[size_t(j)];
          }
        }
        agg[size_t(i)] = pick;
      }
      // leftovers: coupled ones alone, uncoupled ones eight to an aggregate
      i32 open = -1, fill = 0;
      for (i32 i = 0; i < n; ++i) {
        if (agg[size_t(i)] >= 0) continue;
        if (!lonely(i)) {
          agg[size_t(i)] = na++;
          continue;
        }
        if (open < 0 || fill >= 8) {
          open = na++;
          fill = 0;
        }
        agg[size_t(i)] = open;
        ++fill;
      }
      if (!opt_.reaggregate || round == 3 || 2 * static_cast<i64>(na) <= n) break;
      theta *= 0.5;
    }
    if (na >= n || static_cast<f64>(na) > 0.85 * n) return false;

    Level C;
    C.pos.assign(size_t(na), V3{});
    std::vector<i32> cnt(size_t(na), 0);
    for (i32 i = 0; i < n; ++i) {
      V3& q = C.pos[size_t(agg[size_t(i)])];
      q.x += F.pos[size_t(i)].x;
      q.y += F.pos[size_t(i)].y;
      q.z += F.pos[size_t(i)].z;
      ++cnt[size_t(agg[size_t(i)])];
    }
    for (i32 a = 0; a < na; ++a) {
      const f64 s = 1.0 / cnt[size_t(a)];
      C.pos[size_t(a)].x *= s;
      C.pos[size_t(a)].y *= s;
      C.pos[size_t(a)].z *= s;
    }
    F.agg = agg;
    F.off.resize(size_t(n));
    for (i32 i = 0; i < n; ++i) F.off[size_t(i)] = vsub(F.pos[size_t(i)], C.pos[size_t(agg[size_t(i)])]);

    // prolongation: each aggregate's rigid motion, smoothed by one damped block-Jacobi step
    using Entry = std::pair<i32, std::array<f64, 36>>;
    {
      const f64 omega = smoothed ? (4.0 / 3.0) / std::max(1e-300, lmax(F)) : 0.0;
      std::vector<std::vector<Entry>> rows(static_cast<size_t>(n));
      std::vector<i32> slot(size_t(na), -1);
      for (i32 i = 0; i < n; ++i) {
        auto& row = rows[size_t(i)];
        auto add = [&](i32 c, const f64* blk, f64 sc) {
          i32& at = slot[size_t(c)];
          if (at < 0) {
            at = static_cast<i32>(row.size());
            row.push_back({c, {}});
          }
          auto& b = row[size_t(at)].second;
          for (int q = 0; q < 36; ++q) b[size_t(q)] += sc * blk[q];
        };
        f64 Pi[36];
        rigid(F.off[size_t(i)], Pi);
        add(agg[size_t(i)], Pi, 1.0);
        const f64* Di = &F.Dinv[36 * size_t(i)];
        const i32 kend = omega != 0.0 ? A.rowptr[size_t(i) + 1] : A.rowptr[size_t(i)];
        for (i32 k = A.rowptr[size_t(i)]; k < kend; ++k) {
          const i32 j = A.col[size_t(k)];
          f64 Pj[36], T[36], U[36];
          rigid(F.off[size_t(j)], Pj);
          mul6(&A.val[36 * size_t(k)], Pj, T);
          mul6(Di, T, U);
          add(agg[size_t(j)], U, -omega);
        }
        for (const auto& e : row) slot[size_t(e.first)] = -1;
      }
      F.Prow.assign(size_t(n) + 1, 0);
      F.Pcol.clear();
      F.Pval.clear();
      for (i32 i = 0; i < n; ++i) {
        auto& row = rows[size_t(i)];
        std::sort(row.begin(), row.end(), [](const Entry& a, const Entry& b) { return a.first < b.first; });
        for (const auto& e : row) {
          F.Pcol.push_back(e.first);
          F.Pval.insert(F.Pval.end(), e.second.begin(), e.second.end());
        }
        F.Prow[size_t(i) + 1] = static_cast<i32>(F.Pcol.size());
      }
    }
    if (smoothed) build_work_ += 12 * A.blocks();
    for (i32 i = 0; i < n; ++i)
      for (i32 k = A.rowptr[size_t(i)]; k < A.rowptr[size_t(i) + 1]; ++k) {
        const i32 j = A.col[size_t(k)];
        build_work_ += 6 * static_cast<i64>(F.Prow[size_t(j) + 1] - F.Prow[size_t(j)]);
      }
    // P^T by coarse node: (fine node, P entry), fine nodes ascending
    F.Rrow.assign(size_t(na) + 1, 0);
    for (i32 J : F.Pcol) ++F.Rrow[size_t(J) + 1];
    for (i32 a = 0; a < na; ++a) F.Rrow[size_t(a) + 1] += F.Rrow[size_t(a)];
    F.Rent.assign(F.Pcol.size(), {0, 0});
    {
      std::vector<i32> put(F.Rrow.begin(), F.Rrow.end() - 1);
      for (i32 i = 0; i < n; ++i)
        for (i32 e = F.Prow[size_t(i)]; e < F.Prow[size_t(i) + 1]; ++e) F.Rent[size_t(put[size_t(F.Pcol[size_t(e)])]++)] = {i, e};
    }
    // Galerkin product: A P by fine row, then P^T (A P) by coarse row
    std::vector<std::vector<Entry>> ap(static_cast<size_t>(n));
    {
      std::vector<i32> slot(size_t(na), -1);
      for (i32 i = 0; i < n; ++i) {
        auto& row = ap[size_t(i)];
        for (i32 k = A.rowptr[size_t(i)]; k < A.rowptr[size_t(i) + 1]; ++k) {
          const i32 j = A.col[size_t(k)];
          for (i32 e = F.Prow[size_t(j)]; e < F.Prow[size_t(j) + 1]; ++e) {
            const i32 J = F.Pcol[size_t(e)];
            i32& at = slot[size_t(J)];
            if (at < 0) {
              at = static_cast<i32>(row.size());
              row.push_back({J, {}});
            }
            mul6_add(&A.val[36 * size_t(k)], &F.Pval[36 * size_t(e)], row[size_t(at)].second.data(), false);
          }
        }
        for (const auto& x : row) slot[size_t(x.first)] = -1;
      }
    }
    for (i32 i = 0; i < n; ++i) build_work_ += 6 * static_cast<i64>(F.Prow[size_t(i) + 1] - F.Prow[size_t(i)]) * static_cast<i64>(ap[size_t(i)].size());
    std::vector<std::vector<Entry>> crow(static_cast<size_t>(na));
    {
      std::vector<i32> slot(size_t(na), -1);
      for (i32 I = 0; I < na; ++I) {
        auto& out = crow[size_t(I)];
        for (i32 r = F.Rrow[size_t(I)]; r < F.Rrow[size_t(I) + 1]; ++r) {
          const i32 i = F.Rent[size_t(r)].first, e = F.Rent[size_t(r)].second;
          for (const auto& x : ap[size_t(i)]) {
            i32& at = slot[size_t(x.first)];
            if (at < 0) {
              at = static_cast<i32>(out.size());
              out.push_back({x.first, {}});
            }
            mul6_add(&F.Pval[36 * size_t(e)], x.second.data(), out[size_t(at)].second.data(), true);
          }
        }
        for (const auto& y : out) slot[size_t(y.first)] = -1;
      }
    }
    ap.clear();
    Builder B(na);
    for (i32 I = 0; I < na; ++I)
      for (const auto& y : crow[size_t(I)]) {
        f64* dst = B.at(I, y.first);
        for (int q = 0; q < 36; ++q) dst[q] += y.second[size_t(q)];
      }
    crow.clear();
    C.A = B.finish();
    if (!smoothed && opt_.coarse_scale != 1.0)
      for (f64& v : C.A.val) v *= opt_.coarse_scale;
    if (!finalize(C)) return false;
    lv_.push_back(std::move(C));
    return true;
  }

  void factor() {
    const Level& L = lv_.back();
    if (L.A.n > 4 * opt_.coarse_max) {
      chol_n_ = 0;
      chol_.clear();
      return;
    }
    const i32 m = 6 * L.A.n;
    chol_n_ = m;
    build_work_ += static_cast<i64>(m) * m * m / 108;
    chol_.assign(size_t(m) * size_t(m), 0.0);
    auto at = [&](i32 r, i32 c) -> f64& { return chol_[size_t(r) * size_t(m) + size_t(c)]; };
    for (i32 i = 0; i < L.A.n; ++i)
      for (i32 k = L.A.rowptr[size_t(i)]; k < L.A.rowptr[size_t(i) + 1]; ++k) {
        const i32 j = L.A.col[size_t(k)];
        const f64* b = &L.A.val[36 * size_t(k)];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 6; ++c) at(6 * i + r, 6 * j + c) += b[r * 6 + c];
      }
    f64 dmax = 0.0;
    for (i32 i = 0; i < m; ++i) dmax = std::max(dmax, std::abs(at(i, i)));
    for (i32 j = 0; j < m; ++j) {
      f64 d = at(j, j);
      for (i32 k = 0; k < j; ++k) d -= at(j, k) * at(j, k);
      if (!(d > 1e-12 * dmax)) d = std::max(1e-12 * dmax, 1e-300);
      const f64 s = std::sqrt(d);
      at(j, j) = s;
      for (i32 i = j + 1; i < m; ++i) {
        f64 v = 0.5 * (at(i, j) + at(j, i));
        for (i32 k = 0; k < j; ++k) v -= at(i, k) * at(j, k);
        at(i, j) = v / s;
      }
    }
    for (i32 i = 0; i < m; ++i)
      for (i32 k = i + 1; k < m; ++k) at(i, k) = at(k, i);
  }

  void solve_dense(const f64* b, f64* x) const {
    const i32 m = chol_n_;
    auto at = [&](i32 r, i32 c) { return chol_[size_t(r) * size_t(m) + size_t(c)]; };
    for (i32 i = 0; i < m; ++i) {
      f64 v = b[i];
      for (i32 k = 0; k < i; ++k) v -= at(i, k) * x[k];
      x[i] = v / at(i, i);
    }
    for (i32 i = m - 1; i >= 0; --i) {
      f64 v = x[i];
      for (i32 k = i + 1; k < m; ++k) v -= at(i, k) * x[k];
      x[i] = v / at(i, i);
    }
  }

  static void single_block(const f64* B, f32* out, bool transpose) {
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) out[r * 6 + c] = static_cast<f32>(transpose ? B[c * 6 + r] : B[r * 6 + c]);
  }

  void to_single() {
    for (Level& L : lv_) {
      const size_t nb = size_t(L.A.blocks());
      L.Af.assign(36 * nb, 0.0f);
      for (size_t k = 0; k < nb; ++k) single_block(&L.A.val[36 * k], &L.Af[36 * k], false);
      L.Df.assign(36 * size_t(L.A.n), 0.0f);
      for (i32 i = 0; i < L.A.n; ++i) single_block(&L.Dinv[36 * size_t(i)], &L.Df[36 * size_t(i)], false);
      L.Pf.assign(36 * L.Pcol.size(), 0.0f);
      for (size_t e = 0; e < L.Pcol.size(); ++e) single_block(&L.Pval[36 * e], &L.Pf[36 * e], false);
      L.Rf.assign(36 * L.Rent.size(), 0.0f);
      for (size_t r = 0; r < L.Rent.size(); ++r) single_block(&L.Pval[36 * size_t(L.Rent[r].second)], &L.Rf[36 * r], true);
      const size_t m = 6 * size_t(L.A.n);
      L.xf.assign(m, 0.0f);
      L.bf.assign(m, 0.0f);
      L.rf.assign(m, 0.0f);
      L.of.assign(m, 0.0f);
      L.cf.assign(m, 0.0f);
    }
  }

  // hybrid Gauss-Seidel: sequential inside fixed 512-row blocks, the other blocks' values from
  // before the sweep; `fresh` (x = 0) skips what is still zero
  void sweep(const Level& L, f32* x, const f32* b, bool fresh, bool backward) const {
    const Bsr& A = L.A;
    const i32 n = A.n;
    if (!fresh) std::copy(x, x + 6 * size_t(n), L.of.begin());
    const f32* xo = L.of.data();
    for (i64 r0 = 0; r0 < n; r0 += kGrain) {
      const i64 r1 = std::min<i64>(n, r0 + kGrain);
      for (i64 t = 0; t < r1 - r0; ++t) {
        const i64 i = backward ? r1 - 1 - t : r0 + t;
        f32 s[6];
        for (int q = 0; q < 6; ++q) s[q] = b[6 * i + q];
        for (i32 k = A.rowptr[size_t(i)] + 1; k < A.rowptr[size_t(i) + 1]; ++k) {
          const i64 j = A.col[size_t(k)];
          if (j >= r0 && j < r1) {
            if (fresh && j > i) continue;
            fsub(&L.Af[36 * size_t(k)], x + 6 * j, s);
          } else if (!fresh) {
            fsub(&L.Af[36 * size_t(k)], xo + 6 * j, s);
          }
        }
        f32 y[6] = {0, 0, 0, 0, 0, 0};
        fadd(&L.Df[36 * size_t(i)], s, y);
        for (int q = 0; q < 6; ++q) x[6 * i + q] = y[q];
      }
    }
  }

  void cycle(size_t l, const f32* b, f32* x) const {
    const Level& L = lv_[l];
    const i32 n = L.A.n;
    if (l + 1 == lv_.size() && chol_n_ == 0) {
      for (int s = 0; s < 4; ++s) sweep(L, x, b, s == 0, false);
      for (int s = 0; s < 4; ++s) sweep(L, x, b, false, true);
      return;
    }
    if (l + 1 == lv_.size()) {
      std::vector<f64> bd(6 * size_t(n)), xd(6 * size_t(n));
      for (size_t k = 0; k < 6 * size_t(n); ++k) bd[k] = b[k];
      solve_dense(bd.data(), xd.data());
      for (size_t k = 0; k < 6 * size_t(n); ++k) x[k] = static_cast<f32>(xd[k]);
      return;
    }
    for (int s = 0; s < opt_.sweeps; ++s) sweep(L, x, b, s == 0, false);
    const Bsr& A = L.A;
    f32* r = L.rf.data();
    for (i32 i = 0; i < n; ++i) {
      f32 s[6];
      for (int q = 0; q < 6; ++q) s[q] = b[6 * size_t(i) + q];
      for (i32 k = A.rowptr[size_t(i)]; k < A.rowptr[size_t(i) + 1]; ++k) fsub(&L.Af[36 * size_t(k)], x + 6 * size_t(A.col[size_t(k)]), s);
      for (int q = 0; q < 6; ++q) r[6 * size_t(i) + q] = s[q];
    }
    const Level& C = lv_[l + 1];
    f32* bc = C.bf.data();
    for (i32 a = 0; a < C.A.n; ++a) {
      f32 y[6] = {0, 0, 0, 0, 0, 0};
      for (i32 e = L.Rrow[size_t(a)]; e < L.Rrow[size_t(a) + 1]; ++e) fadd(&L.Rf[36 * size_t(e)], r + 6 * size_t(L.Rent[size_t(e)].first), y);
      for (int q = 0; q < 6; ++q) bc[6 * size_t(a) + q] = y[q];
    }
    f32* xc = C.xf.data();
    cycle(l + 1, bc, xc);
    for (i32 i = 0; i < n; ++i) {
      f32 y[6];
      for (int q = 0; q < 6; ++q) y[q] = x[6 * size_t(i) + q];
      for (i32 e = L.Prow[size_t(i)]; e < L.Prow[size_t(i) + 1]; ++e) fadd(&L.Pf[36 * size_t(e)], xc + 6 * size_t(L.Pcol[size_t(e)]), y);
      for (int q = 0; q < 6; ++q) x[6 * size_t(i) + q] = y[q];
    }
    for (int s = 0; s < opt_.sweeps; ++s) sweep(L, x, b, false, true);
  }

  AmgOpt opt_;
  std::vector<Level> lv_;
  std::vector<f64> chol_;
  i32 chol_n_ = 0;
  i64 work_ = 0, build_work_ = 0;
};

// ---- the stress problem ------------------------------------------------------------------------

const svxg_material& mat_of(const std::vector<svxg_material>& mats, int id) {
  return (id >= 0 && size_t(id) < mats.size()) ? mats[size_t(id)] : mats[0];
}

// sum of a[k] b[k] in chunks of 4096 entries, the chunks' sums added in chunk order
f64 chunked_dot(const std::vector<f64>& a, const std::vector<f64>& b) {
  const i64 m = static_cast<i64>(a.size());
  if (m <= 0) return 0.0;
  const i64 g = 8 * kGrain;
  const i64 nc = (m + g - 1) / g;
  auto part = [&](i64 k0, i64 k1) {
    f64 s = 0.0;
    for (i64 k = k0; k < k1; ++k) s += a[size_t(k)] * b[size_t(k)];
    return s;
  };
  if (nc == 1) return part(0, m);
  f64 s = 0.0;
  for (i64 c = 0; c < nc; ++c) s += part(c * g, std::min(m, (c + 1) * g));
  return s;
}

}  // namespace

struct orc_problem {
  std::vector<svxg_node> nodes;
  std::vector<svxg_bond> bonds;
  std::vector<svxg_material> mats;
  svxg_stress_options opt{};
  bool assembled = false, pc_current = false, jacobi_only = false;
  i64 assembly_work = 0;
  i32 pc_n = 0, nfree = 0;
  f64 id_scale = 1.0;
  std::vector<f64> jinv;
  std::vector<i32> dof;
  Bsr K;
  Multigrid amg;
  struct Run {
    bool active = false;
    std::vector<f64> x, r, z, p, q;
    f64 rz = 0.0, bn = 0.0;
    int iters = 0;
  } run;

  void bond_matrices(const svxg_bond& b, f64 D[36], f64 Ba[36], f64 Bb[36]) const {
    const svxg_material& A = mat_of(mats, b.ma);
    const svxg_material& B = mat_of(mats, b.mb);
    const f64 la = std::max(b.la, 1e-4), lb = std::max(b.lb, 1e-4);
    const f64 invE = la / A.E + lb / B.E;
    const f64 invG = la / A.G + lb / B.G;
    const f64 kt[3] = {b.area / invE, (5.0 / 6.0) * b.area / invG, (5.0 / 6.0) * b.area / invG};
    const f64 kr[3] = {0.8 * (b.s1 + b.s2) / invG, b.s2 / invE, b.s1 / invE};
    const V3 Q[3] = {v3(b.n), v3(b.t1), v3(b.t2)};
    for (int q = 0; q < 36; ++q) D[q] = 0.0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        f64 st = 0.0, sr = 0.0;
        for (int k = 0; k < 3; ++k) {
          st += vget(Q[k], r) * kt[k] * vget(Q[k], c);
          sr += vget(Q[k], r) * kr[k] * vget(Q[k], c);
        }
        D[r * 6 + c] = st;
        D[(3 + r) * 6 + 3 + c] = sr;
      }
    rigid(vsub(v3(b.p), v3(nodes[size_t(b.a)].c)), Ba);
    if (b.b >= 0) rigid(vsub(v3(b.p), v3(nodes[size_t(b.b)].c)), Bb);
  }

  void build_matrix() {
    const i32 n = static_cast<i32>(nodes.size());
    dof.assign(size_t(n), -1);
    nfree = 0;
    for (i32 i = 0; i < n; ++i)
      if (!nodes[size_t(i)].fixed) dof[size_t(i)] = nfree++;
    Builder B(nfree);
    for (i32 i = 0; i < n; ++i)
      if (dof[size_t(i)] >= 0) B.at(dof[size_t(i)], dof[size_t(i)]);
    f64 D[36], Ba[36], Bb[36];
    for (svxg_bond& b : bonds) {
      b.in_k = 0;
      if (b.broken || nodes[size_t(b.a)].gone || (b.b >= 0 && nodes[size_t(b.b)].gone)) continue;
      b.in_k = 1;
      bond_matrices(b, D, Ba, Bb);
      const i32 ia = dof[size_t(b.a)];
      const i32 ib = b.b >= 0 ? dof[size_t(b.b)] : -1;
      if (ia >= 0) atbd_add(Ba, D, Ba, 1.0, B.at(ia, ia));
      if (ib >= 0) atbd_add(Bb, D, Bb, 1.0, B.at(ib, ib));
      if (ia >= 0 && ib >= 0) {
        atbd_add(Ba, D, Bb, -1.0, B.at(ia, ib));
        atbd_add(Bb, D, Ba, -1.0, B.at(ib, ia));
      }
    }
    K = B.finish();
    f64 tr = 0.0;
    i32 cnt = 0;
    for (i32 i = 0; i < nfree; ++i) {
      const f64* d = K.diag(i);
      const f64 t = d[0] + d[7] + d[14];
      if (t > 0) {
        tr += t / 3.0;
        ++cnt;
      }
    }
    id_scale = cnt ? tr / cnt : 1.0;
    for (i32 i = 0; i < n; ++i) {
      const i32 d = dof[size_t(i)];
      if (d < 0) continue;
      f64* Dg = K.diag(d);
      const bool empty = Dg[0] == 0.0 && Dg[7] == 0.0 && Dg[14] == 0.0;
      if (nodes[size_t(i)].gone || empty)
        for (int q = 0; q < 6; ++q) Dg[q * 7] = id_scale;
    }
  }

  bool assemble(const svxg_stress_options& o) {
    opt = o;
    assembled = false;
    pc_current = true;
    run.active = false;
    build_matrix();
    assembly_work = 0;
    for (const svxg_bond& b : bonds)
      if (b.in_k) assembly_work += 48;
    std::vector<V3> pos(static_cast<size_t>(nfree));
    for (size_t i = 0; i < nodes.size(); ++i)
      if (dof[i] >= 0) pos[size_t(dof[i])] = v3(nodes[i].c);
    jacobi_only = nfree < opt.amg_min_nodes;
    if (jacobi_only) {
      amg = Multigrid{};
      pc_n = 0;
      jinv.assign(36 * size_t(nfree), 0.0);
      for (i32 i = 0; i < nfree; ++i)
        if (!inv6(K.diag(i), &jinv[36 * size_t(i)]))
          for (int q = 0; q < 6; ++q) jinv[36 * size_t(i) + size_t(q * 7)] = 1.0 / id_scale;
      assembly_work += 12 * static_cast<i64>(nfree);
      assembled = true;
      return true;
    }
    AmgOpt ao;
    ao.coarse_max = opt.amg.coarse_max;
    ao.strength = opt.amg.strength;
    ao.smoothed = opt.amg.smoothed != 0;
    ao.coarse_scale = opt.amg.coarse_scale;
    ao.max_levels = opt.amg.max_levels;
    ao.sweeps = opt.amg.sweeps;
    ao.coarsen_dense = opt.amg.coarsen_dense != 0;
    ao.reaggregate = opt.amg.reaggregate != 0;
    if (!amg.build(K, pos, ao)) return false;
    assembly_work += amg.build_work();
    pc_n = nfree;
    jinv.clear();
    assembled = true;
    return true;
  }

  void release() {
    assembled = pc_current = jacobi_only = false;
    pc_n = nfree = 0;
    run = Run{};
    K = Bsr{};
    amg = Multigrid{};
    jinv.clear();
    dof.clear();
    for (svxg_bond& b : bonds) b.in_k = 0;
  }

  void remove_bond(i32 bi) {
    svxg_bond& b = bonds[size_t(bi)];
    b.broken = 1;
    if (!assembled || !b.in_k) return;
    b.in_k = 0;
    f64 D[36], Ba[36], Bb[36];
    bond_matrices(b, D, Ba, Bb);
    const i32 ia = dof[size_t(b.a)];
    const i32 ib = b.b >= 0 ? dof[size_t(b.b)] : -1;
    if (ia >= 0) atbd_add(Ba, D, Ba, -1.0, K.find(ia, ia));
    if (ib >= 0) atbd_add(Bb, D, Bb, -1.0, K.find(ib, ib));
    if (ia >= 0 && ib >= 0) {
      if (f64* k = K.find(ia, ib)) atbd_add(Ba, D, Bb, 1.0, k);
      if (f64* k = K.find(ib, ia)) atbd_add(Bb, D, Ba, 1.0, k);
    }
    if (!jacobi_only) pc_current = false;
    if (jacobi_only)
      for (i32 d : {ia, ib}) {
        if (d < 0) continue;
        if (!inv6(K.diag(d), &jinv[36 * size_t(d)]))
          for (int q = 0; q < 36; ++q) jinv[36 * size_t(d) + size_t(q)] = (q % 7 == 0) ? 1.0 / id_scale : 0.0;
      }
    run.active = false;
  }

  void retire_nodes(const i32* ids, i32 count) {
    if (count <= 0) return;
    std::vector<uint8_t> now(nodes.size(), 0);
    for (i32 t = 0; t < count; ++t) {
      const i32 i = ids[t];
      now[size_t(i)] = nodes[size_t(i)].gone ? 0 : 1;
      nodes[size_t(i)].gone = 1;
    }
    for (i32 b = 0; b < static_cast<i32>(bonds.size()); ++b) {
      const svxg_bond& B = bonds[size_t(b)];
      if (now[size_t(B.a)] || (B.b >= 0 && now[size_t(B.b)])) remove_bond(b);
    }
    if (assembled)
      for (i32 t = 0; t < count; ++t) {
        const i32 i = ids[t];
        const i32 d = dof[size_t(i)];
        if (d < 0) continue;
        if (now[size_t(i)] && !jacobi_only) pc_current = false;
        f64* Dg = K.find(d, d);
        for (int q = 0; q < 36; ++q) Dg[q] = (q % 7 == 0) ? id_scale : 0.0;
        if (jacobi_only)
          for (int q = 0; q < 36; ++q) jinv[36 * size_t(d) + size_t(q)] = (q % 7 == 0) ? 1.0 / id_scale : 0.0;
      }
    run.active = false;
  }

  void precondition(const f64* r, f64* z) const {
    if (!jacobi_only) {
      amg.apply(r, z);
      return;
    }
    for (i32 i = 0; i < nfree; ++i) mv6(&jinv[36 * size_t(i)], r + 6 * size_t(i), z + 6 * size_t(i));
  }

  void begin(const f64* f, const f64* u0) {
    run.active = false;
    if (!assembled && !assemble(opt)) return;
    const size_t n = nodes.size();
    const size_t m = 6 * size_t(nfree);
    run.x.assign(m, 0.0);
    std::vector<f64> fb(m, 0.0);
    for (size_t i = 0; i < n; ++i) {
      const i32 d = dof[i];
      if (d < 0) continue;
      for (int q = 0; q < 6; ++q) {
        fb[6 * size_t(d) + size_t(q)] = f[6 * i + size_t(q)];
        if (u0) run.x[6 * size_t(d) + size_t(q)] = u0[6 * i + size_t(q)];
      }
    }
    run.r.assign(m, 0.0);
    run.z.assign(m, 0.0);
    run.p.assign(m, 0.0);
    run.q.assign(m, 0.0);
    K.apply(run.x.data(), run.q.data());
    f64 bn = 0.0;
    for (size_t k = 0; k < m; ++k) {
      run.r[k] = fb[k] - run.q[k];
      bn += fb[k] * fb[k];
    }
    run.bn = std::sqrt(bn);
    if (!(bn > 0.0)) {
      std::fill(run.x.begin(), run.x.end(), 0.0);
      std::fill(run.r.begin(), run.r.end(), 0.0);
    }
    precondition(run.r.data(), run.z.data());
    run.p = run.z;
    run.rz = 0.0;
    for (size_t k = 0; k < m; ++k) run.rz += run.r[k] * run.z[k];
    run.iters = 0;
    run.active = true;
  }

  svxg_pcg_result iterate(int maxit, f64 rtol) {
    svxg_pcg_result res{};
    if (!run.active) return res;
    const size_t m = run.x.size();
    f64 rn = std::sqrt(chunked_dot(run.r, run.r));
    const f64 target = rtol * run.bn;
    if (!std::isfinite(run.bn) || !std::isfinite(rn)) {
      res.rel_res = INFINITY;
      res.breakdown = 1;
      run.active = false;
      return res;
    }
    res.rel_res = run.bn > 0 ? rn / run.bn : 0.0;
    if (rn <= target || !(run.bn > 0) || m == 0) {
      res.converged = 1;
      return res;
    }
    for (int it = 0; it < maxit; ++it) {
      K.apply(run.p.data(), run.q.data());
      const f64 pq = chunked_dot(run.p, run.q);
      ++res.iters;
      ++run.iters;
      if (!(pq > 0)) {
        res.converged = 0;
        res.breakdown = 1;
        break;
      }
      const f64 alpha = run.rz / pq;
      for (size_t k = 0; k < m; ++k) {
        run.x[k] += alpha * run.p[k];
        run.r[k] -= alpha * run.q[k];
      }
      rn = std::sqrt(chunked_dot(run.r, run.r));
      res.rel_res = rn / run.bn;
      if (rn <= target) {
        res.converged = 1;
        break;
      }
      precondition(run.r.data(), run.z.data());
      const f64 rz1 = chunked_dot(run.r, run.z);
      if (!(rz1 > 0)) {
        res.breakdown = 1;
        break;
      }
      const f64 beta = rz1 / run.rz;
      run.rz = rz1;
      for (size_t k = 0; k < m; ++k) run.p[k] = run.z[k] + beta * run.p[k];
    }
    return res;
  }

  void current(f64* u) const {
    const size_t n = nodes.size();
    std::fill(u, u + 6 * n, 0.0);
    if (run.x.size() != 6 * size_t(nfree) || dof.size() != n) return;
    for (size_t i = 0; i < n; ++i) {
      const i32 d = dof[i];
      if (d < 0) continue;
      for (int q = 0; q < 6; ++q) u[6 * i + size_t(q)] = run.x[6 * size_t(d) + size_t(q)];
    }
  }

  svxg_bond_load bond_load(i32 bi, const f64* u) const {
    const svxg_bond& b = bonds[size_t(bi)];
    const svxg_material& A = mat_of(mats, b.ma);
    const svxg_material& B = mat_of(mats, b.mb);
    f64 Ba[36], Bb[36];
    rigid(vsub(v3(b.p), v3(nodes[size_t(b.a)].c)), Ba);
    f64 xa[6] = {0, 0, 0, 0, 0, 0}, xb[6] = {0, 0, 0, 0, 0, 0};
    if (!nodes[size_t(b.a)].fixed) mv6(Ba, &u[6 * size_t(b.a)], xa);
    if (b.b >= 0) {
      rigid(vsub(v3(b.p), v3(nodes[size_t(b.b)].c)), Bb);
      if (!nodes[size_t(b.b)].fixed) mv6(Bb, &u[6 * size_t(b.b)], xb);
    }
    const V3 dt{xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2]};
    const V3 dr{xb[3] - xa[3], xb[4] - xa[4], xb[5] - xa[5]};
    const f64 la = std::max(b.la, 1e-4), lb = std::max(b.lb, 1e-4);
    const f64 invE = la / A.E + lb / B.E;
    const f64 invG = la / A.G + lb / B.G;
    svxg_bond_load L;
    L.N = b.area / invE * vdot(dt, v3(b.n));
    L.V1 = (5.0 / 6.0) * b.area / invG * vdot(dt, v3(b.t1));
    L.V2 = (5.0 / 6.0) * b.area / invG * vdot(dt, v3(b.t2));
    L.T = 0.8 * (b.s1 + b.s2) / invG * vdot(dr, v3(b.n));
    L.M1 = b.s2 / invE * vdot(dr, v3(b.t1));
    L.M2 = b.s1 / invE * vdot(dr, v3(b.t2));
    return L;
  }
};

namespace {

void strengths(const svxg_bond& b, f64 fragility, const svxg_material* mats, i32 nmats, f64 S[5]) {
  const f64 s = b.strength / std::max(1e-6, fragility);
  if (b.sectioned) {
    S[0] = s * b.ft;
    S[1] = s * b.fb;
    S[2] = s * b.fc;
    S[3] = s * b.coh;
    S[4] = static_cast<f64>(b.mu);
    return;
  }
  const svxg_material& A = (b.ma < nmats) ? mats[b.ma] : mats[0];
  const svxg_material& B = (b.mb < nmats) ? mats[b.mb] : mats[0];
  S[0] = s * std::min(A.ft, B.ft);
  S[1] = s * std::min(A.fb, B.fb);
  S[2] = s * std::min(A.fc, B.fc);
  S[3] = s * std::min(A.cohesion, B.cohesion);
  S[4] = std::min(A.friction, B.friction);
}

}  // namespace

extern "C" {

orc_problem* orc_create(const svxg_node* nodes, int32_t n, const svxg_bond* bonds, int32_t m, const svxg_material* mats,
                        int32_t nmats) {
  orc_problem* h = new orc_problem;
  h->nodes.assign(nodes, nodes + n);
  h->bonds.assign(bonds, bonds + m);
  h->mats.assign(mats, mats + nmats);
  orc_default_options(&h->opt);
  return h;
}

void orc_destroy(orc_problem* h) { delete h; }

int orc_assemble(orc_problem* h, const svxg_stress_options* opt) { return h->assemble(*opt) ? 1 : 0; }
void orc_release(orc_problem* h) { h->release(); }

int orc_solve(orc_problem* h, const double* f, double* u, double rtol, int32_t maxit, int32_t warm, svxg_pcg_result* res) {
  *res = svxg_pcg_result{};
  if (!h->assembled && !h->assemble(h->opt)) {
    res->rel_res = INFINITY;
    res->breakdown = 1;
    return 0;
  }
  h->begin(f, warm ? u : nullptr);
  *res = h->iterate(maxit < 0 ? 100000 : maxit, rtol);
  h->current(u);
  h->run.active = false;
  return 1;
}

void orc_begin(orc_problem* h, const double* f, const double* u0) { h->begin(f, u0); }
void orc_iterate(orc_problem* h, int32_t maxit, double rtol, svxg_pcg_result* res) { *res = h->iterate(maxit, rtol); }
void orc_current(orc_problem* h, double* u) { h->current(u); }
void orc_remove_bond(orc_problem* h, int32_t b) { h->remove_bond(b); }
void orc_retire_nodes(orc_problem* h, const int32_t* ids, int32_t count) { h->retire_nodes(ids, count); }
void orc_set_bond_broken(orc_problem* h, int32_t b, int32_t broken) { h->bonds[size_t(b)].broken = broken ? 1 : 0; }
void orc_bond_load(orc_problem* h, int32_t b, const double* u, svxg_bond_load* out) { *out = h->bond_load(b, u); }

// StressProblem::extend_warm_start (stress.cpp:213-250): passes over the intact internal bonds; a
// node without a motion yet takes the mean, over its bonded neighbours that have one, of the
// neighbour's translation plus rotation x (its centre - the neighbour's), and rotation as it is
void orc_extend_warm_start(const svxg_node* nodes, int32_t n, const svxg_bond* bonds, int32_t m, double* u,
                           int32_t first_new) {
  if (first_new >= n) return;
  std::vector<f64> acc(6 * size_t(n), 0.0), w(size_t(n), 0.0);
  std::vector<uint8_t> known(size_t(n), 0);
  for (i32 i = 0; i < first_new; ++i) known[size_t(i)] = nodes[i].gone ? 0 : 1;
  for (int pass = 0; pass < 4; ++pass) {
    std::fill(acc.begin(), acc.end(), 0.0);
    std::fill(w.begin(), w.end(), 0.0);
    for (i32 k = 0; k < m; ++k) {
      const svxg_bond& b = bonds[k];
      if (b.broken || b.b < 0) continue;
      i32 from = -1, to = -1;
      if (known[size_t(b.a)] && !known[size_t(b.b)] && b.b >= first_new) from = b.a, to = b.b;
      else if (known[size_t(b.b)] && !known[size_t(b.a)] && b.a >= first_new) from = b.b, to = b.a;
      if (from < 0) continue;
      const V3 r = vsub(v3(nodes[to].c), v3(nodes[from].c));
      const V3 th{u[6 * size_t(from) + 3], u[6 * size_t(from) + 4], u[6 * size_t(from) + 5]};
      const V3 t = vcross(th, r);
      for (int q = 0; q < 3; ++q) acc[6 * size_t(to) + size_t(q)] += u[6 * size_t(from) + size_t(q)] + vget(t, q);
      for (int q = 3; q < 6; ++q) acc[6 * size_t(to) + size_t(q)] += u[6 * size_t(from) + size_t(q)];
      w[size_t(to)] += 1.0;
    }
    bool any = false;
    for (i32 i = first_new; i < n; ++i) {
      if (known[size_t(i)] || w[size_t(i)] == 0.0) continue;
      for (int q = 0; q < 6; ++q) u[6 * size_t(i) + size_t(q)] = acc[6 * size_t(i) + size_t(q)] / w[size_t(i)];
      known[size_t(i)] = 1;
      any = true;
    }
    if (!any) break;
  }
}

void orc_bond_strength(const svxg_bond* b, double fragility, const svxg_material* mats, int32_t nmats, double out[5]) {
  strengths(*b, fragility, mats, nmats, out);
}

double orc_bond_utilization(const svxg_bond* bp, const svxg_bond_load* Lp, double fragility, const svxg_material* mats,
                            int32_t nmats, uint8_t* mode) {
  const svxg_bond& b = *bp;
  const svxg_bond_load& L = *Lp;
  f64 S[5];
  strengths(b, fragility, mats, nmats, S);
  constexpr f64 kWeak = 1e-9;
  const f64 ft = std::max(S[0], kWeak), fb = std::max(S[1], kWeak), fc = std::max(S[2], kWeak);
  const f64 A = std::max(b.area, 1e-12);
  const f64 sN = L.N / A;
  const f64 sb = (b.s2 > 0 ? std::abs(L.M1) * b.c2 / b.s2 : 0.0) + (b.s1 > 0 ? std::abs(L.M2) * b.c1 / b.s1 : 0.0);
  const f64 phi_t = sN >= 0.0 ? sN / ft + sb / fb : std::max(0.0, sb + sN) / fb;
  const f64 phi_c = std::max(0.0, sb - sN) / fc;
  const f64 J = std::max(b.s1 + b.s2, 1e-18);
  const f64 tau = 1.5 * std::sqrt(L.V1 * L.V1 + L.V2 * L.V2) / A + std::abs(L.T) * b.rmax / J;
  const f64 cap = std::max(S[3] + S[4] * std::max(0.0, -sN), kWeak);
  const f64 phi_s = tau / cap;
  f64 phi = phi_t;
  uint8_t md = 1;
  if (phi_c > phi) {
    phi = phi_c;
    md = 2;
  }
  if (phi_s > phi) {
    phi = phi_s;
    md = 3;
  }
  if (mode) *mode = phi > 0 ? md : 0;
  return phi;
}

// every bond's loads and checks from one u (orc_bond_load + orc_bond_utilization of each bond)
void orc_bond_checks(orc_problem* h, const double* u, double fragility, svxg_bond_load* loads, double* phi,
                     uint8_t* mode) {
  for (size_t b = 0; b < h->bonds.size(); ++b) {
    const svxg_bond_load L = h->bond_load(static_cast<i32>(b), u);
    if (loads) loads[b] = L;
    if (phi) phi[b] = orc_bond_utilization(&h->bonds[b], &L, fragility, h->mats.data(), static_cast<i32>(h->mats.size()),
                                           mode ? &mode[b] : nullptr);
  }
}

void orc_lerp_load(const svxg_bond_load* prev, const svxg_bond_load* now, double k, svxg_bond_load* out) {
  const f64* p = &prev->N;
  const f64* q = &now->N;
  f64* r = &out->N;
  for (int c = 0; c < 6; ++c) r[c] = p[c] + k * (q[c] - p[c]);
}

int32_t orc_hinge_of(const svxg_bond* bp, const svxg_bond_load* Lp, const double rot_a[3], const double rot_b[3],
                     const svxg_material* mats, int32_t nmats, const uint64_t ductile[2], double fragility,
                     double hinge_shape, svxg_hinge_cut* out) {
  const svxg_bond& b = *bp;
  const svxg_bond_load& L = *Lp;
  auto is_ductile = [&](int m) { return m >= 0 && m < 128 && ((ductile[m >> 6] >> (m & 63)) & 1ull); };
  if (!is_ductile(b.ma) || !is_ductile(b.mb)) return 0;
  f64 S[5];
  strengths(b, fragility, mats, nmats, S);
  const f64 area = std::max(b.area, 1e-12);
  const f64 sN = L.N / area;
  const f64 w1 = b.s2 > 0 ? b.c2 / b.s2 : 0.0;
  const f64 w2 = b.s1 > 0 ? b.c1 / b.s1 : 0.0;
  const f64 sb = std::abs(L.M1) * w1 + std::abs(L.M2) * w2;
  const f64 fb = std::max(S[1], 1e-9), ft = std::max(S[0], 1e-9), fc = std::max(S[2], 1e-9);
  const f64 pb = sb / fb;
  const f64 pn = sN >= 0.0 ? sN / ft : -sN / fc;
  const f64 J = std::max(b.s1 + b.s2, 1e-18);
  const f64 tau = 1.5 * std::sqrt(L.V1 * L.V1 + L.V2 * L.V2) / area + std::abs(L.T) * b.rmax / J;
  const f64 ps = tau / std::max(S[3] + S[4] * std::max(0.0, -sN), 1e-9);
  // bending governs by half again over pull and shear, or it is no hinge
  if (!(pb > 0.0) || pb < 1.5 * pn || pb < 1.5 * ps) return 0;
  const V3 n = v3(b.n), t1 = v3(b.t1), t2 = v3(b.t2);
  const V3 ra = v3(rot_a);
  const V3 rb = b.b < 0 ? V3{0.0, 0.0, 0.0} : v3(rot_b);
  V3 w = vsub(rb, ra);
  const f64 wn = vdot(w, n);
  w = vsub(w, V3{n.x * wn, n.y * wn, n.z * wn});
  const V3 mt{t1.x * L.M1 + t2.x * L.M2, t1.y * L.M1 + t2.y * L.M2, t1.z * L.M1 + t2.z * L.M2};
  svxg_hinge_cut hc{};
  hc.a = b.a;
  hc.b = b.b;
  for (int c = 0; c < 3; ++c) {
    hc.n[c] = b.n[c];
    hc.p[c] = b.p[c];
  }
  V3 axis;
  if (vdot(w, w) > 1e-30) {
    const f64 inv = 1.0 / std::sqrt(vdot(w, w));
    axis = V3{w.x * inv, w.y * inv, w.z * inv};
    V3 y = vcross(axis, n);
    const f64 yl = std::sqrt(vdot(y, y));
    if (yl > 1e-9) {
      const f64 iy = 1.0 / yl;
      y = V3{y.x * iy, y.y * iy, y.z * iy};
      const f64 edge = std::abs(vdot(y, t1)) * b.c1 + std::abs(vdot(y, t2)) * b.c2;
      for (int c = 0; c < 3; ++c) hc.p[c] = b.p[c] - vget(y, c) * (0.9 * edge);
    }
  } else if (vdot(mt, mt) > 1e-30) {
    const f64 inv = 1.0 / std::sqrt(vdot(mt, mt));
    axis = V3{mt.x * inv, mt.y * inv, mt.z * inv};
  } else {
    return 0;
  }
  hc.axis[0] = axis.x;
  hc.axis[1] = axis.y;
  hc.axis[2] = axis.z;
  const f64 a1 = std::abs(vdot(axis, t1)), a2 = std::abs(vdot(axis, t2));
  hc.mp = std::max(0.0, hinge_shape) * fb / std::max(a1 * w1 + a2 * w2, 1e-12);
  hc.pull = ft * area;
  *out = hc;
  return 1;
}

int32_t orc_judge(orc_problem* h, const double* u, const svxg_judge_options* opt, const svxg_judge_state* st,
                  svxg_bond_load* base, uint8_t* has, double* phi_out, svxg_judge_result* res, svxg_break* picks) {
  const i32 m = static_cast<i32>(h->bonds.size());
  const i32 nm = static_cast<i32>(h->mats.size());
  const bool piece = opt->rule == SVXG_JUDGE_PIECE;
  std::vector<std::pair<f64, i32>> over;
  std::vector<uint8_t> modes(size_t(m), 0);
  f64 maxphi = 0.0;
  for (i32 k = 0; k < m; ++k) {
    const svxg_bond& b = h->bonds[size_t(k)];
    if (b.broken || (piece && h->nodes[size_t(b.a)].gone)) {
      if (phi_out) phi_out[k] = 0.0;
      continue;
    }
    const svxg_bond_load L = h->bond_load(k, u);
    svxg_bond_load Le = L;
    if (st->shock && opt->dif != 1.0 && base && has[k]) orc_lerp_load(&base[k], &L, opt->dif, &Le);
    const f64 phi = orc_bond_utilization(&b, &Le, opt->fragility, h->mats.data(), nm, &modes[size_t(k)]);
    if (!st->transient && base) {
      base[k] = L;
      has[k] = 1;
    }
    if (phi_out) phi_out[k] = phi;
    maxphi = std::max(maxphi, phi);
    if (phi >= 1.0) over.push_back({phi, k});
  }
  *res = svxg_judge_result{};
  res->max_phi = maxphi;
  res->over = static_cast<i32>(over.size());
  if (over.empty() || st->rounds >= opt->max_rounds) return 0;
  std::sort(over.begin(), over.end(), [](const std::pair<f64, i32>& x, const std::pair<f64, i32>& y) {
    return x.first != y.first ? x.first > y.first : x.second < y.second;
  });
  const f64 thr = std::max(1.0, opt->break_band * over.front().first);
  const size_t n_over = over.size();
  const size_t quota = piece ? static_cast<size_t>(std::ceil(std::clamp(opt->quota_fraction, 0.0, 1.0) * f64(n_over)))
                             : std::max<size_t>(1, (n_over + 3) / 4);
  i32 count = 0;
  for (size_t q = 0; q < n_over; ++q) {
    if ((over[q].first < thr && q >= quota) || count >= opt->max_breaks) break;
    const i32 k = over[q].second;
    const svxg_bond& b = h->bonds[size_t(k)];
    svxg_break r{};
    r.bond = k;
    r.round = st->rounds;
    r.phi = over[q].first;
    r.mode = modes[size_t(k)];
    if (opt->plastic_hinges && (!piece || b.b >= 0)) {
      const double zero[3] = {0.0, 0.0, 0.0};
      const double* ra = u + 6 * size_t(b.a) + 3;
      const double* rb = b.b >= 0 ? u + 6 * size_t(b.b) + 3 : zero;
      const svxg_bond_load L = h->bond_load(k, u);
      if (orc_hinge_of(&b, &L, ra, rb, h->mats.data(), nm, opt->ductile, opt->fragility, opt->hinge_shape, &r.cut)) {
        r.hinge = 1;
        ++res->hinges;
      }
    }
    picks[count++] = r;
  }
  res->picked = count;
  return count;
}

void orc_bond_sections(const svxg_node* nodes, int32_t, const svxg_section_head* heads, int32_t m,
                       const svxg_section_face* faces, int32_t, const svxg_voxel_box* box,
                       const svxg_material* mats, int32_t nmats, svxg_bond* out) {
  const f64 h = box->h, A1 = h * h;
  auto at = [&](const int32_t p[3], uint8_t* v, uint8_t* dmg) {
    i64 idx = 0, mul = 1;
    for (int c = 0; c < 3; ++c) {
      const i32 r = p[c] - box->lo[c];
      if (r < 0 || r >= box->dim[c]) {
        *v = 0;
        *dmg = 0;
        return;
      }
      idx += mul * r;
      mul *= box->dim[c];
    }
    *v = box->vox[idx];
    *dmg = box->damage ? box->damage[idx] : 0;
  };
  auto material = [&](uint8_t v) -> const svxg_material& {
    const int id = (v & 0x7F) - 1;
    return mats[(id >= 0 && id < nmats) ? id : 0];
  };
  auto comp = [](V3& a, int i) -> f64& { return i == 0 ? a.x : (i == 1 ? a.y : a.z); };
  auto unit = [](const V3& a) {
    const f64 l = std::sqrt(vdot(a, a));
    return l > 0 ? V3{a.x * (1.0 / l), a.y * (1.0 / l), a.z * (1.0 / l)} : V3{0, 0, 0};
  };
  for (i32 k = 0; k < m; ++k) {
    const svxg_section_head& hd = heads[k];
    const svxg_section_face* F = faces + hd.face0;
    const i32 nf = hd.nfaces;
    const V3 ca = v3(nodes[hd.a].c);
    const bool pair = hd.b >= 0;
    const V3 cb = pair ? v3(nodes[hd.b].c) : V3{0, 0, 0};
    // the interface normal: the faces' mean normal from a to b; a support's face axis
    V3 n{0, 0, 0};
    if (pair) {
      for (i32 f = 0; f < nf; ++f) comp(n, F[f].axis) += F[f].sign;
      if (vdot(n, n) < 1e-12) n = vsub(cb, ca);
      if (vdot(n, n) < 1e-12) n = V3{0, 0, 1};
      n = unit(n);
    } else {
      comp(n, hd.axis) = hd.sign;
    }
    // frame_of (world_internal.hpp:64-69)
    const f64 ax = std::abs(n.x), ay = std::abs(n.y), az = std::abs(n.z);
    const V3 e = (ax <= ay && ax <= az) ? V3{1, 0, 0} : (ay <= az ? V3{0, 1, 0} : V3{0, 0, 1});
    V3 t1 = unit(vcross(n, e));
    V3 t2 = vcross(n, t1);
    std::vector<V3> xs(static_cast<size_t>(nf));
    std::vector<f64> ws(static_cast<size_t>(nf));
    f64 W = 0.0;
    V3 pc{0, 0, 0};
    for (i32 f = 0; f < nf; ++f) {
      V3 x{h * F[f].p[0], h * F[f].p[1], h * F[f].p[2]};
      comp(x, F[f].axis) += 0.5 * h;
      xs[size_t(f)] = x;
      V3 nn = n;
      ws[size_t(f)] = A1 * std::max(0.05, std::abs(comp(nn, F[f].axis)));
      W += ws[size_t(f)];
      pc = V3{pc.x + x.x * ws[size_t(f)], pc.y + x.y * ws[size_t(f)], pc.z + x.z * ws[size_t(f)]};
    }
    pc = V3{pc.x * (1.0 / W), pc.y * (1.0 / W), pc.z * (1.0 / W)};
    // principal axes of the section (the eigenvector of the larger eigenvalue, sqrt only)
    f64 s11 = 0.0, s22 = 0.0, s12 = 0.0;
    for (i32 f = 0; f < nf; ++f) {
      const V3 d = vsub(xs[size_t(f)], pc);
      const f64 y1 = vdot(d, t1), y2 = vdot(d, t2);
      s11 += ws[size_t(f)] * y1 * y1;
      s22 += ws[size_t(f)] * y2 * y2;
      s12 += ws[size_t(f)] * y1 * y2;
    }
    if (std::abs(s12) > 1e-9 * (s11 + s22)) {
      const f64 half = 0.5 * (s11 - s22);
      const f64 l1 = 0.5 * (s11 + s22) + std::sqrt(half * half + s12 * s12);
      f64 e1 = s12, e2 = l1 - s11;
      if (std::abs(l1 - s22) > std::abs(e2)) {
        e1 = l1 - s22;
        e2 = s12;
      }
      const f64 en = std::sqrt(e1 * e1 + e2 * e2);
      if (en > 0.0) {
        e1 /= en;
        e2 /= en;
        const V3 u1{t1.x * e1 + t2.x * e2, t1.y * e1 + t2.y * e2, t1.z * e1 + t2.z * e2};
        const V3 u2{t2.x * e1 - t1.x * e2, t2.y * e1 - t1.y * e2, t2.z * e1 - t1.z * e2};
        t1 = u1;
        t2 = u2;
      }
    }
    svxg_bond B{};
    B.a = hd.a;
    B.b = hd.b;
    B.ma = hd.ma;
    B.mb = hd.mb;
    B.faces = nf;
    B.tag = k;
    for (int c = 0; c < 3; ++c) {
      B.n[c] = vget(n, c);
      B.t1[c] = vget(t1, c);
      B.t2[c] = vget(t2, c);
      B.p[c] = vget(pc, c);
    }
    B.area = std::max(W, 0.25 * nf * A1);
    f64 s1 = 0.0, s2 = 0.0, m1 = 0.0, m2 = 0.0;
    for (i32 f = 0; f < nf; ++f) {
      const V3 d = vsub(xs[size_t(f)], pc);
      const f64 y1 = vdot(d, t1), y2 = vdot(d, t2);
      s1 += ws[size_t(f)] * y1 * y1;
      s2 += ws[size_t(f)] * y2 * y2;
      m1 = std::max(m1, std::abs(y1));
      m2 = std::max(m2, std::abs(y2));
    }
    B.s1 = s1 + B.area * A1 / 12.0;
    B.s2 = s2 + B.area * A1 / 12.0;
    B.c1 = m1 + 0.5 * h;
    B.c2 = m2 + 0.5 * h;
    B.rmax = std::sqrt(B.c1 * B.c1 + B.c2 * B.c2);
    if (pair) {
      const V3 d = vsub(cb, ca);
      const f64 L = std::max({std::abs(vdot(d, n)), 0.5 * std::sqrt(vdot(d, d)), 0.5 * h});
      B.la = std::clamp(vdot(vsub(pc, ca), n), 0.1 * L, 0.9 * L);
      B.lb = L - B.la;
    } else {
      B.la = std::max(0.25 * h, vdot(vsub(pc, ca), n));
      B.lb = 0.5 * h;
    }
    B.strength = std::min(hd.strength_a, hd.strength_b);
    // section_strengths: per face the weaker material of its two voxels (an air side takes the
    // other's), times the condition of the more damaged; the faces' mean
    f64 ft = 0, fb = 0, fc = 0, coh = 0, mu = 0;
    for (i32 f = 0; f < nf; ++f) {
      uint8_t va, da, vb, db;
      at(F[f].p, &va, &da);
      int32_t q[3] = {F[f].p[0], F[f].p[1], F[f].p[2]};
      q[F[f].axis] += 1;
      at(q, &vb, &db);
      const svxg_material& Ma = material(va ? va : vb);
      const svxg_material& Mb = material(vb ? vb : va);
      const f64 cond = std::max(1e-6, 1.0 - static_cast<f64>(std::max(da, db)) / 255.0);
      ft += cond * std::min(Ma.ft, Mb.ft);
      fb += cond * std::min(Ma.fb, Mb.fb);
      fc += cond * std::min(Ma.fc, Mb.fc);
      coh += cond * std::min(Ma.cohesion, Mb.cohesion);
      mu += std::min(Ma.friction, Mb.friction);
    }
    if (nf > 0) {
      const f64 w = 1.0 / static_cast<f64>(nf);
      B.sectioned = 1;
      B.ft = static_cast<f32>(ft * w);
      B.fb = static_cast<f32>(fb * w);
      B.fc = static_cast<f32>(fc * w);
      B.coh = static_cast<f32>(coh * w);
      B.mu = static_cast<f32>(mu * w);
    }
    out[k] = B;
  }
}

// The general case (SecAcc::finish, world_internal.hpp:120-127; finish_general, :250-362;
// section_strengths_general, :402-441): faces of a placed lattice and junction samples, each a small
// square (point, normal, area, side) in the bond's frame. A bond with no samples on an unplaced
// lattice is finish_lattice's: orc_bond_sections above on that lattice's voxels.
void orc_bond_sections_general(const svxg_node* nodes, int32_t n, const svxg_section_head_general* heads, int32_t m,
                               const svxg_section_face* faces, int32_t nfaces, const svxg_junction_sample* samples,
                               int32_t, const svxg_lattice* lat, int32_t, int32_t S, const svxg_material* mats,
                               int32_t nmats, svxg_bond* out) {
  auto comp = [](V3& a, int i) -> f64& { return i == 0 ? a.x : (i == 1 ? a.y : a.z); };
  auto add = [](const V3& a, const V3& b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; };
  auto scale = [](const V3& a, f64 k) { return V3{a.x * k, a.y * k, a.z * k}; };
  auto unit = [&](const V3& a) {
    const f64 l = std::sqrt(vdot(a, a));
    return l > 0 ? scale(a, 1.0 / l) : V3{0, 0, 0};
  };
  // LatticeXf::to / dir_to (rigid.hpp:91-93): R s row by row, then the offset
  auto rot = [](const svxg_lattice& L, const V3& s) {
    const double* R = L.R;
    return V3{R[0] * s.x + R[1] * s.y + R[2] * s.z, R[3] * s.x + R[4] * s.y + R[5] * s.z,
              R[6] * s.x + R[7] * s.y + R[8] * s.z};
  };
  auto to = [&](const svxg_lattice& L, const V3& s) {
    if (L.identity) return s;
    const V3 r = rot(L, s);
    return V3{L.off[0] + r.x, L.off[1] + r.y, L.off[2] + r.z};
  };
  auto dir_to = [&](const svxg_lattice& L, const V3& d) { return L.identity ? d : rot(L, d); };
  auto voxel = [](const svxg_lattice& L, const int32_t p[3], uint8_t* v, uint8_t* dmg) {
    i64 idx = 0, mul = 1;
    for (int c = 0; c < 3; ++c) {
      const i32 r = p[c] - L.lo[c];
      if (r < 0 || r >= L.dim[c]) {
        *v = 0;
        *dmg = 0;
        return;
      }
      idx += mul * r;
      mul *= L.dim[c];
    }
    *v = L.vox[idx];
    *dmg = L.damage ? L.damage[idx] : 0;
  };
  auto material = [&](uint8_t v) -> const svxg_material& {
    const int id = (v & 0x7F) - 1;
    return mats[(id >= 0 && id < nmats) ? id : 0];
  };
  for (i32 k = 0; k < m; ++k) {
    const svxg_section_head_general& hd = heads[k];
    const svxg_lattice& G = lat[hd.grid];
    const i32 nf = hd.nfaces, nj = hd.nsamples;
    const svxg_section_face* F = faces + hd.face0;
    const svxg_junction_sample* J = samples + hd.sample0;
    svxg_bond B{};
    if (nj == 0 && G.identity) {
      // finish_lattice and section_strengths on the faces' own lattice
      svxg_section_head one{};
      one.a = hd.a;
      one.b = hd.b;
      one.axis = hd.axis;
      one.sign = hd.sign;
      one.face0 = 0;
      one.nfaces = nf;
      one.ma = hd.ma;
      one.mb = hd.mb;
      one.strength_a = hd.strength_a;
      one.strength_b = hd.strength_b;
      svxg_voxel_box box{};
      for (int c = 0; c < 3; ++c) {
        box.lo[c] = G.lo[c];
        box.dim[c] = G.dim[c];
      }
      box.h = G.h;
      box.vox = G.vox;
      box.damage = G.damage;
      orc_bond_sections(nodes, n, &one, 1, F, nf, &box, mats, nmats, &B);
      B.tag = k;
      out[k] = B;
      continue;
    }
    // each face and sample as a square: centre, normal (from a to b), area, side
    const size_t ne = size_t(nf) + size_t(nj);
    std::vector<V3> xs(ne), nk(ne);
    std::vector<f64> ar(ne), sd(ne);
    const f64 h = G.h, A1 = h * h;
    for (i32 f = 0; f < nf; ++f) {
      V3 x{h * F[f].p[0], h * F[f].p[1], h * F[f].p[2]};
      comp(x, F[f].axis) += 0.5 * h;
      xs[size_t(f)] = to(G, x);
      V3 nn{0, 0, 0};
      comp(nn, F[f].axis) = F[f].sign;
      nk[size_t(f)] = dir_to(G, nn);
      ar[size_t(f)] = A1;
      sd[size_t(f)] = h;
    }
    for (i32 q = 0; q < nj; ++q) {
      const svxg_junction_sample& js = J[q];
      const svxg_lattice& X = lat[js.vg];
      const f64 hs = X.h, sj = hs / static_cast<f64>(std::max(1, S));
      // the sample's point: the centre of cell (i, j) of an S x S split of the face (push 0)
      const int ax = js.face >> 1, a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
      const f64 sg = (js.face & 1) ? 1.0 : -1.0;
      const int i = js.sub / S, j = js.sub % S;
      V3 pt{hs * js.v[0], hs * js.v[1], hs * js.v[2]};
      comp(pt, ax) += sg * (0.5 * hs + 0.0);
      comp(pt, a1) += ((i + 0.5) / static_cast<f64>(S) - 0.5) * hs;
      comp(pt, a2) += ((j + 0.5) / static_cast<f64>(S) - 0.5) * hs;
      const size_t e = size_t(nf) + size_t(q);
      xs[e] = to(X, pt);
      V3 fn{0, 0, 0};
      comp(fn, ax) = (js.face & 1) ? 1.0 : -1.0;
      nk[e] = scale(dir_to(X, fn), static_cast<f64>(js.sign));
      ar[e] = sj * sj * static_cast<f64>(js.weight);
      sd[e] = sj;
    }
    f64 atot = 0.0;
    for (size_t e = 0; e < ne; ++e) atot += ar[e];
    const bool pair = hd.b >= 0;
    const V3 ca = v3(nodes[hd.a].c);
    const V3 cb = pair ? v3(nodes[hd.b].c) : V3{0, 0, 0};
    // the interface normal: the area-weighted mean normal (a lattice support: its face axis)
    V3 nrm{0, 0, 0};
    if (pair || nj) {
      for (size_t e = 0; e < ne; ++e) nrm = add(nrm, scale(nk[e], ar[e]));
      const f64 tiny = 1e-12 * atot * atot;
      if (vdot(nrm, nrm) <= tiny && pair) nrm = vsub(cb, ca);
      if (vdot(nrm, nrm) <= tiny) nrm = V3{0, 0, 1};
      nrm = unit(nrm);
    } else {
      V3 e{0, 0, 0};
      comp(e, hd.axis) = hd.sign;
      nrm = dir_to(G, e);
    }
    const f64 ax = std::abs(nrm.x), ay = std::abs(nrm.y), az = std::abs(nrm.z);
    const V3 least = (ax <= ay && ax <= az) ? V3{1, 0, 0} : (ay <= az ? V3{0, 1, 0} : V3{0, 0, 1});
    V3 t1 = unit(vcross(nrm, least));
    V3 t2 = vcross(nrm, t1);
    std::vector<f64> ws(ne);
    f64 W = 0.0;
    V3 pc{0, 0, 0};
    for (size_t e = 0; e < ne; ++e) {
      ws[e] = ar[e] * std::max(0.05, std::abs(vdot(nrm, nk[e])));
      W += ws[e];
      pc = add(pc, scale(xs[e], ws[e]));
    }
    pc = scale(pc, 1.0 / W);
    {
      f64 s11 = 0.0, s22 = 0.0, s12 = 0.0;
      for (size_t e = 0; e < ne; ++e) {
        const V3 d = vsub(xs[e], pc);
        const f64 y1 = vdot(d, t1), y2 = vdot(d, t2);
        s11 += ws[e] * y1 * y1;
        s22 += ws[e] * y2 * y2;
        s12 += ws[e] * y1 * y2;
      }
      if (std::abs(s12) > 1e-9 * (s11 + s22)) {
        const f64 half = 0.5 * (s11 - s22);
        const f64 l1 = 0.5 * (s11 + s22) + std::sqrt(half * half + s12 * s12);
        f64 e1 = s12, e2 = l1 - s11;
        if (std::abs(l1 - s22) > std::abs(e2)) {
          e1 = l1 - s22;
          e2 = s12;
        }
        const f64 en = std::sqrt(e1 * e1 + e2 * e2);
        if (en > 0.0) {
          e1 /= en;
          e2 /= en;
          const V3 u1 = add(scale(t1, e1), scale(t2, e2));
          const V3 u2 = vsub(scale(t2, e1), scale(t1, e2));
          t1 = u1;
          t2 = u2;
        }
      }
    }
    B.a = hd.a;
    B.b = hd.b;
    B.ma = hd.ma;
    B.mb = hd.mb;
    B.faces = static_cast<i32>(ne);
    B.tag = k;
    for (int c = 0; c < 3; ++c) {
      B.n[c] = vget(nrm, c);
      B.t1[c] = vget(t1, c);
      B.t2[c] = vget(t2, c);
      B.p[c] = vget(pc, c);
    }
    B.area = std::max(W, 0.25 * atot);
    f64 s1 = 0.0, s2 = 0.0, m1 = 0.0, m2 = 0.0, own = 0.0;
    for (size_t e = 0; e < ne; ++e) {
      const V3 d = vsub(xs[e], pc);
      const f64 y1 = vdot(d, t1), y2 = vdot(d, t2);
      s1 += ws[e] * y1 * y1;
      s2 += ws[e] * y2 * y2;
      m1 = std::max(m1, std::abs(y1) + 0.5 * sd[e]);
      m2 = std::max(m2, std::abs(y2) + 0.5 * sd[e]);
      own += ws[e] * sd[e] * sd[e];
    }
    own /= W;
    B.s1 = s1 + B.area * own / 12.0;
    B.s2 = s2 + B.area * own / 12.0;
    B.c1 = m1;
    B.c2 = m2;
    B.rmax = std::sqrt(B.c1 * B.c1 + B.c2 * B.c2);
    if (pair) {
      const V3 d = vsub(cb, ca);
      const f64 L = std::max({std::abs(vdot(d, nrm)), 0.5 * std::sqrt(vdot(d, d)), 0.5 * h});
      B.la = std::clamp(vdot(vsub(pc, ca), nrm), 0.1 * L, 0.9 * L);
      B.lb = L - B.la;
    } else {
      B.la = std::max(0.25 * h, vdot(vsub(pc, ca), nrm));
      B.lb = 0.5 * h;
    }
    B.strength = std::min(hd.strength_a, hd.strength_b);
    // section_strengths_general: a face weighs 1, a sample 1 / S^2 (an S x S split of a face); each
    // the weaker material of its two voxels times the more damaged one's condition; the mean
    f64 ft = 0, fb = 0, fc = 0, coh = 0, mu = 0, wsum = 0;
    auto one = [&](uint8_t va, uint8_t da, uint8_t vb, uint8_t db, f64 w) {
      const svxg_material& Ma = material(va ? va : vb);
      const svxg_material& Mb = material(vb ? vb : va);
      const f64 cond = std::max(1e-6, 1.0 - static_cast<f64>(std::max(da, db)) / 255.0);
      ft += w * cond * std::min(Ma.ft, Mb.ft);
      fb += w * cond * std::min(Ma.fb, Mb.fb);
      fc += w * cond * std::min(Ma.fc, Mb.fc);
      coh += w * cond * std::min(Ma.cohesion, Mb.cohesion);
      mu += w * std::min(Ma.friction, Mb.friction);
      wsum += w;
    };
    const f64 sw = 1.0 / static_cast<f64>(std::max(1, S * S));
    for (i32 f = 0; f < nf; ++f) {
      uint8_t va, da, vb, db;
      voxel(G, F[f].p, &va, &da);
      int32_t q[3] = {F[f].p[0], F[f].p[1], F[f].p[2]};
      q[F[f].axis] += 1;
      voxel(G, q, &vb, &db);
      one(va, da, vb, db, 1.0);
    }
    for (i32 q = 0; q < nj; ++q) {
      uint8_t va, da, vb, db;
      voxel(lat[J[q].vg], J[q].v, &va, &da);
      if (J[q].kind == 1) {
        vb = va;
        db = da;
      } else {
        voxel(lat[J[q].og], J[q].o, &vb, &db);
      }
      one(va, da, vb, db, sw * static_cast<f64>(1.0f));
    }
    if (ne > 0 && wsum > 0.0) {
      const f64 w = 1.0 / wsum;
      B.sectioned = 1;
      B.ft = static_cast<f32>(ft * w);
      B.fb = static_cast<f32>(fb * w);
      B.fc = static_cast<f32>(fc * w);
      B.coh = static_cast<f32>(coh * w);
      B.mu = static_cast<f32>(mu * w);
    }
    out[k] = B;
  }
}

int32_t orc_graph_components(int32_t n, const svxg_bond* bonds, int32_t m, const uint8_t* seed, int32_t* comp) {
  // adjacency of the intact bonds between nodes, then floods: the seeds first (component 0), then
  // every node not reached yet, in index order
  std::vector<std::vector<i32>> adj(static_cast<size_t>(n));
  for (i32 k = 0; k < m; ++k) {
    const svxg_bond& b = bonds[k];
    if (b.broken || b.b < 0) continue;
    adj[size_t(b.a)].push_back(b.b);
    adj[size_t(b.b)].push_back(b.a);
  }
  std::fill(comp, comp + n, -1);
  std::vector<i32> todo;
  auto flood = [&](i32 c) {
    while (!todo.empty()) {
      const i32 k = todo.back();
      todo.pop_back();
      for (i32 j : adj[size_t(k)])
        if (comp[j] < 0) {
          comp[j] = c;
          todo.push_back(j);
        }
    }
  };
  for (i32 i = 0; i < n; ++i)
    if (seed[i] && comp[i] < 0) {
      comp[i] = 0;
      todo.push_back(i);
    }
  flood(0);
  i32 nc = 1;
  for (i32 i = 0; i < n; ++i) {
    if (comp[i] >= 0) continue;
    comp[i] = nc;
    todo.push_back(i);
    flood(nc++);
  }
  return nc;
}

void orc_info(orc_problem* h, svxg_problem_info* out) {
  *out = svxg_problem_info{};
  out->nodes = static_cast<int32_t>(h->nodes.size());
  out->bonds = static_cast<int32_t>(h->bonds.size());
  out->nfree = h->nfree;
  out->block_jacobi = h->jacobi_only;
  out->assembled = h->assembled;
  out->preconditioner_current = h->assembled && h->pc_current;
  out->blocks = h->K.blocks();
  out->work_per_iteration = 2 * h->K.blocks() + h->amg.work_per_apply();
  out->assembly_work = h->assembly_work;
  out->levels = static_cast<int32_t>(h->amg.levels());
  for (size_t l = 0; l < h->amg.levels() && l < 16; ++l) {
    out->level_rows[l] = h->amg.level(l).A.n;
    out->level_blocks[l] = h->amg.level(l).A.blocks();
    out->level_pblocks[l] = static_cast<int64_t>(h->amg.level(l).Pcol.size());
  }
}

void orc_get_matrix(orc_problem* h, int32_t* rowptr, int32_t* col, double* val) {
  if (rowptr) std::copy(h->K.rowptr.begin(), h->K.rowptr.end(), rowptr);
  if (col) std::copy(h->K.col.begin(), h->K.col.end(), col);
  if (val) std::copy(h->K.val.begin(), h->K.val.end(), val);
}

void orc_apply_matrix(orc_problem* h, const double* x, double* y) { h->K.apply(x, y); }
void orc_apply_preconditioner(orc_problem* h, const double* r, double* z) { h->precondition(r, z); }
void orc_get_bonds(orc_problem* h, svxg_bond* out) { std::copy(h->bonds.begin(), h->bonds.end(), out); }
void orc_get_nodes(orc_problem* h, svxg_node* out) { std::copy(h->nodes.begin(), h->nodes.end(), out); }
void orc_set_threads(int32_t) {}

// StressOptions{} / AmgOptions{} defaults (stress.hpp:68-72, amg.hpp:46-69).
int32_t orc_inv6(const double* A, double* out) { return inv6(A, out) ? 1 : 0; }

void orc_default_options(svxg_stress_options* o) {
  o->rtol = 2e-3;
  o->amg_min_nodes = 160;
  o->precision = SVXG_FP64;
  o->amg.strength = 0.1;
  o->amg.coarse_scale = 0.6;
  o->amg.coarse_max = 48;
  o->amg.smoothed = 1;
  o->amg.max_levels = 16;
  o->amg.sweeps = 1;
  o->amg.cheb_degree = 3;
  o->amg.coarsen_dense = 1;
  o->amg.reaggregate = 1;
}

}  // extern "C"

// ---- a structure's extraction: the walk from a seed ---------------------------------------------
// World::Impl::extract's walk (world.cpp:425-620) and append_nodes (world.cpp:739-865) for the world
// grid alone, over a box of voxels: the sequential walk as the reference takes it - a queue of
// members, each member's voxels in its chunk's scan order, a voxel's six faces in turn - with its
// SectionLog (world_internal.hpp:533-620) and cluster_items (:466-491).
namespace {

constexpr i32 kChunkEdge = 32;  // core/include/svx/world/grid.hpp:41

// SplitMix64's output function (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014), as world_internal.hpp:17-23 uses it
uint64_t orc_mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
uint64_t orc_key3(i64 x, i64 y, i64 z) {  // grid.hpp:56-59
  const i64 off = i64(1) << 20;
  return (uint64_t(x + off) << 42) | (uint64_t(y + off) << 21) | uint64_t(z + off);
}
i32 orc_fdiv(i32 a, i32 b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
uint64_t orc_acc_key(i32 a, i32 b, int axis, int sign) {  // world_internal.hpp:447-450
  if (b >= 0) return (uint64_t(std::min(a, b)) << 32) | (uint64_t(std::max(a, b)) << 1) | 1u;
  return (uint64_t(a) << 32) | (uint64_t(axis * 2 + (sign > 0 ? 1 : 0)) << 1);
}

// keys to dense indices in the order they first came (an open-addressed table, as
// world_internal.hpp:493-532's KeyIndex does it)
struct WalkKeys {
  std::vector<uint64_t> keys;
  std::vector<i32> table;
  i32 operator()(uint64_t k) {
    if (keys.size() * 2 >= table.size()) {
      std::vector<i32> t(std::max<size_t>(1024, table.size() * 2), -1);
      for (size_t e = 0; e < keys.size(); ++e) {
        size_t i = size_t(orc_mix64(keys[e])) & (t.size() - 1);
        while (t[i] >= 0) i = (i + 1) & (t.size() - 1);
        t[i] = i32(e);
      }
      table.swap(t);
    }
    size_t i = size_t(orc_mix64(k)) & (table.size() - 1);
    while (table[i] >= 0) {
      if (keys[size_t(table[i])] == k) return table[i];
      i = (i + 1) & (table.size() - 1);
    }
    table[i] = i32(keys.size());
    keys.push_back(k);
    return table[i];
  }
};

struct WalkHead {
  i32 a = -1, b = -1;
  int axis = 0;
  int sign = 1;
  uint8_t mb = 0;
};
struct WalkFace {
  i32 p[3];
  i32 s;
  int axis;
  int sign;
};

// SectionLog: fine bonds in the order made, faces in the order met
struct WalkLog {
  std::vector<WalkHead> heads;
  std::vector<WalkFace> faces;
  WalkKeys index;
  i32 at(i32 a, i32 b, int axis, int sign) {
    const i32 s = index(orc_acc_key(a, b, axis, sign));
    if (s < i32(heads.size())) return s;
    WalkHead H;
    H.a = b >= 0 ? std::min(a, b) : a;
    H.b = b >= 0 ? std::max(a, b) : b;
    H.axis = axis;
    H.sign = sign;
    heads.push_back(H);
    return s;
  }
  void face(i32 s, const i32 lower[3], int axis, int sign_from_a) {
    WalkFace F;
    for (int c = 0; c < 3; ++c) F.p[c] = lower[c];
    F.s = s;
    F.axis = axis;
    F.sign = sign_from_a;
    faces.push_back(F);
  }
};

}  // namespace

extern "C" int32_t orc_extract_structure(const svxg_voxel_box* box, const uint8_t* broken, const int32_t* frag,
                                         const svxg_fragment* frags, int32_t nfrag, int32_t seed,
                                         const svxg_extract_options* opt, const svxg_material* mats, int32_t nmats,
                                         svxg_extract_result* res, int32_t* members_out, int32_t* member_node,
                                         svxg_node* nodes_out, uint64_t* node_ident, svxg_section_head* heads_out,
                                         svxg_section_face* faces_out, svxg_bond* bonds_out, uint64_t* bond_ident) {
  const i32 nx = box->dim[0], ny = box->dim[1], nz = box->dim[2];
  const i64 nvox = i64(nx) * ny * nz;
  auto vidx = [&](const i32 g[3], i64* v) {
    const i32 x = g[0] - box->lo[0], y = g[1] - box->lo[1], z = g[2] - box->lo[2];
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return false;
    *v = x + i64(nx) * (y + i64(ny) * z);
    return true;
  };
  // each fragment's voxels in its chunk's scan order (x slowest: local index x * 1024 + y * 32 + z)
  std::vector<i32> vstart(size_t(nfrag) + 1, 0);
  for (i64 v = 0; v < nvox; ++v)
    if (frag[v] >= 0) ++vstart[size_t(frag[v]) + 1];
  for (i32 f = 0; f < nfrag; ++f) vstart[size_t(f) + 1] += vstart[size_t(f)];
  std::vector<i64> vlist(size_t(vstart[size_t(nfrag)]));
  {
    std::vector<i32> fill(vstart.begin(), vstart.end() - 1);
    for (i32 x = 0; x < nx; ++x)
      for (i32 y = 0; y < ny; ++y)
        for (i32 z = 0; z < nz; ++z) {
          const i64 v = x + i64(nx) * (y + i64(ny) * z);
          if (frag[v] >= 0) vlist[size_t(fill[size_t(frag[v])]++)] = v;
        }
  }
  auto mat_of = [](uint8_t v) { return uint8_t((v & 0x7F) - 1); };
  // the walk (world.cpp:475-620)
  std::vector<i32> slot(static_cast<size_t>(nfrag), -1), members;
  const V3 seed_pos = v3(frags[seed].com);
  slot[size_t(seed)] = 0;
  members.push_back(seed);
  bool truncated = false, any_support = false, touched_unloaded = false;
  auto meet = [&](i32 G) -> i32 {
    i32& s = slot[size_t(G)];
    if (s == -1) {
      const V3 d = vsub(v3(frags[G].com), seed_pos);
      const i64 other = frags[G].owner;
      const bool owned = other != 0;
      const bool fresh = owned && other >= opt->fresh_from;
      const bool room = i32(members.size()) < opt->max_nodes;
      if (!fresh && room && (owned || std::sqrt(vdot(d, d)) <= opt->max_radius)) {
        s = i32(members.size());
        members.push_back(G);
      } else {
        s = -2;
        truncated = true;
      }
    }
    return s;
  };
  WalkLog fine;
  for (size_t qi = 0; qi < members.size(); ++qi) {
    const i32 F = members[qi];
    const i32 nF = i32(qi);
    for (i32 k = vstart[size_t(F)]; k < vstart[size_t(F) + 1]; ++k) {
      const i64 v = vlist[size_t(k)];
      const i32 p[3] = {i32(v % nx) + box->lo[0], i32((v / nx) % ny) + box->lo[1], i32(v / (i64(nx) * ny)) + box->lo[2]};
      const uint8_t brk_p = broken ? broken[v] : 0;
      for (int a = 0; a < 3; ++a)
        for (int sg = -1; sg <= 1; sg += 2) {
          i32 q[3] = {p[0], p[1], p[2]};
          q[a] += sg;
          i64 w = -1;
          const bool in = vidx(q, &w);
          if (in && frag[w] == F) continue;
          const uint8_t vq = in ? box->vox[w] : 0;
          const bool unloaded = !in && opt->outside == 1;
          if (!unloaded) {
            if (vq == 0) continue;
            const bool face_broken = sg > 0 ? ((brk_p >> a) & 1) != 0 : broken && ((broken[w] >> a) & 1) != 0;
            if (face_broken) continue;
          }
          const i32* lower = sg > 0 ? p : q;
          if (unloaded) touched_unloaded = true;
          if ((vq & 0x80) || unloaded) {
            const i32 A = fine.at(nF, -1, a, sg);
            fine.heads[size_t(A)].mb = unloaded ? mat_of(box->vox[v]) : mat_of(vq);
            fine.face(A, lower, a, 1);
            any_support = true;
            continue;
          }
          const i32 g2 = frag[w];
          if (g2 < 0) continue;
          const i32 s = meet(g2);
          if (s >= 0) {
            if (sg > 0) fine.face(fine.at(nF, s, a, 1), lower, a, nF < s ? 1 : -1);
          } else {
            const i32 A = fine.at(nF, -1, a, sg);
            fine.heads[size_t(A)].mb = mat_of(vq);
            fine.face(A, lower, a, 1);
            any_support = true;
          }
        }
    }
  }
  // append_nodes: clusters (cluster_cell, world.cpp:733-737; cluster_items)
  const i32 m = i32(members.size());
  const i32 cell = m <= opt->cluster_nodes ? 0 : (i64(m) <= 6 * i64(opt->cluster_nodes) ? 8 : 16);
  std::vector<uint64_t> key(static_cast<size_t>(m));
  for (i32 f = 0; f < m; ++f) {
    if (cell <= 0) {
      key[size_t(f)] = uint64_t(f);
      continue;
    }
    const svxg_fragment& fi = frags[members[size_t(f)]];
    i32 cc[3], l[3];
    for (int c = 0; c < 3; ++c) {
      cc[c] = orc_fdiv(fi.first[c], kChunkEdge);
      l[c] = fi.first[c] - cc[c] * kChunkEdge;
    }
    const uint64_t c = opt->cluster_cubes
                           ? uint64_t(((l[0] / cell) << 4) | ((l[1] / cell) << 2) | (l[2] / cell))
                           : uint64_t((l[0] / cell) * 64 + (l[1] / cell) * 8 + l[2] / cell);
    key[size_t(f)] = orc_mix64(orc_key3(cc[0], cc[1], cc[2]) ^ (c << 58));
  }
  std::vector<i32> par(static_cast<size_t>(m));
  for (i32 i = 0; i < m; ++i) par[size_t(i)] = i;
  auto find = [&](i32 x) {
    while (par[size_t(x)] != x) {
      par[size_t(x)] = par[size_t(par[size_t(x)])];
      x = par[size_t(x)];
    }
    return x;
  };
  for (const WalkHead& H : fine.heads) {
    if (H.a < 0 || H.a >= m || H.b < 0 || H.b >= m) continue;
    if (key[size_t(H.a)] != key[size_t(H.b)]) continue;
    const i32 ra = find(H.a), rb = find(H.b);
    if (ra != rb) par[size_t(std::max(ra, rb))] = std::min(ra, rb);
  }
  std::vector<i32> fnode(static_cast<size_t>(m), -1), id(size_t(m), -1);
  i32 K = 0;
  for (i32 i = 0; i < m; ++i) {
    const i32 r = find(i);
    if (id[size_t(r)] < 0) id[size_t(r)] = K++;
    fnode[size_t(i)] = id[size_t(r)];
  }
  std::vector<std::vector<i32>> cm(static_cast<size_t>(K));
  for (i32 f = 0; f < m; ++f) cm[size_t(fnode[size_t(f)])].push_back(f);
  std::vector<svxg_node> nodes(static_cast<size_t>(K));
  std::vector<uint8_t> nmat(static_cast<size_t>(K));
  std::vector<f64> nstrength(static_cast<size_t>(K));
  std::vector<uint64_t> ident(static_cast<size_t>(K));
  for (i32 c = 0; c < K; ++c) {
    f64 mass = 0.0, heavy = -1.0, strength = 1e30;
    V3 com;
    uint8_t mat = 0;
    for (i32 f : cm[size_t(c)]) {
      const svxg_fragment& fi = frags[members[size_t(f)]];
      mass += fi.mass;
      com.x += fi.com[0] * fi.mass;
      com.y += fi.com[1] * fi.mass;
      com.z += fi.com[2] * fi.mass;
      if (fi.mass > heavy) {
        heavy = fi.mass;
        mat = fi.mat;
      }
      strength = std::min(strength, fi.strength);
    }
    svxg_node nd{};
    if (mass > 0) {
      const f64 inv = 1.0 / mass;
      nd.c[0] = com.x * inv;
      nd.c[1] = com.y * inv;
      nd.c[2] = com.z * inv;
    } else {
      nd.c[0] = com.x;
      nd.c[1] = com.y;
      nd.c[2] = com.z;
    }
    nd.mass = mass;
    nodes[size_t(c)] = nd;
    nmat[size_t(c)] = mat;
    nstrength[size_t(c)] = strength;
    const svxg_fragment& f0 = frags[members[size_t(cm[size_t(c)].front())]];
    i32 cc[3], l[3];
    for (int q = 0; q < 3; ++q) {
      cc[q] = orc_fdiv(f0.first[q], kChunkEdge);
      l[q] = f0.first[q] - cc[q] * kChunkEdge;
    }
    const i32 first = l[0] * kChunkEdge * kChunkEdge + l[1] * kChunkEdge + l[2];
    ident[size_t(c)] = orc_mix64(orc_key3(cc[0], cc[1], cc[2]) * 0x9E3779B97F4A7C15ull ^ uint64_t(first));
  }
  // SectionLog::merge: node-level bonds in the order their first fine bond was made
  struct Merged {
    WalkHead H;
    std::vector<WalkFace> faces;
  };
  std::vector<Merged> out;
  std::vector<i32> to(fine.heads.size(), -1);
  std::vector<uint8_t> flip(fine.heads.size(), 0);
  {
    WalkKeys mindex;
    for (size_t f = 0; f < fine.heads.size(); ++f) {
      const WalkHead& F = fine.heads[f];
      const i32 a = fnode[size_t(F.a)];
      const i32 b = F.b >= 0 ? fnode[size_t(F.b)] : -1;
      if (b >= 0 && a == b) continue;
      const i32 mi = mindex(orc_acc_key(a, b, F.axis, F.sign));
      if (mi == i32(out.size())) {
        Merged M;
        M.H = F;
        M.H.a = b >= 0 ? std::min(a, b) : a;
        M.H.b = b >= 0 ? std::max(a, b) : b;
        out.push_back(std::move(M));
      }
      to[f] = mi;
      flip[f] = b >= 0 && a != out[size_t(mi)].H.a;
    }
  }
  {
    std::vector<std::vector<i32>> per(fine.heads.size());
    for (size_t e = 0; e < fine.faces.size(); ++e) per[size_t(fine.faces[e].s)].push_back(i32(e));
    for (size_t f = 0; f < fine.heads.size(); ++f) {
      if (to[f] < 0) continue;
      Merged& M = out[size_t(to[f])];
      for (i32 e : per[f]) {
        WalkFace F = fine.faces[size_t(e)];
        F.sign = (flip[f] ? -1 : 1) * F.sign;
        M.faces.push_back(F);
      }
    }
  }
  // the bonds (world.cpp:824-852): ends in the order of their identities, materials, strengths
  const i32 nb = i32(out.size());
  std::vector<svxg_section_head> heads(static_cast<size_t>(nb));
  std::vector<svxg_section_face> faces;
  std::vector<uint64_t> bid(static_cast<size_t>(nb));
  for (i32 k = 0; k < nb; ++k) {
    Merged& M = out[size_t(k)];
    if (M.H.b >= 0 && ident[size_t(M.H.a)] > ident[size_t(M.H.b)]) {
      std::swap(M.H.a, M.H.b);
      for (WalkFace& F : M.faces) F.sign = -F.sign;
    }
    svxg_section_head& h = heads[size_t(k)];
    std::memset(&h, 0, sizeof h);
    h.a = M.H.a;
    h.b = M.H.b;
    h.axis = M.H.axis;
    h.sign = M.H.sign;
    h.face0 = i32(faces.size());
    h.nfaces = i32(M.faces.size());
    h.ma = nmat[size_t(M.H.a)];
    h.strength_a = nstrength[size_t(M.H.a)];
    if (M.H.b >= 0) {
      h.mb = nmat[size_t(M.H.b)];
      h.strength_b = nstrength[size_t(M.H.b)];
    } else {
      h.mb = M.H.mb;
      h.strength_b = 1e9;
    }
    for (const WalkFace& F : M.faces) {
      svxg_section_face o;
      std::memset(&o, 0, sizeof o);
      for (int c = 0; c < 3; ++c) o.p[c] = F.p[c];
      o.axis = int8_t(F.axis);
      o.sign = int8_t(F.sign);
      faces.push_back(o);
    }
    const uint64_t ia = ident[size_t(M.H.a)], ib = M.H.b >= 0 ? ident[size_t(M.H.b)] : 0;
    const int ax = M.H.b >= 0 ? 0 : M.H.axis, sn = M.H.b >= 0 ? 1 : M.H.sign;
    bid[size_t(k)] = orc_mix64(ia * 0xD6E8FEB86659FD93ull ^ orc_mix64(ib + uint64_t(ax * 2 + (sn > 0 ? 1 : 0))));
  }
  const i32 nf = i32(faces.size());
  svxg_extract_result r{};
  r.members = m;
  r.nodes = K;
  r.bonds = nb;
  r.faces = nf;
  r.cell = cell;
  r.truncated = truncated ? 1 : 0;
  r.any_support = any_support ? 1 : 0;
  r.touched_unloaded = touched_unloaded ? 1 : 0;
  const bool want = members_out || member_node || nodes_out || node_ident || heads_out || faces_out || bonds_out || bond_ident;
  const bool short_room = want && (m > res->members || K > res->nodes || nb > res->bonds || nf > res->faces);
  *res = r;
  if (short_room) return SVXG_BAD_INPUT;
  if (members_out) std::copy(members.begin(), members.end(), members_out);
  if (member_node) std::copy(fnode.begin(), fnode.end(), member_node);
  if (nodes_out) std::copy(nodes.begin(), nodes.end(), nodes_out);
  if (node_ident) std::copy(ident.begin(), ident.end(), node_ident);
  if (heads_out) std::copy(heads.begin(), heads.end(), heads_out);
  if (faces_out) std::copy(faces.begin(), faces.end(), faces_out);
  if (bond_ident) std::copy(bid.begin(), bid.end(), bond_ident);
  if (bonds_out && nb > 0)
    orc_bond_sections(nodes.data(), K, heads.data(), nb, faces.data(), nf, box, mats, nmats, bonds_out);
  return SVXG_OK;
}

extern "C" uint64_t orc_mix64_value(uint64_t x) { return orc_mix64(x); }
