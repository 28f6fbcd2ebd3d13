// Setup-phase kernels of the svxg stress solve, and the bond checks.
//
// Assembly follows StressProblem::build_matrix / bond_matrices (core/src/stress/stress.cpp:59-132)
// and blk6::atbd_add / rigid_block (core/src/solve/amg.cpp:143-170) in the same floating-point
// operation order (this file is compiled with -ffp-contract=off), and sums each block's
// contributions in bond order as Bsr6Builder does: K is bit-identical to the reference's.
//
// The multigrid setup is the reference's smoothed aggregation (amg.cpp:228-533) made parallel:
//   strength        ||A_ij|| >= theta sqrt(||A_ii|| ||A_jj||) (symmetrised: max of ij and ji)
//   aggregation     roots are a distance-2 maximal independent set of the strong graph (the
//                   parallel form of the reference's greedy pass 1, which also keeps roots three
//                   links apart); members join their root; pass 2 joins the strongest aggregated
//                   neighbour (amg.cpp:350-366); pass 3 makes singletons and groups uncoupled rows
//                   eight at a time (amg.cpp:367-384); theta halves up to three times while
//                   2 na > n (amg.cpp:385-388)
//   prolongator     P = (I - w D^-1 A) P_rigid, w = 4/3 / lambda_max (amg.cpp:391-452), the
//                   reference's block product order
//   Galerkin        A P by fine row, then P^T (A P) by coarse row, each block summed in the
//                   reference's order (amg.cpp:472-529)
//   coarsest        dense Cholesky with the reference's symmetrisation and pivot floor
//                   (amg.cpp:535-578), then its explicit inverse for a parallel apply
// Bond checks: bond_load + bond_utilization (stress.cpp:14-57, 477-502), same operation order.
#include <hip/hip_fp16.h>

#include "bond_math.hpp"

namespace svxg {

namespace {

constexpr int kBondsPerWG = 7;  // 36 lanes a bond / a block

// LDS writes of a wavefront visible to its other lanes
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The lane forms of the prolongator and Galerkin fills take a wavefront an entry: its 64 lanes look
// at 64 of the entry's candidate terms at once (a coupling, a restriction entry: whether it meets
// the entry's column, by a lookup or a binary search each) and list the ones that do, in order, in
// LDS; then lanes 0..35, one a block element, add the listed terms in that order - each element's
// arithmetic that of the thread form, term for term. The terms' lookups overlap instead of following
// one another.
constexpr int kFillWaves = kWG / 64;
// the listed terms of a wavefront's chunk: up to 64, by rank; a = the term's index, b = its partner
// entry (or column)
struct FillList {
  int32_t a[kFillWaves][64];
  int32_t b[kFillWaves][64];
};
// lane `lane` lists (a, b) when `hit`, at its rank among the wave's hits; returns the hit count
__device__ __forceinline__ int fill_list(FillList& L, int wv, int lane, bool hit, int a, int b) {
  const unsigned long long mask = __ballot(hit);
  if (hit) {
    const int k = int(__popcll(mask & ((1ull << lane) - 1ull)));
    L.a[wv][k] = a;
    L.b[wv][k] = b;
  }
  wave_sync();
  return int(__popcll(mask));
}
// the first index in [lo, hi) of sorted cols holding c, or -1
__device__ __forceinline__ int find_sorted(const int32_t* __restrict__ cols, int lo, int hi, int c) {
  const int end = hi;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && cols[lo] == c) ? lo : -1;
}

// the bond's local stiffness in the global frame, element (R, C) (stress.cpp:59-79)
__device__ double bond_D(const svxg_bond& b, const svxg_material& A, const svxg_material& B, int R, int C) {
  const bool t = R < 3 && C < 3, rr = R >= 3 && C >= 3;
  if (!t && !rr) return 0.0;
  const double la = fmax(b.la, 1e-4), lb = fmax(b.lb, 1e-4);
  const double invE = la / A.E + lb / B.E;
  const double invG = la / A.G + lb / B.G;
  double k[3];
  if (t) {
    k[0] = b.area / invE;
    k[1] = (5.0 / 6.0) * b.area / invG;
    k[2] = (5.0 / 6.0) * b.area / invG;
  } else {
    k[0] = 0.8 * (b.s1 + b.s2) / invG;
    k[1] = b.s2 / invE;
    k[2] = b.s1 / invE;
  }
  const int r = t ? R : R - 3, c = t ? C : C - 3;
  const double* Q[3] = {b.n, b.t1, b.t2};
  double s = 0.0;
  for (int q = 0; q < 3; ++q) s += Q[q][r] * k[q] * Q[q][c];
  return s;
}

// element (R, C) of s * L^T (M X) (blk6::atbd_add, amg.cpp:155-170) from 6x6 blocks in LDS
__device__ __forceinline__ double atbd_elem(const double* L, const double* M, const double* X, int R, int C) {
  double T[6];
  for (int k = 0; k < 6; ++k) {
    double acc = 0.0;
    for (int j = 0; j < 6; ++j) acc += M[k * 6 + j] * X[j * 6 + C];
    T[k] = acc;
  }
  double acc = 0.0;
  for (int k = 0; k < 6; ++k) acc += L[k * 6 + R] * T[k];
  return acc;
}

// a bond's D, Ba, Bb into LDS (lane e of the bond's 36 computes element e of each)
__device__ void bond_blocks_lds(const svxg_bond& b, const double* nodec, const svxg_material* mats, int nmats, int e,
                                double* sD, double* sA, double* sB) {
  const svxg_material& MA = bond_mat(mats, nmats, b.ma);
  const svxg_material& MB = bond_mat(mats, nmats, b.mb);
  const int R = e / 6, C = e % 6;
  sD[e] = bond_D(b, MA, MB, R, C);
  const double* ca = nodec + 3 * size_t(b.a);
  const double ra[3] = {b.p[0] - ca[0], b.p[1] - ca[1], b.p[2] - ca[2]};
  sA[e] = rigid_elem(ra, R, C);
  if (b.b >= 0) {
    const double* cb = nodec + 3 * size_t(b.b);
    const double rb[3] = {b.p[0] - cb[0], b.p[1] - cb[1], b.p[2] - cb[2]};
    sB[e] = rigid_elem(rb, R, C);
  } else {
    sB[e] = 0.0;
  }
}

// Frobenius norm summed in an order symmetric under transposition (diagonal, then each pair
// (rc, cr) together): ||A|| and ||A^T|| come out bitwise equal
__device__ __forceinline__ double fro_sym(const double* M) {
  double s = 0.0;
  for (int r = 0; r < 6; ++r) s += M[r * 7] * M[r * 7];
  for (int r = 0; r < 6; ++r)
    for (int c = r + 1; c < 6; ++c) s += M[r * 6 + c] * M[r * 6 + c] + M[c * 6 + r] * M[c * 6 + r];
  return sqrt(s);
}

// position of column c in row r (diagonal first, then ascending), -1 if absent
__device__ __forceinline__ int find_block(const int32_t* rowptr, const int32_t* col, int r, int c) {
  int lo = rowptr[r], hi = rowptr[r + 1];
  if (lo >= hi) return -1;
  if (col[lo] == c) return lo;
  ++lo;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return (lo < rowptr[r + 1] && col[lo] == c) ? lo : -1;
}

__device__ __forceinline__ void sort_ints(int32_t* a, int n) {
  for (int i = 1; i < n; ++i) {
    const int32_t v = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > v) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = v;
  }
}

__device__ __forceinline__ int unique_sorted(int32_t* a, int n) {
  if (n == 0) return 0;
  int u = 1;
  for (int i = 1; i < n; ++i)
    if (a[i] != a[u - 1]) a[u++] = a[i];
  return u;
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

constexpr int32_t MIS_UND = 0, MIS_ROOT = 1, MIS_EXC = 2;

__device__ __forceinline__ uint64_t mis_tuple(int state, int i) {
  if (state == MIS_EXC) return 0ull;
  const uint64_t rank = state == MIS_ROOT ? 3ull : 2ull;
  const uint64_t pr = hash32(uint32_t(i) * 2654435761U + 12345U) & 0x7fffffffU;
  return (rank << 62) | (pr << 31) | uint64_t(uint32_t(i));
}

// exclusive scan of a[0..len) in place, a[len] = total, one workgroup
__device__ void block_scan(int32_t* a, int len, int32_t* total_out) {
  // four consecutive values a thread, a wavefront prefix sum of the threads' sums by shuffles, the
  // wavefronts' totals through LDS: three barriers a chunk of 4 kWG values
  __shared__ int32_t wsum[kWG / 64];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = int(threadIdx.x) & 63, w = int(threadIdx.x) >> 6;
  for (int base = 0; base < len; base += 4 * kWG) {
    const int i0 = base + 4 * int(threadIdx.x);
    int32_t v[4];
    for (int k = 0; k < 4; ++k) v[k] = i0 + k < len ? a[i0 + k] : 0;
    const int32_t mine = v[0] + v[1] + v[2] + v[3];
    int32_t incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int32_t run = carry + incl - mine;
    for (int q = 0; q < w; ++q) run += wsum[q];
    for (int k = 0; k < 4; ++k) {
      if (i0 + k < len) a[i0 + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == kWG - 1) carry = run;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a[len] = carry;
    if (total_out) *total_out = carry;
  }
}

}  // namespace

// ---- assembly -------------------------------------------------------------------------------------

__global__ void __launch_bounds__(kWG) k_bond_products(const AssemblyDev* __restrict__ J, const Tile* __restrict__ tiles,
                                                       int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const AssemblyDev& a = J[tl.g];
  __shared__ double sD[kBondsPerWG][36], sA[kBondsPerWG][36], sB[kBondsPerWG][36];
  const int lb = threadIdx.x / 36, e = threadIdx.x % 36;
  const bool on = lb < kBondsPerWG && lb < tl.nrows;
  const int bi = tl.row0 + (on ? lb : 0);
  const int32_t* sl = a.bond_slot + 4 * size_t(bi);
  const bool used = on && (sl[0] >= 0 || sl[1] >= 0 || sl[2] >= 0 || sl[3] >= 0);
  if (used) bond_blocks_lds(a.bonds[bi], a.nodec, a.mats, a.nmats, e, sD[lb], sA[lb], sB[lb]);
  __syncthreads();
  if (!used) return;
  const int R = e / 6, C = e % 6;
  double* out = a.prod + 144 * size_t(bi);
  if (sl[0] >= 0) out[e] = 1.0 * atbd_elem(sA[lb], sD[lb], sA[lb], R, C);
  if (sl[1] >= 0) out[36 + e] = 1.0 * atbd_elem(sB[lb], sD[lb], sB[lb], R, C);
  if (sl[2] >= 0) out[72 + e] = -1.0 * atbd_elem(sA[lb], sD[lb], sB[lb], R, C);
  if (sl[3] >= 0) out[108 + e] = -1.0 * atbd_elem(sB[lb], sD[lb], sA[lb], R, C);
}

// each block: the sum of its contributions in bond order, from zero (Bsr6Builder)
__global__ void __launch_bounds__(kWG) k_gather_blocks(const AssemblyDev* __restrict__ J, const Tile* __restrict__ tiles,
                                                       int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const AssemblyDev& a = J[tl.g];
  const int lb = threadIdx.x / 36, e = threadIdx.x % 36;
  if (lb >= kBondsPerWG || lb >= tl.nrows) return;
  const int k = tl.row0 + lb;
  double v = 0.0;
  for (int c = a.cptr[k]; c < a.cptr[k + 1]; ++c) {
    const int32_t ci = a.cidx[c];
    v += a.prod[144 * size_t(ci >> 2) + 36 * size_t(ci & 3) + e];
  }
  a.K[36 * size_t(k) + e] = v;
}

__global__ void k_diag_trace(const AssemblyDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const AssemblyDev& a = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  const double* d = a.K + 36 * size_t(a.rowptr[i]);
  a.trace[i] = d[0] + d[7] + d[14];
}

// id_scale = mean of tr(K_ii[0:3,0:3]) / 3 over rows with a positive trace, summed in row order
// (stress.cpp:111-122). One workgroup a graph: fifteen wavefronts load a chunk of 960 traces and
// divide them by 3 into LDS (rows without a positive trace give +0.0, which leaves the running sum's
// bits unchanged: it starts at +0.0 and only grows) while lane 0 of the first adds the previous
// chunk's in row order - the sequential sum, with its operands ready in LDS.
constexpr int kIdWG = 1024, kIdRows = kIdWG - 64;
__global__ void __launch_bounds__(kIdWG) k_id_scale(const AssemblyDev* __restrict__ J, int count) {
  const int g = blockIdx.x;
  if (g >= count) return;
  const AssemblyDev& a = J[g];
  const int t = int(threadIdx.x), w = t >> 6;
  __shared__ double buf[2][kIdRows];
  __shared__ int cnts[kIdWG / 64];
  const int nchunk = (a.nfree + kIdRows - 1) / kIdRows;
  double tr = 0.0;
  int cnt = 0;  // (a wavefront's count of positive traces)
  for (int c = 0; c <= nchunk; ++c) {
    if (w > 0) {
      if (c < nchunk) {
        const int i = c * kIdRows + (t - 64);
        const double tv = i < a.nfree ? a.trace[i] : 0.0;
        const bool pos = tv > 0;
        buf[c & 1][t - 64] = pos ? tv / 3.0 : 0.0;
        cnt += __popcll(__ballot(pos));
      }
    } else if (t == 0 && c > 0) {
      const double* b = buf[(c - 1) & 1];
      const int m = min(kIdRows, a.nfree - (c - 1) * kIdRows);
      for (int k = 0; k < m; ++k) tr += b[k];
    }
    __syncthreads();
  }
  if ((t & 63) == 0) cnts[w] = cnt;
  __syncthreads();
  if (t == 0) {
    int n = 0;
    for (int k = 1; k < kIdWG / 64; ++k) n += cnts[k];
    *a.id_scale = n ? tr / n : 1.0;
  }
}

__global__ void k_identity_rows(const AssemblyDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const AssemblyDev& a = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  double* d = a.K + 36 * size_t(a.rowptr[i]);
  const bool empty = d[0] == 0.0 && d[7] == 0.0 && d[14] == 0.0;
  if (a.row_gone[i] || empty) {
    const double s = *a.id_scale;
    for (int q = 0; q < 6; ++q) d[q * 7] = s;
  }
}

// Bonds leaving K in list order (StressProblem::remove_bond, stress.cpp:260-284), spread: a workgroup
// a removed bond forms its four products (aa, bb, ab, ba) into scratch; then a thread a (block,
// element) of K adds the products that land on its block in the removal list's order - each element
// of K gets the same additions in the same order as removing the bonds one after another, so the
// bits are the sequential removal's.
__global__ void __launch_bounds__(64) k_removal_products(const AssemblyDev* __restrict__ J, const int32_t* __restrict__ ids,
                                                         const int32_t* __restrict__ tjob, double* __restrict__ prod,
                                                         int nids) {
  const int t = blockIdx.x;
  if (t >= nids) return;
  const AssemblyDev& a = J[tjob[t]];
  __shared__ double sD[36], sA[36], sB[36];
  const int e = threadIdx.x;
  const int bi = ids[t];
  if (e < 36) bond_blocks_lds(a.bonds[bi], a.nodec, a.mats, a.nmats, e, sD, sA, sB);
  __syncthreads();
  if (e < 36) {
    const int R = e / 6, C = e % 6;
    double* o = prod + 144 * size_t(t);
    o[e] = atbd_elem(sA, sD, sA, R, C);
    o[36 + e] = atbd_elem(sB, sD, sB, R, C);
    o[72 + e] = atbd_elem(sA, sD, sB, R, C);
    o[108 + e] = atbd_elem(sB, sD, sA, R, C);
  }
}

// group g: block slot[g] of job gjob[g], its products cont[gstart[g] .. gstart[g + 1]) (4 t + which,
// in list order); aa and bb subtracted, ab and ba added back (stress.cpp:269-274)
__global__ void __launch_bounds__(kWG) k_removal_apply(const AssemblyDev* __restrict__ J, const int32_t* __restrict__ gjob,
                                                       const int32_t* __restrict__ gslot,
                                                       const int32_t* __restrict__ gstart,
                                                       const int32_t* __restrict__ cont, const double* __restrict__ prod,
                                                       int ngroups) {
  const int64_t t = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (t >= 36 * int64_t(ngroups)) return;
  const int g = int(t / 36), e = int(t % 36);
  double* K = J[gjob[g]].K + 36 * size_t(gslot[g]) + e;
  double v = *K;
  for (int k = gstart[g]; k < gstart[g + 1]; ++k) {
    const int c = cont[k];
    const double p = prod[36 * size_t(c) + e];
    v += ((c & 3) < 2 ? -1.0 : 1.0) * p;
  }
  *K = v;
}

__global__ void k_retire_rows(const AssemblyDev* __restrict__ J, const int32_t* __restrict__ rows, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 36 * count) return;
  const AssemblyDev& a = J[0];
  const int i = rows[t / 36], e = t % 36;
  a.K[36 * size_t(a.rowptr[i]) + e] = (e % 7 == 0) ? *a.id_scale : 0.0;
}

// The 6x6 Gauss-Jordan inverse with partial pivoting of blk6::inv6 (amg.cpp:105-135), a wavefront a
// block: W = [A | I] held in LDS, lane l owning elements l and l + 64 of W's 72 (row-major, 12 a
// row); every element sees the operations blk6::inv6 applies to it, in its order (pivot search, row
// swap, row scaling, elimination; rows whose multiplier is 0 left alone), so the inverse has its
// bits. Returns its verdict (false: zero matrix, or a pivot at most 1e-15 of the largest magnitude;
// the same on every lane); on success the lanes write the inverse to out.
__device__ bool inv6_wave(const double* A, double* W, int lane, double* __restrict__ out) {
  // (element x0 = lane of every lane, x1 = lane + 64 of lanes 0..7)
  const int x0 = lane, x1 = lane + 64;
  const bool two = x1 < 72;
  const int r0 = x0 / 12, k0 = x0 % 12, r1 = x1 / 12, k1 = x1 % 12;
  W[x0] = k0 < 6 ? A[r0 * 6 + k0] : (k0 - 6 == r0 ? 1.0 : 0.0);
  if (two) W[x1] = k1 < 6 ? A[r1 * 6 + k1] : (k1 - 6 == r1 ? 1.0 : 0.0);
  wave_sync();
  double big = 0.0;
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) big = fmax(big, fabs(W[r * 12 + c]));
  if (big == 0.0) return false;
  for (int c = 0; c < 6; ++c) {
    int best = c;
    for (int r = c + 1; r < 6; ++r)
      if (fabs(W[r * 12 + c]) > fabs(W[best * 12 + c])) best = r;
    if (fabs(W[best * 12 + c]) <= 1e-15 * big) return false;
    if (best != c) {
      const double v0 = r0 == c ? W[best * 12 + k0] : (r0 == best ? W[c * 12 + k0] : W[x0]);
      const double v1 = two ? (r1 == c ? W[best * 12 + k1] : (r1 == best ? W[c * 12 + k1] : W[x1])) : 0.0;
      wave_sync();
      W[x0] = v0;
      if (two) W[x1] = v1;
      wave_sync();
    }
    const double d = 1.0 / W[c * 12 + c];
    wave_sync();
    if (r0 == c) W[x0] *= d;
    if (two && r1 == c) W[x1] *= d;
    wave_sync();
    const double m0 = W[r0 * 12 + c], p0 = W[c * 12 + k0];
    const double m1 = two ? W[r1 * 12 + c] : 0.0, p1 = two ? W[c * 12 + k1] : 0.0;
    wave_sync();
    if (r0 != c && m0 != 0.0) W[x0] -= m0 * p0;
    if (two && r1 != c && m1 != 0.0) W[x1] -= m1 * p1;
    wave_sync();
  }
  if (k0 >= 6) out[r0 * 6 + (k0 - 6)] = W[x0];
  if (two && k1 >= 6) out[r1 * 6 + (k1 - 6)] = W[x1];
  return true;
}

// a level's inverse diagonal blocks (amg.cpp:227-244): a wavefront a block, four a workgroup, 64
// workgroups a row tile; a singular block is regularised by 1e-9 of its diagonal's absolute sum
// (mode 0) or replaced by the identity scale's inverse (mode 1: block-Jacobi's retired rows)
__global__ void __launch_bounds__(kWG) k_inv_diag(const DiagDev* __restrict__ J, const Tile* __restrict__ tiles,
                                                  int ntiles) {
  const int tile = int(blockIdx.x) >> 6, part = int(blockIdx.x) & 63;
  if (tile >= ntiles) return;
  const Tile tl = tiles[tile];
  const DiagDev& a = J[tl.g];
  const int w = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
  const int rl = 4 * part + w;
  if (rl >= tl.nrows) return;  // (a whole wavefront)
  const int i = tl.row0 + rl;
  __shared__ double Dl[kWG / 64][36], W[kWG / 64][72];
  if (lane < 36) Dl[w][lane] = a.A[36 * size_t(a.rowptr[i]) + size_t(lane)];
  wave_sync();
  double* out = a.Dinv + 36 * size_t(i);
  if (inv6_wave(Dl[w], W[w], lane, out)) return;
  if (a.mode == 1) {
    const double s = 1.0 / *a.id_scale;
    if (lane < 36) out[lane] = (lane % 7 == 0) ? s : 0.0;
    return;
  }
  double tr = 0.0;
  for (int q = 0; q < 6; ++q) tr += fabs(Dl[w][q * 7]);
  const double eps = tr > 0 ? 1e-9 * tr : 1.0;
  wave_sync();
  if (lane < 6) Dl[w][lane * 7] += eps;
  wave_sync();
  if (inv6_wave(Dl[w], W[w], lane, out)) return;
  if (lane < 36) out[lane] = 0.0;
  if (lane == 0) atomicAdd(a.fail, 1);
}

// ---- aggregation ------------------------------------------------------------------------------------

// block strengths (and the diagonal's)
__global__ void k_strength(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  for (int k = c.rowptr[i]; k < c.rowptr[i + 1]; ++k) c.s[k] = fro_sym(c.A + 36 * size_t(k));
}

// strong couplings (symmetric), initial MIS states (uncoupled rows: excluded, grouped later)
__global__ void k_strong(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  const int k0 = c.rowptr[i], k1 = c.rowptr[i + 1];
  const double di = c.s[k0];
  for (int k = k0; k < k1; ++k) {
    const int j = c.col[k];
    if (j == i) {
      c.strong[k] = 0;
      continue;
    }
    const int kt = find_block(c.rowptr, c.col, j, i);
    const double sk = kt >= 0 ? fmax(c.s[k], c.s[kt]) : c.s[k];
    const double dj = c.s[c.rowptr[j]];
    c.strong[k] = sk >= c.theta * sqrt(di * dj) ? 1 : 0;
  }
  c.state[i] = (k1 - k0 <= 1) ? MIS_EXC : MIS_UND;
}

// Round r counts its undecided rows into counts[mis_slot(r)], the two slots taken in turn: a graph
// whose previous round left nothing undecided skips the round (a round then changes nothing), its
// slot set to 0, so the host may queue many rounds between two looks
__device__ __forceinline__ int mis_slot(int round) { return (round & 1) ? 6 : 0; }
__device__ __forceinline__ bool mis_settled(const CoarsenDev& c, int round) {
  return round > 0 && c.counts[mis_slot(round - 1)] == 0;
}

__global__ void k_mis_t1(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles, int round) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  const bool settled = mis_settled(c, round);
  if (tl.row0 == 0 && threadIdx.x == 0) c.counts[mis_slot(round)] = 0;  // (this round's undecided rows)
  if (settled || int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  uint64_t m = mis_tuple(c.state[i], i);
  for (int k = c.rowptr[i]; k < c.rowptr[i + 1]; ++k)
    if (c.strong[k]) {
      const int j = c.col[k];
      const uint64_t t = mis_tuple(c.state[j], j);
      m = t > m ? t : m;
    }
  c.t1[i] = m;
}

// the largest first-hop tuple over a row's strong neighbourhood (the distance-2 maximum), and with it
// the row's decision: an undecided row whose own tuple is that maximum becomes a root, one next to a
// root (the maximum's tag) is excluded; the undecided rows left are counted (a row's decision reads
// only its own distance-2 maximum, so it is taken in the same launch)
__global__ void __launch_bounds__(kWG) k_mis_t2(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles,
                                                int ntiles, int round) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (mis_settled(c, round)) return;
  __shared__ int und;
  if (threadIdx.x == 0) und = 0;
  __syncthreads();
  if (int(threadIdx.x) < tl.nrows) {
    const int i = tl.row0 + threadIdx.x;
    uint64_t m = c.t1[i];
    for (int k = c.rowptr[i]; k < c.rowptr[i + 1]; ++k)
      if (c.strong[k]) {
        const uint64_t t = c.t1[c.col[k]];
        m = t > m ? t : m;
      }
    if (c.state[i] == MIS_UND) {
      if (m == mis_tuple(MIS_UND, i)) c.state[i] = MIS_ROOT;
      else if ((m >> 62) == 3ull) c.state[i] = MIS_EXC;
      else atomicAdd(&und, 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && und) atomicAdd(&c.counts[mis_slot(round)], und);
}

// pass 1: a root is its own aggregate's anchor; its strong neighbours join it
__global__ void k_agg_pass1(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  int root = -1;
  if (c.state[i] == MIS_ROOT) {
    root = i;
  } else {
    uint64_t best = 0;
    for (int k = c.rowptr[i]; k < c.rowptr[i + 1]; ++k) {
      const int j = c.col[k];
      if (c.strong[k] && c.state[j] == MIS_ROOT) {
        const uint64_t t = mis_tuple(MIS_ROOT, j);
        if (t > best) {
          best = t;
          root = j;
        }
      }
    }
  }
  c.rootof[i] = root;
}

// pass 2 (on pass 1's result): the strongest aggregated neighbour's aggregate; then the anchor
// (roots, singletons) and uncoupled-row flags for the numbering scans
__global__ void k_agg_pass2(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  const int k0 = c.rowptr[i], k1 = c.rowptr[i + 1];
  const bool lonely = k1 - k0 <= 1;
  int root = c.rootof[i];
  if (root < 0 && !lonely) {
    double best = -1.0;
    for (int k = k0; k < k1; ++k) {
      const int j = c.col[k];
      if (j == i || c.rootof[j] < 0) continue;
      if (c.s[k] > best) {
        best = c.s[k];
        root = c.rootof[j];
      }
    }
  }
  c.rootof2[i] = root;
  c.scan_a[i] = (c.state[i] == MIS_ROOT || (root < 0 && !lonely)) ? 1 : 0;
  c.scan_b[i] = lonely ? 1 : 0;
}

__global__ void k_scan_pair(const CoarsenDev* __restrict__ J, int count) {
  if (int(blockIdx.x) >= count) return;
  const CoarsenDev& c = J[blockIdx.x];
  block_scan(c.scan_a, c.n, &c.counts[1]);
  __syncthreads();
  block_scan(c.scan_b, c.n, &c.counts[2]);
}

// aggregate ids: anchors in row order, then the uncoupled rows' groups of eight
__global__ void k_assign(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  const int nanchor = c.scan_a[c.n];
  int a;
  if (c.scan_b[i + 1] - c.scan_b[i]) a = nanchor + c.scan_b[i] / 8;
  else if (c.scan_a[i + 1] - c.scan_a[i]) a = c.scan_a[i];
  else a = c.scan_a[c.rootof2[i]];
  c.agg[i] = a;
}

__global__ void k_member_count(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  atomicAdd(&c.acnt[c.agg[tl.row0 + threadIdx.x]], 1);
}

__global__ void k_scan_acnt(const CoarsenDev* __restrict__ J, int count) {
  if (int(blockIdx.x) >= count) return;
  const CoarsenDev& c = J[blockIdx.x];
  block_scan(c.acnt, c.na, nullptr);
}

__global__ void k_member_fill(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  const int a = c.agg[i];
  const int at = atomicAdd(&c.cursor[a], 1);
  c.amem[c.acnt[a] + at] = i;
}

// coarse positions: the members' mean (summed in row order, times 1 / count: amg.cpp:391-398)
__global__ void k_coarse_pos(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int a = tl.row0 + threadIdx.x;
  int32_t* m = c.amem + c.acnt[a];
  const int cnt = c.acnt[a + 1] - c.acnt[a];
  sort_ints(m, cnt);
  double x = 0.0, y = 0.0, z = 0.0;
  for (int t = 0; t < cnt; ++t) {
    const double* p = c.pos + 3 * size_t(m[t]);
    x += p[0];
    y += p[1];
    z += p[2];
  }
  const double s = 1.0 / cnt;
  c.posc[3 * size_t(a)] = x * s;
  c.posc[3 * size_t(a) + 1] = y * s;
  c.posc[3 * size_t(a) + 2] = z * s;
}

__global__ void k_offsets(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  const double* pc = c.posc + 3 * size_t(c.agg[i]);
  for (int q = 0; q < 3; ++q) c.off[3 * size_t(i) + q] = c.pos[3 * size_t(i) + q] - pc[q];
}

// ---- prolongator ---------------------------------------------------------------------------------

// candidates: the aggregates of the row's couplings (smoothed) or of the row itself, sorted,
// unique, in cand[rowptr[i] ...]; Prow[i] = their count
__global__ void k_p_count(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  const int k0 = c.rowptr[i], k1 = c.rowptr[i + 1];
  int32_t* cd = c.cand + k0;
  int n = 0;
  if (c.omega != 0.0) {
    for (int k = k0; k < k1; ++k) cd[n++] = c.agg[c.col[k]];
  } else {
    cd[n++] = c.agg[i];
  }
  sort_ints(cd, n);
  c.Prow[i] = unique_sorted(cd, n);
}

__global__ void k_scan_rows(const CoarsenDev* __restrict__ J, int count, int which) {
  if (int(blockIdx.x) >= count) return;
  const CoarsenDev& c = J[blockIdx.x];
  switch (which) {
    case 0: block_scan(c.Prow, c.n, &c.counts[4]); break;
    case 1: block_scan(c.Rrow, c.na, nullptr); break;
    case 2: block_scan(c.cptr, c.n, &c.counts[7]); break;
    case 3: block_scan(c.AProw, c.n, &c.counts[5]); break;
    case 4: block_scan(c.cptr, c.na, &c.counts[7]); break;
    case 5: block_scan(c.Crow, c.na, &c.counts[6]); break;
  }
}

// P's columns from the row's sorted unique candidates, and each entry's row
__global__ void k_p_cols(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  const int32_t* cd = c.cand + c.rowptr[i];
  for (int e = c.Prow[i]; e < c.Prow[i + 1]; ++e) {
    c.Pcol[e] = cd[e - c.Prow[i]];
    c.rid[e] = i;
  }
}

// P's blocks, one entry a thread: (J == agg(i) ? P_rigid,i : 0) - w sum over couplings k of row i
// with agg(col k) = J of D_i^-1 (A_k P_rigid,col k), each element in the reference's summation order
__global__ void k_p_fill_thread(const CoarsenDev* __restrict__ J, const Tile* __restrict__ etiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = etiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int e = tl.row0 + threadIdx.x;
  const int i = c.rid[e];
  const int Jc = c.Pcol[e];
  const double* Di = c.Dinv + 36 * size_t(i);
  const double* offi = c.off + 3 * size_t(i);
  double acc[36];
  for (int q = 0; q < 36; ++q) acc[q] = (Jc == c.agg[i]) ? rigid_elem(offi, q / 6, q % 6) : 0.0;
  if (c.omega != 0.0) {
    for (int k = c.rowptr[i]; k < c.rowptr[i + 1]; ++k) {
      const int j = c.col[k];
      if (c.agg[j] != Jc) continue;
      const double* A = c.A + 36 * size_t(k);
      const double* offj = c.off + 3 * size_t(j);
      for (int col = 0; col < 6; ++col) {
        double T[6];
        for (int r = 0; r < 6; ++r) {
          double s = 0.0;
          for (int p = 0; p < 6; ++p) s += A[r * 6 + p] * rigid_elem(offj, p, col);
          T[r] = s;
        }
        for (int r = 0; r < 6; ++r) {
          double s = 0.0;
          for (int p = 0; p < 6; ++p) s += Di[r * 6 + p] * T[p];
          acc[r * 6 + col] += -c.omega * s;
        }
      }
    }
  }
  double* out = c.P + 36 * size_t(e);
  for (int q = 0; q < 36; ++q) out[q] = acc[q];
}

// P's blocks: (J == agg(i) ? P_rigid,i : 0) - w sum over couplings k of row i with agg(col k) = J
// of D_i^-1 (A_k P_rigid,col k), each element in the reference's summation order. A wavefront an
// entry (lanes 0..35: element (r, c)), four entries a workgroup (entry tiles of kFillEntries); D_i^-1
// and each listed coupling's block and offset are staged in LDS, a lane an element, and every lane
// reads the whole block there
__global__ void __launch_bounds__(kWG) k_p_fill_lanes(const CoarsenDev* __restrict__ J, const Tile* __restrict__ etiles,
                                                int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = etiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  const int wv = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
  if (wv >= tl.nrows) return;  // (a whole wavefront)
  __shared__ FillList L;
  __shared__ double Dl[kFillWaves][36], Al[kFillWaves][40];  // (Al[36..38]: the coupled row's offset)
  const int e = tl.row0 + wv;
  const int i = c.rid[e];
  const int Jc = c.Pcol[e];
  const int q = lane < 36 ? lane : 0, r = q / 6, col = q % 6;
  double acc = (Jc == c.agg[i]) ? rigid_elem(c.off + 3 * size_t(i), r, col) : 0.0;
  if (c.omega != 0.0) {
    if (lane < 36) Dl[wv][lane] = c.Dinv[36 * size_t(i) + size_t(lane)];
    const int k0 = c.rowptr[i], k1 = c.rowptr[i + 1];
    for (int kc = k0; kc < k1; kc += 64) {
      const int k = kc + lane;
      int j = 0;
      bool hit = false;
      if (k < k1) {
        j = c.col[k];
        hit = c.agg[j] == Jc;
      }
      const int nh = fill_list(L, wv, lane, hit, k, j);
      for (int t = 0; t < nh; ++t) {
        if (lane < 36) Al[wv][lane] = c.A[36 * size_t(L.a[wv][t]) + size_t(lane)];
        else if (lane < 39) Al[wv][lane] = c.off[3 * size_t(L.b[wv][t]) + size_t(lane - 36)];
        wave_sync();
        if (lane < 36) {
          const double* A = Al[wv];
          const double* offj = Al[wv] + 36;
          const double* Di = Dl[wv];
          double T[6];
          for (int rr = 0; rr < 6; ++rr) {
            double s2 = 0.0;
            for (int p = 0; p < 6; ++p) s2 += A[rr * 6 + p] * rigid_elem(offj, p, col);
            T[rr] = s2;
          }
          double s2 = 0.0;
          for (int p = 0; p < 6; ++p) s2 += Di[r * 6 + p] * T[p];
          acc += -c.omega * s2;
        }
        wave_sync();  // (the block is read before the next one overwrites it)
      }
    }
  }
  if (lane < 36) c.P[36 * size_t(e) + size_t(q)] = acc;
}

// restriction lists: P's entries by coarse column, fine rows ascending
__global__ void k_r_count(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  for (int e = c.Prow[i]; e < c.Prow[i + 1]; ++e) atomicAdd(&c.Rrow[c.Pcol[e]], 1);
}

__global__ void k_r_fill(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  for (int e = c.Prow[i]; e < c.Prow[i + 1]; ++e) {
    const int I = c.Pcol[e];
    const int at = c.Rrow[I] + atomicAdd(&c.cursor[I], 1);
    c.Rfine[at] = i;
    c.Rent[at] = e;
  }
}

// per coarse row: entries sorted by fine row (the pairs move together). A wavefront a row (four
// rows a workgroup, 64 workgroups a coarse-row tile): each entry's rank is the count of smaller
// fine rows in the row (they are distinct), the pairs land at their ranks through LDS; a row longer
// than the LDS holds is sorted by one lane
constexpr int kRSortCap = 1024;
__global__ void __launch_bounds__(kWG) k_r_sort(const CoarsenDev* __restrict__ J, const Tile* __restrict__ ctiles,
                                                int nctiles) {
  const int tile = int(blockIdx.x) >> 6, part = int(blockIdx.x) & 63;
  if (tile >= nctiles) return;
  const Tile tl = ctiles[tile];
  const CoarsenDev& c = J[tl.g];
  const int w = int(threadIdx.x) >> 6, l = int(threadIdx.x) & 63;
  const int rl = 4 * part + w;
  if (rl >= tl.nrows) return;
  const int I = tl.row0 + rl;
  const int b = c.Rrow[I], n = c.Rrow[I + 1] - b;
  int32_t* f = c.Rfine + b;
  int32_t* en = c.Rent + b;
  __shared__ int32_t sf[4][kRSortCap], se[4][kRSortCap];
  if (n <= kRSortCap) {
    for (int t = l; t < n; t += 64) {
      sf[w][t] = f[t];
      se[w][t] = en[t];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int t = l; t < n; t += 64) {
      const int32_t key = sf[w][t];
      int rank = 0;
      for (int u = 0; u < n; ++u) rank += sf[w][u] < key ? 1 : 0;
      f[rank] = key;
      en[rank] = se[w][t];
    }
    return;
  }
  if (l != 0) return;
  for (int t = 1; t < n; ++t) {
    const int32_t fv = f[t], ev = en[t];
    int u = t - 1;
    while (u >= 0 && f[u] > fv) {
      f[u + 1] = f[u];
      en[u + 1] = en[u];
      --u;
    }
    f[u + 1] = fv;
    en[u + 1] = ev;
  }
}

// ---- Galerkin product ----------------------------------------------------------------------------

// candidate counts of A P's row i: sum over couplings of the coupled row's P entries
__global__ void k_ap_cand_count(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int i = tl.row0 + threadIdx.x;
  int n = 0;
  for (int k = c.rowptr[i]; k < c.rowptr[i + 1]; ++k) {
    const int j = c.col[k];
    n += c.Prow[j + 1] - c.Prow[j];
  }
  c.cptr[i] = n;
}

// A P's candidate columns of row i (unsorted): the P columns of each coupled row, in coupling order.
// A wavefront a row (four rows a workgroup, 64 workgroups a row tile): each lane copies one coupled
// row's P columns to its offset (a wavefront prefix sum of their counts)
__global__ void __launch_bounds__(kWG) k_ap_cand_fill(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles,
                                                      int ntiles) {
  const int tile = int(blockIdx.x) >> 6, part = int(blockIdx.x) & 63;
  if (tile >= ntiles) return;
  const Tile tl = tiles[tile];
  const CoarsenDev& c = J[tl.g];
  const int w = int(threadIdx.x) >> 6, l = int(threadIdx.x) & 63;
  const int rl = 4 * part + w;
  if (rl >= tl.nrows) return;
  const int i = tl.row0 + rl;
  int32_t* cd = c.cand + c.cptr[i];
  int base = 0;
  const int k1 = c.rowptr[i + 1];
  for (int k0 = c.rowptr[i]; k0 < k1; k0 += 64) {
    const int k = k0 + l;
    int x0 = 0, len = 0;
    if (k < k1) {
      const int j = c.col[k];
      x0 = c.Prow[j];
      len = c.Prow[j + 1] - x0;
    }
    int incl = len;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (l >= o) incl += t;
    }
    int32_t* dst = cd + base + incl - len;
    for (int x = 0; x < len; ++x) dst[x] = c.Pcol[x0 + x];
    base += __shfl(incl, 63, 64);
  }
}

// candidate counts of the coarse row I: its diagonal and the A P rows of its restriction entries
__global__ void k_c_cand_count(const CoarsenDev* __restrict__ J, const Tile* __restrict__ ctiles, int nctiles) {
  if (int(blockIdx.x) >= nctiles) return;
  const Tile tl = ctiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int I = tl.row0 + threadIdx.x;
  int n = 1;
  for (int r = c.Rrow[I]; r < c.Rrow[I + 1]; ++r) {
    const int i = c.Rfine[r];
    n += c.AProw[i + 1] - c.AProw[i];
  }
  c.cptr[I] = n;
}

// the coarse row I's candidate columns (unsorted): I, then the A P columns of its restriction rows
// in order. A wavefront a row (four rows a workgroup, 64 workgroups a coarse-row tile): each lane
// copies one restriction row's columns to its offset (a wavefront prefix sum of the row lengths)
__global__ void __launch_bounds__(kWG) k_c_cand_fill(const CoarsenDev* __restrict__ J, const Tile* __restrict__ ctiles,
                                                     int nctiles) {
  const int tile = int(blockIdx.x) >> 6, part = int(blockIdx.x) & 63;
  if (tile >= nctiles) return;
  const Tile tl = ctiles[tile];
  const CoarsenDev& c = J[tl.g];
  const int w = int(threadIdx.x) >> 6, l = int(threadIdx.x) & 63;
  const int rl = 4 * part + w;
  if (rl >= tl.nrows) return;
  const int I = tl.row0 + rl;
  int32_t* cd = c.cand + c.cptr[I];
  if (l == 0) cd[0] = I;
  int base = 1;
  const int r1 = c.Rrow[I + 1];
  for (int r0 = c.Rrow[I]; r0 < r1; r0 += 64) {
    const int r = r0 + l;
    int x0 = 0, len = 0;
    if (r < r1) {
      const int i = c.Rfine[r];
      x0 = c.AProw[i];
      len = c.AProw[i + 1] - x0;
    }
    int incl = len;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (l >= o) incl += t;
    }
    int32_t* dst = cd + base + incl - len;
    for (int x = 0; x < len; ++x) dst[x] = c.APcol[x0 + x];
    base += __shfl(incl, 63, 64);
  }
}

constexpr int kHashSlots = 16384;  // LDS hash set (64 KB)
constexpr int kSortSlots = 4096;   // LDS sort buffer (16 KB): at most this many distinct columns a row

// Each row's candidate segment (cand[cptr[r] ..]) made its sorted distinct columns in place, their
// count in out[r] (the diagonal first when which = 1: coarse rows, as Bsr6Builder::finish). A
// wavefront a row (tiles of four rows): an LDS hash set (order-independent, so the result is
// deterministic), the occupied slots compacted by a wavefront prefix sum, each column placed at its
// rank. A row whose distinct columns overflow the wavefront's share of LDS is left with out[r] = -1
// for the workgroup pass (k_unique_rows) that follows.
constexpr int kUHash = 2048;  // hash slots a wavefront
constexpr int kUCap = 1024;   // distinct columns a wavefront places
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__global__ void __launch_bounds__(kWG) k_unique_rows_wave(const CoarsenDev* __restrict__ J,
                                                          const Tile* __restrict__ tiles, int ntiles, int which) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  const int w = int(threadIdx.x) >> 6, l = int(threadIdx.x) & 63;
  if (w >= tl.nrows) return;
  __shared__ int32_t tab[4][kUHash];
  __shared__ int32_t keys[4][kUCap];
  const int r = tl.row0 + w;
  int32_t* out = which == 0 ? c.AProw : c.Crow;
  int32_t* cd = c.cand + c.cptr[r];
  const int len = c.cptr[r + 1] - c.cptr[r];
  int H = 64;
  while (H < 2 * len && H < kUHash) H <<= 1;
  for (int h = l; h < H; h += 64) tab[w][h] = -1;
  wave_lds_fence();
  bool full = false;
  for (int k = l; k < len; k += 64) {
    const int32_t key = cd[k];
    uint32_t h = (uint32_t(key) * 2654435761U) & uint32_t(H - 1);
    for (int probes = 0;; ++probes) {
      if (probes >= H) {
        full = true;
        break;
      }
      const int32_t old = atomicCAS(&tab[w][h], -1, key);
      if (old == -1 || old == key) break;
      h = (h + 1) & uint32_t(H - 1);
    }
  }
  wave_lds_fence();
  const int per = H / 64;
  int mine = 0;
  for (int h = l * per; h < (l + 1) * per; ++h) mine += tab[w][h] >= 0;
  int incl = mine;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (l >= o) incl += t;
  }
  const int u = __shfl(incl, 63, 64);
  if (__any(full) || u > kUCap) {
    if (l == 0) out[r] = -1;
    return;
  }
  int at = incl - mine;
  for (int h = l * per; h < (l + 1) * per; ++h)
    if (tab[w][h] >= 0) keys[w][at++] = tab[w][h];
  wave_lds_fence();
  for (int t = l; t < u; t += 64) {
    const int32_t key = keys[w][t];
    int rank = 0;
    for (int q = 0; q < u; ++q) rank += keys[w][q] < key ? 1 : 0;
    if (which == 1) rank = key == r ? 0 : rank + (key < r ? 1 : 0);
    cd[rank] = key;
  }
  if (l == 0) out[r] = u;
}

// The rows k_unique_rows_wave left (out[r] = -1): one workgroup a row, an LDS hash set, then a
// bitonic sort (at most kSortSlots distinct columns)
__global__ void __launch_bounds__(kWG) k_unique_rows(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles,
                                                     int ntiles, int which) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  __shared__ int32_t tab[kHashSlots];
  __shared__ int32_t keys[kSortSlots];
  __shared__ int32_t wsum[4];
  __shared__ int32_t total;
  int32_t* out = which == 0 ? c.AProw : c.Crow;
  for (int rr = 0; rr < tl.nrows; ++rr) {
    const int r = tl.row0 + rr;
    if (out[r] != -1) continue;
    int32_t* cd = c.cand + c.cptr[r];
    const int len = c.cptr[r + 1] - c.cptr[r];
    int H = 64;
    while (H < 2 * len && H < kHashSlots) H <<= 1;
    for (int h = threadIdx.x; h < H; h += kWG) tab[h] = -1;
    __syncthreads();
    for (int k = threadIdx.x; k < len; k += kWG) {
      const int32_t key = cd[k];
      uint32_t h = (uint32_t(key) * 2654435761U) & uint32_t(H - 1);
      for (;;) {
        const int32_t old = atomicCAS(&tab[h], -1, key);
        if (old == -1 || old == key) break;
        h = (h + 1) & uint32_t(H - 1);
      }
    }
    __syncthreads();
    // compact the occupied slots (slot order; sorted below)
    const int per = (H + kWG - 1) / kWG;
    int mine = 0;
    for (int h = threadIdx.x * per; h < min(H, (threadIdx.x + 1) * per); ++h) mine += tab[h] >= 0;
    int incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if ((threadIdx.x & 63) >= o) incl += t;
    }
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < int(threadIdx.x >> 6); ++w) woff += wsum[w];
    if (threadIdx.x == 0) total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    const int u = total;
    if (u > kSortSlots) {
      if (threadIdx.x == 0) {
        atomicAdd(&c.counts[3], 1);
        out[r] = 0;
      }
      __syncthreads();
      continue;
    }
    int at = woff + incl - mine;
    for (int h = threadIdx.x * per; h < min(H, (threadIdx.x + 1) * per); ++h)
      if (tab[h] >= 0) keys[at++] = tab[h];
    int P2 = 1;
    while (P2 < u) P2 <<= 1;
    for (int k = u + threadIdx.x; k < P2; k += kWG) keys[k] = 0x7fffffff;
    __syncthreads();
    for (int size = 2; size <= P2; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = threadIdx.x; t < P2 / 2; t += kWG) {
          const int lo = 2 * stride * (t / stride) + (t % stride), hi = lo + stride;
          const bool up = ((lo & size) == 0);
          const int32_t a0 = keys[lo], a1 = keys[hi];
          if ((a0 > a1) == up) {
            keys[lo] = a1;
            keys[hi] = a0;
          }
        }
        __syncthreads();
      }
    if (which == 1) {
      // the diagonal first, then ascending (Bsr6Builder::finish)
      if (threadIdx.x == 0) {
        int d = 0;
        while (d < u && keys[d] != r) ++d;
        for (int t = d; t > 0; --t) keys[t] = keys[t - 1];
        keys[0] = r;
      }
      __syncthreads();
    }
    for (int k = threadIdx.x; k < u; k += kWG) cd[k] = keys[k];
    if (threadIdx.x == 0) out[r] = u;
    __syncthreads();
  }
}

// the pattern from the rows' sorted distinct candidates, and each entry's row
__global__ void k_pattern_cols(const CoarsenDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles,
                               int which) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int r = tl.row0 + threadIdx.x;
  const int32_t* cd = c.cand + c.cptr[r];
  const int32_t* ptr = which == 0 ? c.AProw : c.Crow;
  int32_t* cols = which == 0 ? c.APcol : c.Ccol;
  for (int x = ptr[r]; x < ptr[r + 1]; ++x) {
    cols[x] = cd[x - ptr[r]];
    c.rid[x] = r;
  }
}

// (A P)_iJ, one entry a thread: the sum over couplings k of row i (in order) of A_k P_e, e the entry
// of column J in P's row col k (P rows ascending: a binary search)
__global__ void k_ap_fill_thread(const CoarsenDev* __restrict__ J, const Tile* __restrict__ etiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = etiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int x = tl.row0 + threadIdx.x;
  const int i = c.rid[x];
  const int Jc = c.APcol[x];
  double acc[36];
  for (int q = 0; q < 36; ++q) acc[q] = 0.0;
  for (int k = c.rowptr[i]; k < c.rowptr[i + 1]; ++k) {
    const int j = c.col[k];
    int lo = c.Prow[j], hi = c.Prow[j + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (c.Pcol[mid] < Jc) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= c.Prow[j + 1] || c.Pcol[lo] != Jc) continue;
    const double* A = c.A + 36 * size_t(k);
    const double* P = c.P + 36 * size_t(lo);
    for (int r = 0; r < 6; ++r)
      for (int col = 0; col < 6; ++col) {
        double s = 0.0;
        for (int q = 0; q < 6; ++q) s += A[r * 6 + q] * P[q * 6 + col];
        acc[r * 6 + col] = acc[r * 6 + col] + s;
      }
  }
  double* out = c.AP + 36 * size_t(x);
  for (int q = 0; q < 36; ++q) out[q] = acc[q];
}

// (A P)_iJ: the sum over couplings k of row i (in order) of A_k P_e, e the entry of column J in P's
// row col k (P rows ascending: a binary search). A wavefront an entry (lanes 0..35: the elements),
// four entries a workgroup
__global__ void __launch_bounds__(kWG) k_ap_fill_lanes(const CoarsenDev* __restrict__ J, const Tile* __restrict__ etiles,
                                                 int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = etiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  const int wv = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
  if (wv >= tl.nrows) return;
  __shared__ FillList L;
  __shared__ double Xl[kFillWaves][2][36];
  const int x = tl.row0 + wv;
  const int i = c.rid[x];
  const int Jc = c.APcol[x];
  const int q = lane < 36 ? lane : 0, r = q / 6, col = q % 6;
  double acc = 0.0;
  const int k0 = c.rowptr[i], k1 = c.rowptr[i + 1];
  for (int kc = k0; kc < k1; kc += 64) {
    const int k = kc + lane;
    int pe = -1;
    if (k < k1) {
      const int j = c.col[k];
      pe = find_sorted(c.Pcol, c.Prow[j], c.Prow[j + 1], Jc);
    }
    const int nh = fill_list(L, wv, lane, pe >= 0, k, pe);
    for (int t = 0; t < nh; ++t) {
      // (the coupling's A block and P entry staged in LDS, an element a lane)
      if (lane < 36) {
        Xl[wv][0][lane] = c.A[36 * size_t(L.a[wv][t]) + size_t(lane)];
        Xl[wv][1][lane] = c.P[36 * size_t(L.b[wv][t]) + size_t(lane)];
      }
      wave_sync();
      if (lane < 36) {
        const double* A = Xl[wv][0] + 6 * r;
        const double* P = Xl[wv][1] + col;
        double s2 = 0.0;
        for (int u = 0; u < 6; ++u) s2 += A[u] * P[6 * u];
        acc = acc + s2;
      }
      wave_sync();
    }
  }
  if (lane < 36) c.AP[36 * size_t(x) + size_t(q)] = acc;
}

// (P^T A P)_IJ, one entry a thread: the sum over restriction entries (i, e) of I, in order, of
// P_e^T (A P)_iJ; times the coarse scale (unsmoothed levels)
__global__ void k_c_fill_thread(const CoarsenDev* __restrict__ J, const Tile* __restrict__ etiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = etiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int y = tl.row0 + threadIdx.x;
  const int I = c.rid[y];
  const int Jc = c.Ccol[y];
  double acc[36];
  for (int q = 0; q < 36; ++q) acc[q] = 0.0;
  for (int r = c.Rrow[I]; r < c.Rrow[I + 1]; ++r) {
    const int i = c.Rfine[r];
    int lo = c.AProw[i], hi = c.AProw[i + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (c.APcol[mid] < Jc) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= c.AProw[i + 1] || c.APcol[lo] != Jc) continue;
    const double* P = c.P + 36 * size_t(c.Rent[r]);
    const double* X = c.AP + 36 * size_t(lo);
    for (int rr = 0; rr < 6; ++rr)
      for (int col = 0; col < 6; ++col) {
        double s = 0.0;
        for (int q = 0; q < 6; ++q) s += P[q * 6 + rr] * X[q * 6 + col];
        acc[rr * 6 + col] = acc[rr * 6 + col] + s;
      }
  }
  double* out = c.C + 36 * size_t(y);
  for (int q = 0; q < 36; ++q) out[q] = c.coarse_scale != 1.0 ? acc[q] * c.coarse_scale : acc[q];
}

// (P^T A P)_IJ: the sum over restriction entries (i, e) of I, in order, of P_e^T (A P)_iJ; times
// the coarse scale (unsmoothed levels). A wavefront an entry (lanes 0..35: the elements), four
// entries a workgroup
__global__ void __launch_bounds__(kWG) k_c_fill_lanes(const CoarsenDev* __restrict__ J, const Tile* __restrict__ etiles,
                                                int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = etiles[blockIdx.x];
  const CoarsenDev& c = J[tl.g];
  const int wv = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
  if (wv >= tl.nrows) return;
  __shared__ FillList L;
  __shared__ double Xl[kFillWaves][2][36];
  const int y = tl.row0 + wv;
  const int I = c.rid[y];
  const int Jc = c.Ccol[y];
  const int q = lane < 36 ? lane : 0, rr = q / 6, col = q % 6;
  double acc = 0.0;
  const int r0 = c.Rrow[I], r1 = c.Rrow[I + 1];
  for (int rc = r0; rc < r1; rc += 64) {
    const int r = rc + lane;
    int lo = -1, ent = 0;
    if (r < r1) {
      const int i = c.Rfine[r];
      lo = find_sorted(c.APcol, c.AProw[i], c.AProw[i + 1], Jc);
      ent = c.Rent[r];
    }
    const int nh = fill_list(L, wv, lane, lo >= 0, ent, lo);
    for (int t = 0; t < nh; ++t) {
      // (the restriction entry's P block and the A P block staged in LDS, an element a lane)
      if (lane < 36) {
        Xl[wv][0][lane] = c.P[36 * size_t(L.a[wv][t]) + size_t(lane)];
        Xl[wv][1][lane] = c.AP[36 * size_t(L.b[wv][t]) + size_t(lane)];
      }
      wave_sync();
      if (lane < 36) {
        const double* P = Xl[wv][0] + rr;
        const double* X = Xl[wv][1] + col;
        double s2 = 0.0;
        for (int u = 0; u < 6; ++u) s2 += P[6 * u] * X[6 * u];
        acc = acc + s2;
      }
      wave_sync();
    }
  }
  if (lane < 36) c.C[36 * size_t(y) + size_t(q)] = c.coarse_scale != 1.0 ? acc * c.coarse_scale : acc;
}

// ---- power iteration (6 lanes a row, kTileRows-row tiles) ------------------------------------------

__global__ void __launch_bounds__(kWG) k_power_init(const PowerDev* __restrict__ J, const Tile* __restrict__ tiles,
                                                    int ntiles, int count) {
  if (blockIdx.x == 0)
    for (int j = threadIdx.x; j < count; j += kWG) {
      J[j].lam[0] = 1.0;
      J[j].lam[1] = 1.0;
    }
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const PowerDev& p = J[tl.g];
  int rl, q;
  lane_row(rl, q);
  if (rl < 0 || rl >= tl.nrows) return;
  const size_t e = 6 * size_t(tl.row0 + rl) + q;
  p.v[0][e] = 1.0 + 0.001 * double(e % 7);
}

// (A x)_q of `row` over a tile's rows [row0, row0 + nrows) of an FP64 BSR matrix, a block a pair of
// lanes: the tile's blocks are contiguous, 128 a pass; a lane pair loads one block (three rows each) and
// the six x values of its column once, and writes the block's row sums (row q . x, left to right) to
// LDS; lane q of each row then adds its row's sums in block order - the row walk's sum, bit for bit.
// Every lane of the workgroup calls it (it synchronises the workgroup).
__device__ __forceinline__ double lpb_rows64(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                             const double* __restrict__ A, const double* __restrict__ x, int row0,
                                             int nrows, bool on, int row, int q) {
  constexpr int PER = kWG / 2;
  __shared__ double yb[6 * PER];
  const int k0 = rowptr[row0], k1 = rowptr[row0 + nrows];
  int rb0 = 0, rb1 = 0;
  if (on) {
    rb0 = rowptr[row];
    rb1 = rowptr[row + 1];
  }
  const int lb = int(threadIdx.x) >> 1, q0 = (int(threadIdx.x) & 1) * 3;
  double acc = 0.0;
  for (int c = k0; c < k1; c += PER) {
    const int kb = c + lb;
    if (kb < k1) {
      const double* a = A + 36 * size_t(kb) + 6 * q0;
      const double* v = x + 6 * size_t(col[kb]);
      double av[18], xv[6];
      for (int u = 0; u < 18; ++u) av[u] = a[u];
      for (int u = 0; u < 6; ++u) xv[u] = v[u];
      for (int r = 0; r < 3; ++r) {
        double s = 0.0;
        for (int u = 0; u < 6; ++u) s += av[6 * r + u] * xv[u];
        yb[6 * lb + q0 + r] = s;
      }
    }
    __syncthreads();
    if (on) {
      const int e0 = max(rb0, c), e1 = min(rb1, c + PER);
      for (int k = e0; k < e1; ++k) acc += yb[6 * (k - c) + q];
    }
    __syncthreads();
  }
  return acc;
}

// v_{s+1} = D^-1 A (v_s / |v_s|) over a tile, and the tile's partial of |v_{s+1}|^2 (slot parity s & 1:
// step s + 1 reads it while writing the other). S six-lane sub-groups a row: S = 10 serves the long
// rows of coarse levels a wavefront each; S = 1 takes the products a block a lane pair. inv: 1 / |v_s|.
template <int S>
__device__ __forceinline__ void power_step_tile(const PowerDev& p, const Tile& tl, int step, double inv,
                                                double* __restrict__ partial) {
  const double* in = p.v[step & 1];
  double* out = p.v[(step + 1) & 1];
  int rl, sub, q;
  lane_row_s<S>(rl, sub, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int row = tl.row0 + (on ? rl : 0);
  double w = 0.0;
  if constexpr (S == 1) {
    w = lpb_rows64(p.rowptr, p.col, p.A, in, tl.row0, tl.nrows, on, row, q);
  } else if (on) {
    for (int k = p.rowptr[row] + sub; k < p.rowptr[row + 1]; k += S) {
      const double* a = p.A + 36 * size_t(k) + 6 * q;
      const double* x = in + 6 * size_t(p.col[k]);
      double s = 0.0;
      for (int c = 0; c < 6; ++c) s += a[c] * x[c];
      w += s;
    }
  }
  w = row_reduce<S>(w) * inv;
  double wv[6];
  row_gather(w, wv);
  double v = 0.0;
  if (on && sub == 0) {
    const double* D = p.Dinv + 36 * size_t(row) + 6 * q;
    v = D[0] * wv[0] + D[1] * wv[1] + D[2] * wv[2] + D[3] * wv[3] + D[4] * wv[4] + D[5] * wv[5];
    out[6 * size_t(row) + q] = v;
  }
  __shared__ double lds[4];
  const double s = block_sum(v * v, lds);
  if (threadIdx.x == 0) partial[2 * tl.slot + (step & 1)] = s;
}

// |v_s|^2 from a graph's per-tile partials of step s - 1 (slots k0 .. k1, parity `which`): each lane's
// strided sum in slot order, then the workgroup's fixed order (thread 0 holds it)
__device__ __forceinline__ double power_sum(int k0, int k1, const double* __restrict__ partial, int which) {
  __shared__ double lds[4];
  double a = 0.0;
  for (int k = k0 + int(threadIdx.x); k < k1; k += kWG) a += partial[2 * k + which];
  return block_sum(a, lds);
}

// One step of every graph's power iteration in one launch: workgroups [0, n1) take tiles1 (S = 1),
// the rest tiles10 (S = 10). A graph of few tiles (p.self) normalises itself: each of its workgroups
// sums the last step's partials (the same order everywhere, so the same |v_s|), stops when |v_s| is
// not positive (writing zero partials, so that it stays stopped) and its first tile keeps |v_s| in
// lam; a graph of many tiles reads |v_s| from lam, which k_power_finish set (stopped: lam[1] = 0).
__global__ void __launch_bounds__(kWG) k_power_steps(const PowerDev* __restrict__ J, const Tile* __restrict__ tiles1,
                                                     int n1, const Tile* __restrict__ tiles10, int n10,
                                                     const int32_t* __restrict__ gtile, int step,
                                                     double* __restrict__ partial) {
  const int b = int(blockIdx.x);
  if (b >= n1 + n10) return;
  const bool longr = b >= n1;
  const Tile tl = longr ? tiles10[b - n1] : tiles1[b];
  const PowerDev& p = J[tl.g];
  double inv = 1.0;
  if (p.self) {
    if (step > 0) {
      __shared__ double nvs;
      const double ss = power_sum(gtile[tl.g], gtile[tl.g + 1], partial, (step - 1) & 1);
      if (threadIdx.x == 0) nvs = sqrt(ss);
      __syncthreads();
      const double nv = nvs;
      if (!(nv > 0)) {
        if (threadIdx.x == 0) partial[2 * tl.slot + (step & 1)] = 0.0;
        return;
      }
      if (threadIdx.x == 0 && tl.slot == gtile[tl.g]) {
        p.lam[0] = nv;
        p.lam[1] = nv;
      }
      inv = 1.0 / nv;
    }
  } else {
    if (p.lam[1] <= 0.0) return;  // (stopped: a zero vector)
    if (step > 0) inv = 1.0 / p.lam[1];
  }
  if (longr) power_step_tile<10>(p, tl, step, inv, partial);
  else power_step_tile<1>(p, tl, step, inv, partial);
}

// |v_{s+1}| of the graphs of many tiles (close = 0) after step s, or of the graphs that normalise
// themselves after the last step (close = 1): lam = |v_{s+1}| while positive, else the graph stops
// (lam[1] = 0; lam[0] keeps the last estimate). A workgroup a graph.
__global__ void __launch_bounds__(kWG) k_power_finish(const PowerDev* __restrict__ J, const int32_t* __restrict__ gtile,
                                                      int count, const double* __restrict__ partial, int which,
                                                      int close) {
  if (int(blockIdx.x) >= count) return;
  const PowerDev& p = J[blockIdx.x];
  if ((p.self != 0) != (close != 0) || p.lam[1] <= 0.0) return;
  const double s = power_sum(gtile[blockIdx.x], gtile[blockIdx.x + 1], partial, which);
  if (threadIdx.x == 0) {
    const double nv = sqrt(s);
    if (!(nv > 0)) {
      p.lam[1] = 0.0;
      return;
    }
    p.lam[0] = nv;
    p.lam[1] = nv;
  }
}

// ---- dense coarsest level: blocked Cholesky and explicit inverse on the matrix cores ---------------
//
// One workgroup of 16 waves a graph. The coarsest matrix (m = 6 n_c rows, symmetrised, as
// amg.cpp:535-578) is padded with an identity to mp = 16 ceil(m / 16) rows and held as 16 x 16
// tiles in global scratch (L2-resident at the usual m <= 288). A right-looking blocked Cholesky
// L L^T, per block column: one wave factors the diagonal tile in LDS with the reference's pivot
// floor (d >= 1e-12 max|a_ii|) and forms the tile's triangular inverse; the panel below is
// multiplied by that inverse's transpose; the trailing lower triangle takes the panel's rank-16
// update. Then W = L^-1 (row block by row block) and M^-1 = W^T W, written to the m x m inverse.
// Every 16 x 16 x 16 tile product is four v_mfma_f64_16x16x4f64.

typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int kDenseWG = 1024;
constexpr int kDenseWaves = kDenseWG / 64;

// lane layout of v_mfma_f64_16x16x4f64 (lane l = 16 g + i): A[i][k] at k = g, B[k][j] at k = g,
// j = i; the accumulator's register r holds D[g + 4 r][i]
// a load through a global-address-space pointer (the scratch is device memory; through a generic
// pointer the compiler emits flat loads)
__device__ __forceinline__ double gld(const double* p) { return *(const __attribute__((address_space(1))) double*)p; }

__device__ __forceinline__ dbl4 tile_load(const double* T, int ld) {
  const int i = threadIdx.x & 15, g = (threadIdx.x & 63) >> 4;
  dbl4 c;
#pragma unroll
  for (int r = 0; r < 4; ++r) c[r] = gld(T + size_t(g + 4 * r) * size_t(ld) + size_t(i));
  return c;
}
__device__ __forceinline__ void tile_store(double* T, int ld, dbl4 c) {
  const int i = threadIdx.x & 15, g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) T[size_t(g + 4 * r) * size_t(ld) + size_t(i)] = c[r];
}
// c + sign op(X) op(Y) over a 16-wide inner dimension; X, Y row-major tiles with leading dimensions
// ldx, ldy; tX: X read transposed, tY: Y read transposed
template <bool tX, bool tY>
__device__ __forceinline__ dbl4 tile_mma(dbl4 c, const double* X, int ldx, const double* Y, int ldy, double sign) {
  const int i = threadIdx.x & 15, g = (threadIdx.x & 63) >> 4;
  // (all eight operands loaded before the first product: one memory round trip a tile product)
  double a[4], b[4];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int k = 4 * kk + g;
    a[kk] = tX ? X[size_t(k) * size_t(ldx) + size_t(i)] : X[size_t(i) * size_t(ldx) + size_t(k)];
    b[kk] = tY ? Y[size_t(i) * size_t(ldy) + size_t(k)] : Y[size_t(k) * size_t(ldy) + size_t(i)];
    // (X, Y: global scratch or LDS, so the generic pointer stays)
  }
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) c = __builtin_amdgcn_mfma_f64_16x16x4f64(sign * a[kk], b[kk], c, 0, 0, 0);
  return c;
}
// LDS writes of this wave visible to its other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// the padded matrix in L (identity below m), W zeroed, the blocks scattered, the lower triangle
// symmetrised (amg.cpp:535-578); returns max |a_ii| (the pivot floor's scale). A 1024-thread
// workgroup.
__device__ double dense_prep(const DenseDev& d) {
  const int m = d.m, mp = d.mp;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* L = d.L;
  double* W = d.W;
  auto at = [&](int r, int c) -> double& { return L[size_t(r) * size_t(mp) + size_t(c)]; };
  // (rows a wave, columns its lanes)
  for (int r = wave; r < mp; r += kDenseWaves)
    for (int c = lane; c < mp; c += 64) {
      at(r, c) = (r == c && r >= m) ? 1.0 : 0.0;
      W[size_t(r) * size_t(mp) + size_t(c)] = 0.0;
    }
  __syncthreads();
  // the blocks, an element a thread (a block row's blocks have distinct columns: no two write one place)
  for (int i = wave; i < d.n; i += kDenseWaves) {
    const int k0 = d.rowptr[i], k1 = d.rowptr[i + 1];
    for (int t = lane; t < 36 * (k1 - k0); t += 64) {
      const int k = k0 + t / 36, e = t % 36;
      at(6 * i + e / 6, 6 * d.col[k] + e % 6) = d.A[36 * size_t(k) + size_t(e)];
    }
  }
  __syncthreads();
  __shared__ double red[kDenseWaves];
  double dm = 0.0;
  for (int i = wave; i < m; i += kDenseWaves) {
    for (int j = lane; j < i; j += 64) at(i, j) = 0.5 * (at(i, j) + at(j, i));
    if (lane == 0) dm = fmax(dm, fabs(at(i, i)));
  }
  for (int o = 32; o > 0; o >>= 1) dm = fmax(dm, __shfl_xor(dm, o, 64));
  if (lane == 0) red[wave] = dm;
  __syncthreads();
  double dmax = 0.0;
  for (int w = 0; w < kDenseWaves; ++w) dmax = fmax(dmax, red[w]);
  return dmax;
}

// block column kb: wave 0 factors the diagonal tile in LDS (the reference's pivot floor) and forms
// its triangular inverse Ik (also W's diagonal tile); then the waves take the panel below,
// L_ib,kb = A_ib,kb (L_kb,kb)^-T. A 1024-thread workgroup.
__device__ void dense_panel(const DenseDev& d, int kb, double dmax) {
  const int mp = d.mp, NB = mp / 16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* L = d.L;
  double* W = d.W;
  auto at = [&](int r, int c) -> double& { return L[size_t(r) * size_t(mp) + size_t(c)]; };
  __shared__ double Dk[16][17];
  __shared__ double Ik[16][17];
  const int k0 = 16 * kb;
  if (wave == 0) {
    for (int e = lane; e < 256; e += 64) {
      const int r = e >> 4, c = e & 15;
      Dk[r][c] = c <= r ? at(k0 + r, k0 + c) : 0.0;
    }
    wave_lds_sync();
    for (int j = 0; j < 16; ++j) {
      double dd = Dk[j][j];
      if (!(dd > 1e-12 * dmax)) dd = fmax(1e-12 * dmax, 1e-300);
      const double lj = sqrt(dd);
      wave_lds_sync();
      if (lane == j) Dk[j][j] = lj;
      else if (lane > j && lane < 16) Dk[lane][j] /= lj;
      wave_lds_sync();
      for (int e = lane; e < 256; e += 64) {
        const int r = e >> 4, c = e & 15;
        if (c > j && c <= r) Dk[r][c] -= Dk[r][j] * Dk[c][j];
      }
      wave_lds_sync();
    }
    if (lane < 16) {  // column `lane` of the inverse by forward substitution
      const int c = lane;
      for (int r = 0; r < c; ++r) Ik[r][c] = 0.0;
      Ik[c][c] = 1.0 / Dk[c][c];
      for (int r = c + 1; r < 16; ++r) {
        double s = 0.0;
        for (int q = c; q < r; ++q) s += Dk[r][q] * Ik[q][c];
        Ik[r][c] = -s / Dk[r][r];
      }
    }
    wave_lds_sync();
    for (int e = lane; e < 256; e += 64) {
      const int r = e >> 4, c = e & 15;
      at(k0 + r, k0 + c) = Dk[r][c];
      W[size_t(k0 + r) * size_t(mp) + size_t(k0 + c)] = Ik[r][c];
    }
  }
  __syncthreads();
  for (int ib = kb + 1 + wave; ib < NB; ib += kDenseWaves) {
    double* T = &at(16 * ib, k0);
    dbl4 c = {0.0, 0.0, 0.0, 0.0};
    c = tile_mma<false, true>(c, T, mp, &Ik[0][0], 17, 1.0);
    tile_store(T, mp, c);
  }
  __syncthreads();
}

// tile t of a lower triangle of n tile rows: (row a, column t - a(a+1)/2)
__device__ __forceinline__ void tri_tile(int t, int* a_out, int* b_out) {
  int a = int((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (a * (a + 1) / 2 > t) --a;
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  *a_out = a;
  *b_out = t - a * (a + 1) / 2;
}

// the trailing lower triangle's tile t after block column kb: A_ib,jb -= L_ib,kb L_jb,kb^T
__device__ __forceinline__ void dense_update_tile(const DenseDev& d, int kb, int t) {
  const int mp = d.mp, k0 = 16 * kb;
  int a, b;
  tri_tile(t, &a, &b);
  const int ib = kb + 1 + a, jb = kb + 1 + b;
  double* L = d.L;
  double* C = L + size_t(16 * ib) * size_t(mp) + size_t(16 * jb);
  dbl4 c = tile_load(C, mp);
  c = tile_mma<false, true>(c, L + size_t(16 * ib) * size_t(mp) + size_t(k0), mp,
                            L + size_t(16 * jb) * size_t(mp) + size_t(k0), mp, -1.0);
  tile_store(C, mp, c);
}

// M^-1's lower tile t = sum over kb >= ib of W_kb,ib^T W_kb,jb, written and mirrored
__device__ __forceinline__ void dense_wtw_tile(const DenseDev& d, int t) {
  const int m = d.m, mp = d.mp, NB = mp / 16;
  const int lane = threadIdx.x & 63;
  int ib, jb;
  tri_tile(t, &ib, &jb);
  const double* W = d.W;
  dbl4 c = {0.0, 0.0, 0.0, 0.0};
  for (int kb = ib; kb < NB; ++kb)
    c = tile_mma<true, false>(c, &W[size_t(16 * kb) * size_t(mp) + size_t(16 * ib)], mp,
                              &W[size_t(16 * kb) * size_t(mp) + size_t(16 * jb)], mp, 1.0);
  const int i = lane & 15, g = lane >> 4;
  for (int r = 0; r < 4; ++r) {
    const int row = 16 * ib + g + 4 * r, col = 16 * jb + i;
    if (row >= m || col >= m || col > row) continue;
    d.Minv[size_t(row) * size_t(m) + size_t(col)] = c[r];
    d.Minv[size_t(col) * size_t(m) + size_t(row)] = c[r];
  }
}

__global__ void __launch_bounds__(kDenseWG) k_dense_factor(const DenseDev* __restrict__ J, int count) {
  if (int(blockIdx.x) >= count) return;
  const DenseDev& d = J[blockIdx.x];
  const int mp = d.mp, NB = mp / 16;
  const int wave = threadIdx.x >> 6;
  double* L = d.L;
  double* W = d.W;
  const double dmax = dense_prep(d);
  for (int kb = 0; kb < NB; ++kb) {
    dense_panel(d, kb, dmax);
    const int n = NB - kb - 1;
    for (int t = wave; t < n * (n + 1) / 2; t += kDenseWaves) dense_update_tile(d, kb, t);
    __syncthreads();
  }
  // W = L^-1 below the diagonal: W_ib,jb = -W_ib,ib sum_{jb <= kb < ib} L_ib,kb W_kb,jb
  __shared__ double Sw[kDenseWaves][16][17];
  for (int ib = 1; ib < NB; ++ib) {
    for (int jb = wave; jb < ib; jb += kDenseWaves) {
      dbl4 c = {0.0, 0.0, 0.0, 0.0};
      for (int kb = jb; kb < ib; ++kb)
        c = tile_mma<false, false>(c, &L[size_t(16 * ib) * size_t(mp) + size_t(16 * kb)], mp,
                                   &W[size_t(16 * kb) * size_t(mp) + size_t(16 * jb)], mp, 1.0);
      tile_store(&Sw[wave][0][0], 17, c);
      wave_lds_sync();
      dbl4 z = {0.0, 0.0, 0.0, 0.0};
      z = tile_mma<false, false>(z, &W[size_t(16 * ib) * size_t(mp) + size_t(16 * ib)], mp, &Sw[wave][0][0], 17, -1.0);
      wave_lds_sync();
      tile_store(&W[size_t(16 * ib) * size_t(mp) + size_t(16 * jb)], mp, z);
    }
    __syncthreads();
  }
  for (int t = wave; t < NB * (NB + 1) / 2; t += kDenseWaves) dense_wtw_tile(d, t);
}

// ---- the same over many workgroups (coarsest levels above kDenseOneBlocks tile rows) ------------------
//
// A level of C2's size (864 rows, 54 tile rows) leaves one workgroup with some 27k tile products in
// sequence. Spread: per block column, a workgroup a graph factors the diagonal tile and solves the
// panel (the sequential part), then a wave a trailing tile updates the lower triangle; L^-1 goes
// column block by column block, a workgroup each (its rows in order, each row's sum split over the
// 16 waves and added in wave order); M^-1 = W^T W a wave a tile. The arithmetic of a tile product is
// the one-workgroup kernel's; L^-1's sums are grouped differently (same terms, a fixed order), so a
// graph's inverse depends only on its size class, never on its batch.

// dense_prep spread: kDensePrepParts workgroups a graph, part p owning the block rows i = p mod parts
// (their six matrix rows) and the padding rows r = m + p mod parts. First the part's rows of L are
// zeroed (identity on padding rows), W's likewise zeroed, and the part's block rows scattered in;
// then (after the launch) the lower triangle of the part's rows is symmetrised, 0.5 (a_ij + a_ji),
// and max |a_ii| kept by an atomic maximum on its bits (magnitudes order as their bits): the same
// values dense_prep leaves.
constexpr int kDensePrepParts = 16;

__global__ void __launch_bounds__(kWG) k_dense_scatter(const DenseDev* __restrict__ J, double* __restrict__ dmax) {
  const DenseDev& d = J[blockIdx.y];
  const int part = int(blockIdx.x), m = d.m, mp = d.mp;
  double* L = d.L;
  double* W = d.W;
  if (part == 0 && threadIdx.x == 0) dmax[blockIdx.y] = 0.0;
  auto clear = [&](int r) {
    for (int c = int(threadIdx.x); c < mp; c += kWG) {
      L[size_t(r) * size_t(mp) + size_t(c)] = (r == c && r >= m) ? 1.0 : 0.0;
      W[size_t(r) * size_t(mp) + size_t(c)] = 0.0;
    }
  };
  for (int i = part; i < d.n; i += kDensePrepParts)
    for (int q = 0; q < 6; ++q) clear(6 * i + q);
  for (int r = m + part; r < mp; r += kDensePrepParts) clear(r);
  __syncthreads();
  for (int i = part; i < d.n; i += kDensePrepParts) {
    const int k0 = d.rowptr[i], k1 = d.rowptr[i + 1];
    for (int t = int(threadIdx.x); t < 36 * (k1 - k0); t += kWG) {
      const int k = k0 + t / 36, e = t % 36;
      L[size_t(6 * i + e / 6) * size_t(mp) + size_t(6 * d.col[k] + e % 6)] = d.A[36 * size_t(k) + size_t(e)];
    }
  }
}

__global__ void __launch_bounds__(kWG) k_dense_sym(const DenseDev* __restrict__ J, double* __restrict__ dmax) {
  const DenseDev& d = J[blockIdx.y];
  const int part = int(blockIdx.x), mp = d.mp;
  double* L = d.L;
  double dm = 0.0;
  for (int i = part; i < d.n; i += kDensePrepParts)
    for (int q = 0; q < 6; ++q) {
      const int r = 6 * i + q;
      for (int c = int(threadIdx.x); c < r; c += kWG)
        L[size_t(r) * size_t(mp) + size_t(c)] =
            0.5 * (L[size_t(r) * size_t(mp) + size_t(c)] + L[size_t(c) * size_t(mp) + size_t(r)]);
      if (threadIdx.x == 0) dm = fmax(dm, fabs(L[size_t(r) * size_t(mp) + size_t(r)]));
    }
  if (threadIdx.x == 0 && dm > 0.0)
    atomicMax(reinterpret_cast<unsigned long long*>(dmax + blockIdx.y), __double_as_longlong(dm));
}

__global__ void __launch_bounds__(kDenseWG) k_dense_panel(const DenseDev* __restrict__ J, int count,
                                                          const double* __restrict__ dmax, int kb) {
  if (int(blockIdx.x) >= count) return;
  const DenseDev& d = J[blockIdx.x];
  if (kb >= d.mp / 16) return;
  dense_panel(d, kb, dmax[blockIdx.x]);
}

// a wave a tile: grid (tile groups of four, graphs)
__global__ void __launch_bounds__(256) k_dense_update(const DenseDev* __restrict__ J, int kb) {
  const DenseDev& d = J[blockIdx.y];
  const int n = d.mp / 16 - kb - 1;
  const int t = int(blockIdx.x) * 4 + int(threadIdx.x >> 6);
  if (n <= 0 || t >= n * (n + 1) / 2) return;
  dense_update_tile(d, kb, t);
}

// column block jb of W = L^-1: grid (column blocks, graphs)
__global__ void __launch_bounds__(kDenseWG) k_dense_winv(const DenseDev* __restrict__ J) {
  const DenseDev& d = J[blockIdx.y];
  const int mp = d.mp, NB = mp / 16, jb = blockIdx.x;
  if (jb >= NB - 1) return;
  const int wave = threadIdx.x >> 6;
  const double* L = d.L;
  double* W = d.W;
  __shared__ double Sp[kDenseWaves][16][17];
  for (int ib = jb + 1; ib < NB; ++ib) {
    dbl4 c = {0.0, 0.0, 0.0, 0.0};
    for (int kb = jb + wave; kb < ib; kb += kDenseWaves)
      c = tile_mma<false, false>(c, &L[size_t(16 * ib) * size_t(mp) + size_t(16 * kb)], mp,
                                 &W[size_t(16 * kb) * size_t(mp) + size_t(16 * jb)], mp, 1.0);
    tile_store(&Sp[wave][0][0], 17, c);
    __syncthreads();
    if (wave == 0) {
      // (the waves' partial sums added in wave order, then times -W_ib,ib)
      const int i = threadIdx.x & 15, g = (threadIdx.x & 63) >> 4;
      for (int w = 1; w < kDenseWaves && jb + w < ib; ++w)
        for (int r = 0; r < 4; ++r) Sp[0][g + 4 * r][i] += Sp[w][g + 4 * r][i];
      wave_lds_sync();
      dbl4 z = {0.0, 0.0, 0.0, 0.0};
      z = tile_mma<false, false>(z, &W[size_t(16 * ib) * size_t(mp) + size_t(16 * ib)], mp, &Sp[0][0][0], 17, -1.0);
      tile_store(&W[size_t(16 * ib) * size_t(mp) + size_t(16 * jb)], mp, z);
    }
    __syncthreads();
  }
}

// a wave a tile of M^-1's lower triangle: grid (tile groups of four, graphs)
__global__ void __launch_bounds__(256) k_dense_wtw(const DenseDev* __restrict__ J) {
  const DenseDev& d = J[blockIdx.y];
  const int NB = d.mp / 16;
  const int t = int(blockIdx.x) * 4 + int(threadIdx.x >> 6);
  if (t >= NB * (NB + 1) / 2) return;
  dense_wtw_tile(d, t);
}

// ---- conversions -----------------------------------------------------------------------------------

__global__ void k_to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t count) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int64_t k = t; k < count; k += int64_t(gridDim.x) * blockDim.x) out[k] = float(in[k]);
}

// ---- bulk segments (one workgroup a segment at a time) ------------------------------------------------

__global__ void __launch_bounds__(kWG) k_zero_many(const BulkSeg* __restrict__ segs, int n) {
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const BulkSeg g = segs[k];
    uint4* d = static_cast<uint4*>(g.dst);
    const int64_t v = g.count >> 4;
    for (int64_t i = threadIdx.x; i < v; i += kWG) d[i] = make_uint4(0u, 0u, 0u, 0u);
    uint8_t* b = static_cast<uint8_t*>(g.dst);
    for (int64_t i = (v << 4) + threadIdx.x; i < g.count; i += kWG) b[i] = 0;
  }
}

// a block's entries over its largest magnitude, rounded to FP16 (A_ij and A_ji = A_ij^T hold the
// same values, so a symmetric matrix's copy stays symmetric); tr: the block's transpose is stored
// (a restriction entry's block, read as P's block transposed)
__device__ __forceinline__ void half_block(const float* __restrict__ a, uint8_t* __restrict__ h, bool tr) {
  float m = 0.0f;
  for (int i = 0; i < 36; ++i) m = fmaxf(m, fabsf(a[i]));
  const float inv = m > 0.0f ? 1.0f / m : 0.0f;
  for (int i = 0; i < 36; i += 2) {
    // (the transpose's entries (r, c), (r, c + 1) are the block's (c, r), (c + 1, r))
    const int r = i / 6, c = i % 6;
    const __half2 v = tr ? __floats2half2_rn(a[6 * c + r] * inv, a[6 * (c + 1) + r] * inv)
                         : __floats2half2_rn(a[i] * inv, a[i + 1] * inv);
    *reinterpret_cast<__half2*>(h + 2 * i) = v;
  }
  reinterpret_cast<float*>(h + 72)[0] = m;
  reinterpret_cast<float*>(h + 72)[1] = 0.0f;
}

__global__ void __launch_bounds__(kWG) k_half_blocks(const float* __restrict__ A, uint8_t* __restrict__ H, int64_t nnzb) {
  const int64_t k = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (k >= nnzb) return;
  half_block(A + 36 * k, H + kHalfBlock * k, false);
}

// the same for the transposes of the blocks P[ent[e]] (a restriction entry's block), in entry order
__global__ void __launch_bounds__(kWG) k_half_blocks_t(const float* __restrict__ P, const int32_t* __restrict__ ent,
                                                       uint8_t* __restrict__ H, int64_t n) {
  const int64_t e = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (e >= n) return;
  half_block(P + 36 * size_t(ent[e]), H + kHalfBlock * e, true);
}

// many arrays' FP16 copies in one launch: a workgroup a piece of at most kWG blocks
__global__ void __launch_bounds__(kWG) k_half_many(const HalfSeg* __restrict__ segs, int n) {
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const HalfSeg g = segs[k];
    const int64_t e = int64_t(threadIdx.x);
    if (e >= g.count) continue;
    const float* a = g.ent ? g.src + 36 * size_t(g.ent[e]) : g.src + 36 * e;
    half_block(a, g.dst + kHalfBlock * e, g.ent != nullptr);
  }
}

__global__ void __launch_bounds__(kWG) k_f32_many(const BulkSeg* __restrict__ segs, int n) {
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const BulkSeg g = segs[k];
    const double* in = static_cast<const double*>(g.src);
    float* out = static_cast<float*>(g.dst);
    for (int64_t i = threadIdx.x; i < g.count; i += kWG) out[i] = float(in[i]);
  }
}

__global__ void __launch_bounds__(kWG) k_copy_many(const BulkSeg* __restrict__ segs, int n) {
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const BulkSeg g = segs[k];
    const int64_t v = g.count >> 4;
    const uint4* s4 = static_cast<const uint4*>(g.src);
    uint4* d4 = static_cast<uint4*>(g.dst);
    for (int64_t i = threadIdx.x; i < v; i += kWG) d4[i] = s4[i];
    const uint8_t* s1 = static_cast<const uint8_t*>(g.src);
    uint8_t* d1 = static_cast<uint8_t*>(g.dst);
    for (int64_t i = (v << 4) + threadIdx.x; i < g.count; i += kWG) d1[i] = s1[i];
  }
}

// ---- bond checks -------------------------------------------------------------------------------------

__global__ void k_bond_checks(const CheckDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const CheckDev& c = J[tl.g];
  if (int(threadIdx.x) >= tl.nrows) return;
  const int bi = tl.row0 + threadIdx.x;
  const svxg_bond& b = c.bonds[bi];
  const svxg_material& MA = bond_mat(c.mats, c.nmats, b.ma);
  const svxg_material& MB = bond_mat(c.mats, c.nmats, b.mb);
  const svxg_bond_load L = bond_load_dev(b, c.nodes, MA, MB, c.u);
  if (c.loads) c.loads[bi] = L;
  uint8_t md;
  const double phi = bond_phi_dev(b, bond_strength_dev(b, MA, MB, c.fragility), L, &md);
  if (c.phi) c.phi[bi] = phi;
  if (c.mode) c.mode[bi] = md;
}

__global__ void k_scatter_u(const double* __restrict__ x, const int32_t* __restrict__ dof, double* __restrict__ u,
                            int nnodes) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nnodes) return;
  const int i = t / 6, q = t % 6;
  const int d = dof[i];
  u[t] = d >= 0 ? x[6 * size_t(d) + q] : 0.0;
}

// x (free DOFs) from u (all nodes): the inverse of k_scatter_u
__global__ void k_gather_u(const double* __restrict__ u, const int32_t* __restrict__ dof, double* __restrict__ x,
                           int nnodes) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nnodes) return;
  const int d = dof[t / 6];
  if (d >= 0) x[6 * size_t(d) + t % 6] = u[t];
}

// every value of v that is not finite becomes 0
__global__ void k_zero_nonfinite(double* __restrict__ v, int64_t n) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < n && !isfinite(v[t])) v[t] = 0.0;
}

__global__ void __launch_bounds__(kWG) k_scatter_many(const ScatterJob* __restrict__ jobs, int n) {
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const ScatterJob j = jobs[k];
    for (int t = threadIdx.x; t < 6 * j.nnodes; t += kWG) {
      const int d = j.dof[t / 6];
      j.u[t] = d >= 0 ? j.x[6 * size_t(d) + t % 6] : 0.0;
    }
  }
}

// ---- host launchers -----------------------------------------------------------------------------------

void launch_bond_products(hipStream_t s, const AssemblyDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_bond_products, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_gather_blocks(hipStream_t s, const AssemblyDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_gather_blocks, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_diag_trace(hipStream_t s, const AssemblyDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_diag_trace, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_id_scale(hipStream_t s, const AssemblyDev* jobs, int count) {
  if (count > 0) hipLaunchKernelGGL(k_id_scale, dim3(count), dim3(kIdWG), 0, s, jobs, count);
}
void launch_identity_rows(hipStream_t s, const AssemblyDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_identity_rows, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_removal_products(hipStream_t s, const AssemblyDev* jobs, const int32_t* ids, const int32_t* tjob,
                             double* prod, int nids) {
  if (nids > 0) hipLaunchKernelGGL(k_removal_products, dim3(nids), dim3(64), 0, s, jobs, ids, tjob, prod, nids);
}
void launch_removal_apply(hipStream_t s, const AssemblyDev* jobs, const int32_t* gjob, const int32_t* gslot,
                          const int32_t* gstart, const int32_t* cont, const double* prod, int ngroups) {
  if (ngroups > 0)
    hipLaunchKernelGGL(k_removal_apply, dim3(unsigned((36 * int64_t(ngroups) + kWG - 1) / kWG)), dim3(kWG), 0, s, jobs,
                       gjob, gslot, gstart, cont, prod, ngroups);
}
void launch_retire_rows(hipStream_t s, const AssemblyDev* job, const int32_t* rows, int count) {
  if (count > 0) hipLaunchKernelGGL(k_retire_rows, dim3((36 * count + 255) / 256), dim3(256), 0, s, job, rows, count);
}
void launch_inv_diag(hipStream_t s, const DiagDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_inv_diag, dim3(64 * ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_strength(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_strength, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_strong(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_strong, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_mis_round(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, int round) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_mis_t1, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles, round);
  hipLaunchKernelGGL(k_mis_t2, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles, round);
}
void launch_aggregate(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, int count) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_agg_pass1, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
  hipLaunchKernelGGL(k_agg_pass2, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
  hipLaunchKernelGGL(k_scan_pair, dim3(count), dim3(kWG), 0, s, jobs, count);
}
void launch_scan_pair(hipStream_t s, const CoarsenDev* jobs, int count) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_scan_pair, dim3(count), dim3(kWG), 0, s, jobs, count);
}
void launch_assign(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_assign, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_coarse_positions(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, const Tile* ctiles,
                             int nctiles, int count) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_member_count, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
  hipLaunchKernelGGL(k_scan_acnt, dim3(count), dim3(kWG), 0, s, jobs, count);
  hipLaunchKernelGGL(k_member_fill, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
  hipLaunchKernelGGL(k_coarse_pos, dim3(nctiles), dim3(kWG), 0, s, jobs, ctiles, nctiles);
  hipLaunchKernelGGL(k_offsets, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_p_count(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_p_count, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_scan_rows(hipStream_t s, const CoarsenDev* jobs, int count, int which) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_scan_rows, dim3(count), dim3(kWG), 0, s, jobs, count, which);
}
void launch_p_cols(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_p_cols, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_p_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* etiles, int ntiles, bool lanes) {
  if (ntiles <= 0) return;
  if (lanes) hipLaunchKernelGGL(k_p_fill_lanes, dim3(ntiles), dim3(kWG), 0, s, jobs, etiles, ntiles);
  else hipLaunchKernelGGL(k_p_fill_thread, dim3(ntiles), dim3(kWG), 0, s, jobs, etiles, ntiles);
}
void launch_r_build(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, const Tile* ctiles,
                    int nctiles, int count) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_r_count, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
  hipLaunchKernelGGL(k_scan_rows, dim3(count), dim3(kWG), 0, s, jobs, count, 1);
  hipLaunchKernelGGL(k_r_fill, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
  hipLaunchKernelGGL(k_r_sort, dim3(64 * nctiles), dim3(kWG), 0, s, jobs, ctiles, nctiles);
}
void launch_ap_cand_count(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_ap_cand_count, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_ap_cand_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_ap_cand_fill, dim3(64 * ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_c_cand_count(hipStream_t s, const CoarsenDev* jobs, const Tile* ctiles, int nctiles) {
  if (nctiles > 0) hipLaunchKernelGGL(k_c_cand_count, dim3(nctiles), dim3(kWG), 0, s, jobs, ctiles, nctiles);
}
void launch_c_cand_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* ctiles, int nctiles) {
  if (nctiles > 0) hipLaunchKernelGGL(k_c_cand_fill, dim3(64 * nctiles), dim3(kWG), 0, s, jobs, ctiles, nctiles);
}
void launch_unique_rows(hipStream_t s, const CoarsenDev* jobs, const Tile* utiles, int ntiles, int which) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_unique_rows_wave, dim3(ntiles), dim3(kWG), 0, s, jobs, utiles, ntiles, which);
  hipLaunchKernelGGL(k_unique_rows, dim3(ntiles), dim3(kWG), 0, s, jobs, utiles, ntiles, which);
}
void launch_pattern_cols(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, int which) {
  if (ntiles > 0) hipLaunchKernelGGL(k_pattern_cols, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles, which);
}
void launch_ap_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* etiles, int ntiles, bool lanes) {
  if (ntiles <= 0) return;
  if (lanes) hipLaunchKernelGGL(k_ap_fill_lanes, dim3(ntiles), dim3(kWG), 0, s, jobs, etiles, ntiles);
  else hipLaunchKernelGGL(k_ap_fill_thread, dim3(ntiles), dim3(kWG), 0, s, jobs, etiles, ntiles);
}
void launch_c_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* etiles, int ntiles, bool lanes) {
  if (ntiles <= 0) return;
  if (lanes) hipLaunchKernelGGL(k_c_fill_lanes, dim3(ntiles), dim3(kWG), 0, s, jobs, etiles, ntiles);
  else hipLaunchKernelGGL(k_c_fill_thread, dim3(ntiles), dim3(kWG), 0, s, jobs, etiles, ntiles);
}
void launch_power_init(hipStream_t s, const PowerDev* jobs, const Tile* tiles, int ntiles, int count) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_power_init, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles, count);
}
void launch_power_steps(hipStream_t s, const PowerDev* jobs, const Tile* tiles1, int n1, const Tile* tiles10, int n10,
                        const int32_t* gtile, int step, double* partial) {
  if (n1 + n10 <= 0) return;
  hipLaunchKernelGGL(k_power_steps, dim3(n1 + n10), dim3(kWG), 0, s, jobs, tiles1, n1, tiles10, n10, gtile, step,
                     partial);
}
void launch_power_finish(hipStream_t s, const PowerDev* jobs, const int32_t* gtile, int count, const double* partial,
                         int which, bool close) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_power_finish, dim3(count), dim3(kWG), 0, s, jobs, gtile, count, partial, which, close ? 1 : 0);
}
void launch_dense_factor(hipStream_t s, const DenseDev* jobs, int count) {
  if (count > 0) hipLaunchKernelGGL(k_dense_factor, dim3(count), dim3(kDenseWG), 0, s, jobs, count);
}
void launch_dense_factor_wide(hipStream_t s, const DenseDev* jobs, int count, int max_nb, double* dmax) {
  if (count <= 0 || max_nb <= 0) return;
  hipLaunchKernelGGL(k_dense_scatter, dim3(kDensePrepParts, unsigned(count)), dim3(kWG), 0, s, jobs, dmax);
  hipLaunchKernelGGL(k_dense_sym, dim3(kDensePrepParts, unsigned(count)), dim3(kWG), 0, s, jobs, dmax);
  for (int kb = 0; kb < max_nb; ++kb) {
    hipLaunchKernelGGL(k_dense_panel, dim3(count), dim3(kDenseWG), 0, s, jobs, count, dmax, kb);
    const int n = max_nb - kb - 1;
    if (n > 0)
      hipLaunchKernelGGL(k_dense_update, dim3(unsigned((n * (n + 1) / 2 + 3) / 4), unsigned(count)), dim3(256), 0, s,
                         jobs, kb);
  }
  if (max_nb > 1) hipLaunchKernelGGL(k_dense_winv, dim3(unsigned(max_nb - 1), unsigned(count)), dim3(kDenseWG), 0, s, jobs);
  hipLaunchKernelGGL(k_dense_wtw, dim3(unsigned((max_nb * (max_nb + 1) / 2 + 3) / 4), unsigned(count)), dim3(256), 0,
                     s, jobs);
}
void launch_to_f32(hipStream_t s, const double* in, float* out, int64_t count) {
  if (count <= 0) return;
  const int64_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_to_f32, dim3(unsigned(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, in, out, count);
}
static unsigned bulk_grid(int n) { return unsigned(n < 65536 ? n : 65536); }
void launch_zero_many(hipStream_t s, const BulkSeg* segs, int n) {
  if (n > 0) hipLaunchKernelGGL(k_zero_many, dim3(bulk_grid(n)), dim3(kWG), 0, s, segs, n);
}
void launch_half_blocks(hipStream_t s, const float* A, uint8_t* H, int64_t nnzb) {
  if (nnzb > 0) hipLaunchKernelGGL(k_half_blocks, dim3(unsigned((nnzb + kWG - 1) / kWG)), dim3(kWG), 0, s, A, H, nnzb);
}

void launch_half_blocks_t(hipStream_t s, const float* P, const int32_t* ent, uint8_t* H, int64_t n) {
  if (n > 0) hipLaunchKernelGGL(k_half_blocks_t, dim3(unsigned((n + kWG - 1) / kWG)), dim3(kWG), 0, s, P, ent, H, n);
}

void launch_half_many(hipStream_t s, const HalfSeg* segs, int n) {
  if (n > 0) hipLaunchKernelGGL(k_half_many, dim3(bulk_grid(n)), dim3(kWG), 0, s, segs, n);
}
void launch_f32_many(hipStream_t s, const BulkSeg* segs, int n) {
  if (n > 0) hipLaunchKernelGGL(k_f32_many, dim3(bulk_grid(n)), dim3(kWG), 0, s, segs, n);
}
void launch_copy_many(hipStream_t s, const BulkSeg* segs, int n) {
  if (n > 0) hipLaunchKernelGGL(k_copy_many, dim3(bulk_grid(n)), dim3(kWG), 0, s, segs, n);
}
void launch_bond_checks(hipStream_t s, const CheckDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_bond_checks, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}
void launch_scatter_many(hipStream_t s, const ScatterJob* jobs, int n) {
  if (n > 0) hipLaunchKernelGGL(k_scatter_many, dim3(bulk_grid(n)), dim3(kWG), 0, s, jobs, n);
}
void launch_gather_u(hipStream_t s, const double* u, const int32_t* dof, double* x, int nnodes) {
  if (nnodes > 0) hipLaunchKernelGGL(k_gather_u, dim3((6 * nnodes + 255) / 256), dim3(256), 0, s, u, dof, x, nnodes);
}
void launch_zero_nonfinite(hipStream_t s, double* v, int64_t n) {
  if (n > 0) hipLaunchKernelGGL(k_zero_nonfinite, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, v, n);
}
void launch_scatter_u(hipStream_t s, const double* x, const int32_t* dof, double* u, int nnodes) {
  if (nnodes > 0) hipLaunchKernelGGL(k_scatter_u, dim3((6 * nnodes + 255) / 256), dim3(256), 0, s, x, dof, u, nnodes);
}

}  // namespace svxg
