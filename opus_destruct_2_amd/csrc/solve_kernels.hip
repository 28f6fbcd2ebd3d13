// Solve-phase kernels of the svxg stress solve: block-sparse products, the Chebyshev-smoothed
// V-cycle (smoothing, residual, restriction, prolongation, dense coarsest solve), block-Jacobi
// preconditioning and the PCG vector updates with their fixed-order reductions.
//
// The reference's hot loops these replace (core/src/solve/amg.cpp, core/src/stress/stress.cpp):
//   Bsr6::apply amg.cpp:174-183       -> spmv_row (the same per-row sum order: a row's six
//                                        products summed, then added to the running sum)
//   Amg::smooth_f amg.cpp:630-659     -> k_cheb_step (Chebyshev in D^-1 A, row-parallel)
//   residual / restriction / prolong  -> k_residual, k_restrict, k_prolong (amg.cpp:683-714)
//   solve_coarsest amg.cpp:580-592    -> k_dense_solve (the coarsest level's inverse, FP64)
//   StressProblem::iterate            -> k_pcg_spmv, k_axpy_rr, k_update_p, k_finish_*
//     stress.cpp:403-464                 (the dots: per-tile partials summed per graph in order)
#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>

#include <type_traits>

#include "svxg_kernels.hpp"

namespace svxg {

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));
template <typename T>
struct V2;
template <>
struct V2<double> {
  using t = dvec2;
};
template <>
struct V2<float> {
  using t = fvec2;
};

// a load through a global-address-space pointer (every array here is device memory; through a
// generic pointer the compiler emits flat loads)
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// six consecutive values (16-byte aligned for double, 8 for float)
template <typename T>
__device__ __forceinline__ void load6(const T* p, T v[6]) {
  using P = typename V2<T>::t;
  const P* q = reinterpret_cast<const P*>(p);
  const P a = ldg(q), b = ldg(q + 1), c = ldg(q + 2);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
  v[4] = c.x;
  v[5] = c.y;
}

// matrix rows: a row's 6 lanes read a block's 144 / 288 contiguous bytes in three loads of 8 / 16
// bytes each, so a cache line is touched by all three: plain loads keep it in L2 between them (a
// non-temporal hint here made the line leave and come back, half again the matrix's bytes)
template <typename T>
__device__ __forceinline__ void load6_stream(const T* p, T v[6]) {
  load6(p, v);
}

// this lane's share of (A x)_q of block row `row`: its blocks k0 + sub, k0 + sub + S, ... in order,
// each block's row q . x_col summed left to right, the blocks' sums added in turn. Blocks go in
// groups of four whose column indices, then matrix rows and x values, are all loaded before any is
// used (two memory round trips a group instead of two a block; the sums keep their order)
template <typename T, int S>
__device__ __forceinline__ T spmv_row_part(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                           const T* __restrict__ A, const T* __restrict__ x, int row, int q, int sub) {
  constexpr int U = 4;
  T acc = T(0);
  const int k0 = ldg(rowptr + row), k1 = ldg(rowptr + row + 1);
  for (int k = k0 + sub; k < k1; k += U * S) {
    // (past the row's end a group re-reads its last block, unused: the loads stay unconditional)
    int kk[U], j[U];
    T a[U][6], v[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) kk[u] = min(k + u * S, k1 - 1);
#pragma unroll
    for (int u = 0; u < U; ++u) j[u] = ldg(col + kk[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) load6_stream(A + 36 * size_t(kk[u]) + 6 * q, a[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) load6(x + 6 * size_t(j[u]), v[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k + u * S < k1) {
        const T s = a[u][0] * v[u][0] + a[u][1] * v[u][1] + a[u][2] * v[u][2] + a[u][3] * v[u][3] +
                    a[u][4] * v[u][4] + a[u][5] * v[u][5];
        acc += s;
      }
  }
  return acc;
}

// spmv_row_part over the FP16 block copies: row q of each block (six halves at byte 12 q) times x,
// scaled by the block's largest magnitude, in FP32
template <int S>
__device__ __forceinline__ float spmv_row_part_h(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                 const uint8_t* __restrict__ A16, const float* __restrict__ x, int row,
                                                 int q, int sub) {
  constexpr int U = 4;
  float acc = 0.0f;
  const int k0 = ldg(rowptr + row), k1 = ldg(rowptr + row + 1);
  for (int k = k0 + sub; k < k1; k += U * S) {
    int kk[U], j[U];
    uint32_t w[U][3];
    float sc[U], v[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) kk[u] = min(k + u * S, k1 - 1);
#pragma unroll
    for (int u = 0; u < U; ++u) j[u] = ldg(col + kk[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint8_t* blk = A16 + size_t(kHalfBlock) * size_t(kk[u]);
      const uint32_t* rp = reinterpret_cast<const uint32_t*>(blk + 12 * q);
      w[u][0] = ldg(rp);
      w[u][1] = ldg(rp + 1);
      w[u][2] = ldg(rp + 2);
      sc[u] = ldg(reinterpret_cast<const float*>(blk + 72));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) load6(x + 6 * size_t(j[u]), v[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k + u * S < k1) {
        float a[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          __half2 h;
          __builtin_memcpy(&h, &w[u][c], 4);
          const float2 f = __half22float2(h);
          a[2 * c] = f.x;
          a[2 * c + 1] = f.y;
        }
        const float s = a[0] * v[u][0] + a[1] * v[u][1] + a[2] * v[u][2] + a[3] * v[u][3] + a[4] * v[u][4] +
                        a[5] * v[u][5];
        acc += sc[u] * s;
      }
  }
  return acc;
}

// row q of an FP16 block (three words at byte 12 q) and its scale (byte 72)
__device__ __forceinline__ void half_row_load(const uint8_t* __restrict__ blk, int q, uint32_t w[3], float* sc) {
  const uint32_t* rp = reinterpret_cast<const uint32_t*>(blk + 12 * q);
  w[0] = ldg(rp);
  w[1] = ldg(rp + 1);
  w[2] = ldg(rp + 2);
  *sc = ldg(reinterpret_cast<const float*>(blk + 72));
}
// sc (row . v), the row's six products summed left to right in FP32
__device__ __forceinline__ float half_row_dot(const uint32_t w[3], float sc, const float v[6]) {
  float a[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    __half2 h;
    __builtin_memcpy(&h, &w[c], 4);
    const float2 f = __half22float2(h);
    a[2 * c] = f.x;
    a[2 * c + 1] = f.y;
  }
  return sc * (a[0] * v[0] + a[1] * v[1] + a[2] * v[2] + a[3] * v[3] + a[4] * v[4] + a[5] * v[5]);
}

template <typename T, int S>
__device__ __forceinline__ T spmv_row_part_v(const LevelDev& L, const T* __restrict__ A, const T* __restrict__ x,
                                             int row, int q, int sub) {
  if constexpr (std::is_same<T, float>::value) {
    if (L.A16) return spmv_row_part_h<S>(L.rowptr, L.col, L.A16, x, row, q, sub);
  }
  return spmv_row_part<T, S>(L.rowptr, L.col, A, x, row, q, sub);
}

// (A x)_q of block row `row`: sum over the row's blocks of (row q of the block) . x_col
template <typename T>
__device__ __forceinline__ T spmv_row(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                      const T* __restrict__ A, const T* __restrict__ x, int row, int q) {
  return spmv_row_part<T, 1>(rowptr, col, A, x, row, q, 0);
}

// ---- products of a tile's rows, a block a lane ---------------------------------------------------
//
// The rows of a tile own a contiguous range of blocks (a BSR level stores a row's blocks together,
// rows in order). A workgroup takes that range kWG blocks a pass: lane k loads block k's column
// index, the whole block in 16-byte pieces and the six x values of its column, and writes the
// block's six row sums (row q . x, left to right; an FP16 block's scaled by its largest magnitude)
// to LDS; then lane q of each row adds its row's sums in block order. Each block and each gathered
// x is loaded once (row-walking lanes load every x six times), the loads of a pass are all in flight
// together, and the sums are those of spmv_row_part's in the same order: the same bits.

// block kb's row sums y_q = sum_c A_qc x_c for its rows q in [q0, q0 + NQ) (FP16 blocks: all six,
// scale * sum_c a_qc x_c in FP32)
template <typename T, bool HALF, int NQ>
__device__ __forceinline__ void lpb_block(const void* __restrict__ A, const int32_t* __restrict__ col,
                                          const T* __restrict__ x, int kb, int q0, T y[NQ]) {
  const int j = ldg(col + kb);
  if constexpr (HALF) {
    static_assert(NQ == 6, "an FP16 block is taken whole");
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4* bp = reinterpret_cast<const u32x4*>(static_cast<const uint8_t*>(A) + size_t(kHalfBlock) * size_t(kb));
    u32x4 w[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) w[c] = ldg(bp + c);
    float v[6];
    load6(reinterpret_cast<const float*>(x) + 6 * size_t(j), v);
    uint32_t ww[20];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      ww[4 * c] = w[c].x;
      ww[4 * c + 1] = w[c].y;
      ww[4 * c + 2] = w[c].z;
      ww[4 * c + 3] = w[c].w;
    }
    float sc;
    __builtin_memcpy(&sc, &ww[18], 4);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      float a[6];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        __half2 h;
        __builtin_memcpy(&h, &ww[3 * q + c], 4);
        const float2 f = __half22float2(h);
        a[2 * c] = f.x;
        a[2 * c + 1] = f.y;
      }
      const float s = a[0] * v[0] + a[1] * v[1] + a[2] * v[2] + a[3] * v[3] + a[4] * v[4] + a[5] * v[5];
      y[q] = T(sc * s);
    }
  } else {
    using P = typename V2<T>::t;
    const P* bp = reinterpret_cast<const P*>(static_cast<const T*>(A) + 36 * size_t(kb) + 6 * q0);
    P m[3 * NQ];
#pragma unroll
    for (int c = 0; c < 3 * NQ; ++c) m[c] = ldg(bp + c);
    T v[6];
    load6(x + 6 * size_t(j), v);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const P a0 = m[3 * q], a1 = m[3 * q + 1], a2 = m[3 * q + 2];
      y[q] = a0.x * v[0] + a0.y * v[1] + a1.x * v[2] + a1.y * v[3] + a2.x * v[4] + a2.y * v[5];
    }
  }
}

// a block of the restriction read in place from the prolongator (FP32): block ent[kb] of P taken
// transposed - its column q dotted with x_j - in the row walk's order of terms
__device__ __forceinline__ void lpb_block_pt(const float* __restrict__ P, const int32_t* __restrict__ ent,
                                             const int32_t* __restrict__ col, const float* __restrict__ x, int kb,
                                             float y[6]) {
  const int j = ldg(col + kb);
  const int pe = ldg(ent + kb);
  using F2 = typename V2<float>::t;
  const F2* bp = reinterpret_cast<const F2*>(P + 36 * size_t(pe));
  F2 m[18];
#pragma unroll
  for (int c = 0; c < 18; ++c) m[c] = ldg(bp + c);
  float v[6];
  load6(x + 6 * size_t(j), v);
  float a[36];
#pragma unroll
  for (int c = 0; c < 18; ++c) {
    a[2 * c] = m[c].x;
    a[2 * c + 1] = m[c].y;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q)
    y[q] = a[q] * v[0] + a[6 + q] * v[1] + a[12 + q] * v[2] + a[18 + q] * v[3] + a[24 + q] * v[4] + a[30 + q] * v[5];
}

// acc0 + (A x)_q of row `row` (lanes with on = false get 0) over the tile's rows [row0, row0 + nrows),
// the row's block sums added to acc0 in block order; every lane of the workgroup must call it (it
// synchronises the workgroup). H lanes share a block, each
// taking 6 / H of its rows (FP64 blocks: two lanes of 144 bytes each keep more blocks in flight)
template <typename T, bool HALF, bool PT = false>
__device__ __forceinline__ T lpb_rows(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                      const void* __restrict__ A, const T* __restrict__ x, int row0, int nrows,
                                      bool on, int row, int q, T acc0 = T(0), const int32_t* __restrict__ ent = nullptr) {
  static_assert(!PT || (std::is_same<T, float>::value && !HALF), "the in-place transposed blocks are FP32");
  constexpr int H = (!HALF && sizeof(T) == 8) ? 2 : 1;
  constexpr int NQ = 6 / H, PER = kWG / H;
  __shared__ T yb[6 * PER];
  const int k0 = ldg(rowptr + row0), k1 = ldg(rowptr + row0 + nrows);
  int rb0 = 0, rb1 = 0;
  if (on) {
    rb0 = ldg(rowptr + row);
    rb1 = ldg(rowptr + row + 1);
  }
  const int lb = int(threadIdx.x) / H, q0 = (int(threadIdx.x) % H) * NQ;
  T acc = on ? acc0 : T(0);
  for (int c = k0; c < k1; c += PER) {
    const int kb = c + lb;
    if (kb < k1) {
      T y[NQ];
      if constexpr (PT)
        lpb_block_pt(static_cast<const float*>(A), ent, col, x, kb, y);
      else
        lpb_block<T, HALF, NQ>(A, col, x, kb, q0, y);
#pragma unroll
      for (int u = 0; u < NQ; ++u) yb[6 * lb + q0 + u] = y[u];
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

// lpb_rows over a level's matrix in the precision T (FP32: its FP16 copy where one exists)
template <typename T>
__device__ __forceinline__ T lpb_rows_v(const LevelDev& L, const T* __restrict__ A, const T* __restrict__ x, int row0,
                                        int nrows, bool on, int row, int q) {
  if constexpr (std::is_same<T, float>::value) {
    if (L.A16) return lpb_rows<float, true>(L.rowptr, L.col, L.A16, x, row0, nrows, on, row, q);
  }
  return lpb_rows<T, false>(L.rowptr, L.col, A, x, row0, nrows, on, row, q);
}

// (M v)_q with M a 6x6 row-major block and v the row's values gathered from its lanes
template <typename T>
__device__ __forceinline__ T block_row(const T* __restrict__ M, const T v[6], int q) {
  T m[6];
  load6(M + 6 * q, m);
  return m[0] * v[0] + m[1] * v[1] + m[2] * v[2] + m[3] * v[3] + m[4] * v[4] + m[5] * v[5];
}

template <typename T>
__device__ __forceinline__ const LevelArrays<T>& arrays(const LevelDev& L);
template <>
__device__ __forceinline__ const LevelArrays<double>& arrays<double>(const LevelDev& L) {
  return L.f64;
}
template <>
__device__ __forceinline__ const LevelArrays<float>& arrays<float>(const LevelDev& L) {
  return L.f32;
}

// the graph takes part in a launch serving `mode` (state, phase, inner state and product precision
// read in one 16-byte load)
__device__ __forceinline__ bool live(const GraphDev& g, int mode) {
  typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
  const i32x4 f = ldg(reinterpret_cast<const i32x4*>(&g.st->state));
  const int32_t state = f.x, phase = f.y, inner_state = f.z, inner64 = f.w;
  if (state != ST_RUNNING) return false;
  if (mode == SERVE_OUTER) return phase == PH_ITER;
  if (mode == SERVE_NEW) return phase == PH_NEW;
  if (mode == SERVE_REFINE) return phase == PH_OUTER;
  if (inner_state != ST_RUNNING) return false;
  if (!(mode & (inner64 ? SERVE_P64 : SERVE_P32))) return false;
  return (mode & SERVE_ST) ? phase == PH_START || ((mode & SERVE_IT) && phase == PH_INNER)
                           : (mode & SERVE_IT) && phase == PH_INNER;
}

// Chebyshev coefficients of step j of the smoother on [lmax / ratio, lmax]: the direction is
// d_j = c1 d_{j-1} + c2 D^-1 r_j (c1 = 0 at step 0)
__device__ __forceinline__ void cheb_coeffs(double lmax, int j, double& c1, double& c2) {
  const double hi = lmax, lo = lmax / kChebRatio;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  if (j == 0) {
    c1 = 0.0;
    c2 = 1.0 / theta;
    return;
  }
  double rho_prev = rho;
  for (int s = 1; s <= j; ++s) {
    rho_prev = rho;
    rho = 1.0 / (2.0 * sigma - rho_prev);
  }
  c1 = rho * rho_prev;
  c2 = 2.0 * rho / delta;
}

__device__ __forceinline__ void finish_partials(double v0, double v1, double* partial, int slot, bool two) {
  __shared__ double lds[8];
  const double s0 = block_sum(v0, lds);
  double s1 = 0.0;
  if (two) s1 = block_sum(v1, lds + 4);
  if (threadIdx.x == 0) {
    partial[2 * slot] = s0;
    if (two) partial[2 * slot + 1] = s1;
  }
}

// the sum of a graph's per-tile partials (its tiles' slots k0 .. k1), in the same order wherever run
// (each lane's strided sum in slot order, its loads eight at a time; then the workgroup's fixed order)
__device__ __forceinline__ double graph_sum(int k0, int k1, const double* partial, int which) {
  __shared__ double lds[4];
  double a = 0.0;
  int k = k0 + int(threadIdx.x);
  for (; k + 7 * kWG < k1; k += 8 * kWG) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = partial[2 * (k + u * kWG) + which];
#pragma unroll
    for (int u = 0; u < 8; ++u) a += v[u];
  }
  for (; k < k1; k += kWG) a += partial[2 * k + which];
  return block_sum(a, lds);
}

}  // namespace


// ---- V-cycle kernels ---------------------------------------------------------------------------

// One Chebyshev step on `level`: r = b - A x_in (b when fresh: x_in = 0), d = c1 d + c2 D^-1 r,
// x_out = x_in + d. S: six-lane sub-groups a row. FRESH: the first step from x_in = 0 (no matrix
// pass; its own instantiation so that timings of the two kinds of step stay apart). With
// `partial` (level 0's last step, on the PCG's tiles), the partial sums of the PCG's r . z.
template <typename T, int S, bool FRESH>
__device__ __forceinline__ void cheb_step_tile(const GraphDev& g, const Tile& tl, int level, int step, int in_buf, int out_buf, int inner, double* __restrict__ partial) {
  const LevelDev& L = g.lv[level];
  const LevelArrays<T>& a = arrays<T>(L);
  int rl, sub, q;
  lane_row_s<S>(rl, sub, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int row = tl.row0 + (on ? rl : 0);
  const size_t e = 6 * size_t(row) + q;
  bool half = false;
  if constexpr (std::is_same<T, float>::value) half = L.D16 != nullptr;
  // (the loads that do not wait for the product go first: b, x_in, the direction and the row of
  // the inverse diagonal block)
  T xin = T(0), bq = T(0), dold = T(0);
  uint32_t dw[3] = {0u, 0u, 0u};
  float dsc = 0.0f;
  T dm[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if (on) {
    bq = a.b[e];
    if (!FRESH) xin = a.x[in_buf][e];
    if (sub == 0) {
      if (half) half_row_load(L.D16 + size_t(kHalfBlock) * size_t(row), q, dw, &dsc);
      else load6(a.Dinv + 36 * size_t(row) + 6 * q, dm);
      if (step > 0) dold = a.d[e];
    }
  }
  T r = bq;
  if constexpr (!FRESH) {
    T ax = T(0);
    if constexpr (S == 1) ax = lpb_rows_v<T>(L, a.A, a.x[in_buf], tl.row0, tl.nrows, on, row, q);
    else if (on) ax = spmv_row_part_v<T, S>(L, a.A, a.x[in_buf], row, q, sub);
    r = bq - row_reduce<S>(ax);
  }
  T rv[6];
  row_gather(r, rv);
  double c1, c2;
  cheb_coeffs(L.lmax, step, c1, c2);
  T xo = T(0);
  if (on && sub == 0) {
    T z;
    if (half) z = T(half_row_dot(dw, dsc, reinterpret_cast<const float*>(rv)));
    else z = dm[0] * rv[0] + dm[1] * rv[1] + dm[2] * rv[2] + dm[3] * rv[3] + dm[4] * rv[4] + dm[5] * rv[5];
    const T d = (step == 0) ? T(c2) * z : T(c1) * dold + T(c2) * z;
    a.d[e] = d;
    xo = xin + d;
    a.x[out_buf][e] = xo;
  }
  if (partial) {
    // (the V-cycle's last step at level 0, over the PCG's tiles: r . z of the PCG - the residual r of
    // the iteration's precision, FP64 for FP64-product inner solves, times z = x_out - as dot_rz_tile)
    double rz = 0.0;
    if (on && sub == 0) {
      double rp = double(bq);
      if constexpr (std::is_same<T, float>::value) {
        if (g.st->inner64) rp = g.r[e];
      }
      rz = rp * double(xo);
    }
    finish_partials(rz, 0.0, partial, tl.slot, false);
  }

}
template <typename T, int S, bool FRESH>
__global__ void __launch_bounds__(kWG) k_cheb_step(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                   int ntiles, int level, int step, int in_buf, int out_buf,
                                                   int inner, double* __restrict__ partial, ProfAcc pa) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  if (pa.acc && tl.row0 == 0 && threadIdx.x == 0) atomicAdd(pa.acc, pa.gb[tl.g]);
  cheb_step_tile<T, S, FRESH>(g, tl, level, step, in_buf, out_buf, inner, partial);
}
// the V-cycle's last level-0 step with the PCG's r . z partials (its own kernel, so that the plain
// steps' timings stay apart)
template <typename T>
__global__ void __launch_bounds__(kWG) k_cheb_step_rz(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                      int ntiles, int step, int in_buf, int out_buf, int inner,
                                                      double* __restrict__ partial) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  cheb_step_tile<T, 1, false>(g, tl, 0, step, in_buf, out_buf, inner, partial);
}

// r = b - A x
template <typename T, int S>
__device__ __forceinline__ void residual_tile(const GraphDev& g, const Tile& tl, int level, int buf, int inner) {
  const LevelDev& L = g.lv[level];
  const LevelArrays<T>& a = arrays<T>(L);
  int rl, sub, q;
  lane_row_s<S>(rl, sub, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int row = tl.row0 + (on ? rl : 0);
  const size_t e = 6 * size_t(row) + q;
  const T bq = on && sub == 0 ? a.b[e] : T(0);
  T v = T(0);
  if constexpr (S == 1) v = lpb_rows_v<T>(L, a.A, a.x[buf], tl.row0, tl.nrows, on, row, q);
  else if (on) v = spmv_row_part_v<T, S>(L, a.A, a.x[buf], row, q, sub);
  v = row_reduce<S>(v);
  if (!on || sub != 0) return;
  a.r[e] = bq - v;

}
template <typename T, int S>
__global__ void __launch_bounds__(kWG) k_residual(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                  int ntiles, int level, int buf, int inner) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  residual_tile<T, S>(g, tl, level, buf, inner);
}

// column q of a row-major 6x6 block (row q of its transpose): the six lanes of a row read the
// whole block, six consecutive values a load
template <typename T>
__device__ __forceinline__ void load6_col(const T* __restrict__ blk, int q, T* m) {
#pragma unroll
  for (int k = 0; k < 6; ++k) m[k] = ldg(blk + 6 * k + q);
}

// b_c = P^T r: coarse rows (tiles of level + 1), restriction entries of `level`, each the
// transpose of its prolongation block (read in place through Rent)
template <typename T, int S>
__device__ __forceinline__ void restrict_tile(const GraphDev& g, const Tile& tl, int level, int inner) {
  const LevelDev& F = g.lv[level];
  const LevelArrays<T>& af = arrays<T>(F);
  const LevelArrays<T>& ac = arrays<T>(g.lv[level + 1]);
  int rl, sub, q;
  lane_row_s<S>(rl, sub, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int I = tl.row0 + (on ? rl : 0);
  T acc = T(0);
  bool half = false;
  if constexpr (std::is_same<T, float>::value) half = F.R16 != nullptr;
  if constexpr (S == 1) {
    // (the FP16 transposed blocks in restriction order are a matrix by coarse row: a block a lane)
    if (half) {
      acc = T(lpb_rows<float, true>(F.Rrow, F.Rfine, F.R16, reinterpret_cast<const float*>(af.r), tl.row0, tl.nrows,
                                    on, I, q));
      if (on) ac.b[6 * size_t(I) + q] = acc;
      return;
    }
    if constexpr (std::is_same<T, float>::value) {
      // (FP32: the prolongator's blocks read in place through Rent, each taken transposed)
      acc = lpb_rows<float, false, true>(F.Rrow, F.Rfine, af.P, af.r, tl.row0, tl.nrows, on, I, q, 0.0f, F.Rent);
      if (on) ac.b[6 * size_t(I) + q] = acc;
      return;
    }
  }
  if (on && half) {
    // (the FP16 transposed blocks in restriction order: row q of each, no Rent indirection)
    constexpr int U = 4;
    const int e0 = ldg(F.Rrow + I), e1 = ldg(F.Rrow + I + 1);
    for (int e = e0 + sub; e < e1; e += U * S) {
      int i[U], ee[U];
      uint32_t w[U][3];
      float sc[U], v[U][6];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ee[u] = min(e + u * S, e1 - 1);
        i[u] = ldg(F.Rfine + ee[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        half_row_load(F.R16 + size_t(kHalfBlock) * size_t(ee[u]), q, w[u], &sc[u]);
        load6(reinterpret_cast<const float*>(af.r) + 6 * size_t(i[u]), v[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (e + u * S < e1) acc += T(half_row_dot(w[u], sc[u], v[u]));
    }
  } else if (on) {
    // (entries in groups of four, loads first: see spmv_row_part)
    constexpr int U = 4;
    const int e0 = ldg(F.Rrow + I), e1 = ldg(F.Rrow + I + 1);
    for (int e = e0 + sub; e < e1; e += U * S) {
      int i[U], p[U];
      T m[U][6], v[U][6];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int ee = min(e + u * S, e1 - 1);
        i[u] = ldg(F.Rfine + ee);
        p[u] = ldg(F.Rent + ee);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        load6_col(af.P + 36 * size_t(p[u]), q, m[u]);
        load6(af.r + 6 * size_t(i[u]), v[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (e + u * S < e1)
          acc += m[u][0] * v[u][0] + m[u][1] * v[u][1] + m[u][2] * v[u][2] + m[u][3] * v[u][3] +
                 m[u][4] * v[u][4] + m[u][5] * v[u][5];
    }
  }
  acc = row_reduce<S>(acc);
  if (!on || sub != 0) return;
  ac.b[6 * size_t(I) + q] = acc;

}
template <typename T, int S>
__global__ void __launch_bounds__(kWG) k_restrict(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                  int ntiles, int level, int inner) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  restrict_tile<T, S>(g, tl, level, inner);
}

// x += P x_c (x_c: the coarse level's result, x[1]): the prolongator's rows a block a lane
// (lpb_rows from x), each row's entries added to x in entry order
template <typename T>
__device__ __forceinline__ void prolong_tile(const GraphDev& g, const Tile& tl, int level, int buf, int inner) {
  const LevelDev& F = g.lv[level];
  const LevelArrays<T>& af = arrays<T>(F);
  const T* xc = arrays<T>(g.lv[level + 1]).x[1];
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int i = tl.row0 + (on ? rl : 0);
  const size_t e = 6 * size_t(i) + q;
  const T x0 = on ? af.x[buf][e] : T(0);
  bool half = false;
  if constexpr (std::is_same<T, float>::value) half = F.P16 != nullptr;
  T acc;
  if (half)
    acc = T(lpb_rows<float, true>(F.Prow, F.Pcol, F.P16, reinterpret_cast<const float*>(xc), tl.row0, tl.nrows, on, i,
                                  q, float(x0)));
  else
    acc = lpb_rows<T, false>(F.Prow, F.Pcol, af.P, xc, tl.row0, tl.nrows, on, i, q, x0);
  if (on) af.x[buf][e] = acc;
}
template <typename T>
__global__ void __launch_bounds__(kWG) k_prolong(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                 int ntiles, int level, int buf, int inner) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  prolong_tile<T>(g, tl, level, buf, inner);
}

// x[1] = M^-1 b at the coarsest level of each listed graph (FP64 products): one workgroup a graph,
// a wavefront a row at a time, lanes across the row
// x = M^-1 b at a dense coarsest level, rows [r0, r0 + 64): M^-1 is symmetric, so row i is read as
// column i and the 64 lanes' loads of one j are one contiguous run; the j range is cut into
// kDenseSpans spans, which the workgroup's NT / 64 waves share out, and the spans' partial sums meet
// in LDS in span order (so any NT gives the same bits)
constexpr int kDenseRows = 64;
constexpr int kDenseSolveWG = 1024;  // (a wave a span: a 864-row inverse is 54 loads a lane, not 216)
template <typename T, int NT>
__device__ __forceinline__ void dense_rows(const LevelDev& L, const LevelArrays<T>& a, int r0, double* sb,
                                           double* part) {
  const int m = L.dense_m;
  for (int j = threadIdx.x; j < m; j += NT) sb[j] = double(a.b[j]);
  __syncthreads();
  constexpr int nw = NT / 64;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, i = r0 + lane;
  const int span = (m + kDenseSpans - 1) / kDenseSpans;
  for (int sp = w; sp < kDenseSpans; sp += nw) {
    const int j0 = sp * span, j1 = min(m, j0 + span);
    double s = 0.0;
    if (i < m) {
      // (eight loads in flight, the sum in order)
      const double* col = L.Minv + i;
      int j = j0;
      for (; j + 8 <= j1; j += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ldg(col + size_t(j + u) * size_t(m));
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u] * sb[j + u];
      }
      for (; j < j1; ++j) s += ldg(col + size_t(j) * size_t(m)) * sb[j];
    }
    part[sp * 64 + lane] = s;
  }
  __syncthreads();
  if (w == 0 && i < m) {
    double t = part[lane];
    for (int q = 1; q < kDenseSpans; ++q) t += part[q * 64 + lane];
    a.x[1][i] = T(t);
  }
  __syncthreads();
}
template <typename T>
__device__ __forceinline__ void dense_solve_g(const GraphDev& g, int level) {
  __shared__ double sb[kMaxDenseM];
  __shared__ double part[kDenseSpans * 64];
  const LevelDev& L = g.lv[level];
  for (int r0 = 0; r0 < L.dense_m; r0 += kDenseRows) dense_rows<T, kWG>(L, arrays<T>(L), r0, sb, part);
}
// one workgroup a chunk of 64 rows (tile: graph, first row, rows)
template <typename T>
__global__ void __launch_bounds__(kDenseSolveWG) k_dense_solve(const GraphDev* __restrict__ G,
                                                               const Tile* __restrict__ chunks, int count, int level,
                                                               int inner) {
  if (int(blockIdx.x) >= count) return;
  const Tile tl = chunks[blockIdx.x];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  __shared__ double sb[kMaxDenseM];
  __shared__ double part[kDenseSpans * 64];
  const LevelDev& L = g.lv[level];
  dense_rows<T, kDenseSolveWG>(L, arrays<T>(L), tl.row0, sb, part);
}

// The same product spread over more of the machine: one wavefront a (64-row chunk, span) pair, its
// span's partial sums to `part` (kDenseSpans x 64 doubles a chunk); the wavefront that finishes a
// chunk's last span adds the chunk's spans in span order, as dense_rows does: the same bits. For
// batches of few dense chunks (a single large graph: 14 chunks of a 864-row inverse would occupy 14
// of the 256 compute units).
template <typename T>
__global__ void __launch_bounds__(64) k_dense_solve_span(const GraphDev* __restrict__ G, const Tile* __restrict__ chunks,
                                                         int count, int level, int inner, double* __restrict__ part,
                                                         int32_t* __restrict__ cnt) {
  const int c = int(blockIdx.x) / kDenseSpans, sp = int(blockIdx.x) % kDenseSpans;
  if (c >= count) return;
  const Tile tl = chunks[c];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  const LevelDev& L = g.lv[level];
  const LevelArrays<T>& a = arrays<T>(L);
  const int m = L.dense_m;
  const int span = (m + kDenseSpans - 1) / kDenseSpans;
  const int j0 = sp * span, j1 = min(m, j0 + span);
  __shared__ double sb[(kMaxDenseM + kDenseSpans - 1) / kDenseSpans];
  for (int j = j0 + int(threadIdx.x); j < j1; j += 64) sb[j - j0] = double(a.b[j]);
  __syncthreads();
  const int lane = threadIdx.x, i = tl.row0 + lane;
  double s = 0.0;
  if (i < m) {
    // (32 loads in flight, the sum in order)
    const double* col = L.Minv + i;
    for (int jb = j0; jb < j1; jb += 32) {
      double v[32];
#pragma unroll
      for (int u = 0; u < 32; ++u) v[u] = jb + u < j1 ? ldg(col + size_t(jb + u) * size_t(m)) : 0.0;
#pragma unroll
      for (int u = 0; u < 32; ++u)
        if (jb + u < j1) s += v[u] * sb[jb + u - j0];
    }
  }
  double* pc = part + size_t(c) * kDenseSpans * 64;
  pc[sp * 64 + lane] = s;
  __shared__ int last;
  if (threadIdx.x == 0) {
    __threadfence();
    const int old = atomicAdd(cnt + c, 1);
    last = old == kDenseSpans - 1 ? 1 : 0;
    if (last) {
      cnt[c] = 0;
      __threadfence();
    }
  }
  __syncthreads();
  if (!last || i >= m) return;
  double t = pc[lane];
  for (int q = 1; q < kDenseSpans; ++q) t += pc[q * 64 + lane];
  a.x[1][i] = T(t);
}

// z = D^-1 r (block-Jacobi preconditioning of a graph without a hierarchy)
template <typename T>
__device__ __forceinline__ void jacobi_tile(const GraphDev& g, const Tile& tl, int inner) {
  const LevelArrays<T>& a = arrays<T>(g.lv[0]);
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int row = tl.row0 + (on ? rl : 0);
  const size_t e = 6 * size_t(row) + q;
  const T r = on ? a.b[e] : T(0);
  T rv[6];
  row_gather(r, rv);
  if (on) a.x[1][e] = block_row<T>(a.Dinv + 36 * size_t(row), rv, q);

}
template <typename T>
__global__ void __launch_bounds__(kWG) k_jacobi(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                int ntiles, int inner) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  jacobi_tile<T>(g, tl, inner);
}

// ---- PCG kernels (level 0 tiles) -------------------------------------------------------------

template <typename T>
struct PcgVecs {
  T *x, *r, *z, *p, *q;
};
template <typename T>
__device__ __forceinline__ PcgVecs<T> pcg_vecs(const GraphDev& g);
template <>
__device__ __forceinline__ PcgVecs<double> pcg_vecs<double>(const GraphDev& g) {
  return {g.x, g.r, g.z, g.p, g.q};
}
template <>
__device__ __forceinline__ PcgVecs<float> pcg_vecs<float>(const GraphDev& g) {
  return {g.ix, g.ir, g.iz, g.ip, g.iq};
}

// q = K p, partial p . q
template <typename T>
__device__ __forceinline__ void pcg_spmv_tile(const GraphDev& g, const Tile& tl, int inner, double* __restrict__ partial) {
  const LevelDev& L = g.lv[0];
  const PcgVecs<T> v = pcg_vecs<T>(g);
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int row = tl.row0 + (on ? rl : 0);
  const size_t e = 6 * size_t(row) + q;
  const T pe = on ? v.p[e] : T(0);
  const T y = lpb_rows<T, false>(L.rowptr, L.col, arrays<T>(L).A, v.p, tl.row0, tl.nrows, on, row, q);
  double s = 0.0;
  if (on) {
    v.q[e] = y;
    s = double(pe) * double(y);
  }
  finish_partials(s, 0.0, partial, tl.slot, false);

}
template <typename T>
__global__ void __launch_bounds__(kWG) k_pcg_spmv(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                  int ntiles, int inner, double* __restrict__ partial, ProfAcc pa) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  if (pa.acc && tl.row0 == 0 && threadIdx.x == 0) atomicAdd(pa.acc, pa.gb[tl.g]);
  pcg_spmv_tile<T>(g, tl, inner, partial);
}

// FP64-product inner solves (mixed precision): the FP32 V-cycle's input r_i = (float) r is written
// with r, and its output z_i is read as (double) z_i wherever z is, so that no conversion passes run
template <typename T>
__device__ __forceinline__ bool inner64_mode(int inner) {
  return std::is_same<T, double>::value && (inner & SERVE_P64) != 0;
}
template <typename T>
__device__ __forceinline__ T z_at(const GraphDev& g, const PcgVecs<T>& v, size_t e, int inner) {
  if (inner64_mode<T>(inner)) return T(g.iz[e]);
  return v.z[e];
}

// x += alpha p, r -= alpha q, partial r . r
template <typename T>
__device__ __forceinline__ void axpy_rr_tile(const GraphDev& g, const Tile& tl, int inner, double* __restrict__ partial) {
  const PcgVecs<T> v = pcg_vecs<T>(g);
  const double alpha = g.st->alpha;
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  double s = 0.0;
  if (on) {
    const size_t e = 6 * size_t(tl.row0 + rl) + q;
    v.x[e] = T(double(v.x[e]) + alpha * double(v.p[e]));
    const T r = T(double(v.r[e]) - alpha * double(v.q[e]));
    v.r[e] = r;
    if (inner64_mode<T>(inner)) g.ir[e] = float(r);
    s = double(r) * double(r);
  }
  finish_partials(s, 0.0, partial, tl.slot, false);

}
template <typename T>
__global__ void __launch_bounds__(kWG) k_axpy_rr(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                 int ntiles, int inner, double* __restrict__ partial) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  axpy_rr_tile<T>(g, tl, inner, partial);
}

// p = z + beta p
template <typename T>
__device__ __forceinline__ void update_p_tile(const GraphDev& g, const Tile& tl, int inner) {
  const PcgVecs<T> v = pcg_vecs<T>(g);
  const double beta = g.st->beta;
  int rl, q;
  lane_row(rl, q);
  if (rl < 0 || rl >= tl.nrows) return;
  const size_t e = 6 * size_t(tl.row0 + rl) + q;
  v.p[e] = T(double(z_at<T>(g, v, e, inner)) + beta * double(v.p[e]));

}
template <typename T>
__global__ void __launch_bounds__(kWG) k_update_p(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                  int ntiles, int inner) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  update_p_tile<T>(g, tl, inner);
}

// partial r . z (after the preconditioner)
template <typename T>
__device__ __forceinline__ void dot_rz_tile(const GraphDev& g, const Tile& tl, int inner, double* __restrict__ partial) {
  const PcgVecs<T> v = pcg_vecs<T>(g);
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  double s = 0.0;
  if (on) {
    const size_t e = 6 * size_t(tl.row0 + rl) + q;
    s = double(v.r[e]) * double(z_at<T>(g, v, e, inner));
  }
  finish_partials(s, 0.0, partial, tl.slot, false);

}
template <typename T>
__global__ void __launch_bounds__(kWG) k_dot_rz(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                int ntiles, int inner, double* __restrict__ partial) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  dot_rz_tile<T>(g, tl, inner, partial);
}

// partial r . r
template <typename T>
__device__ __forceinline__ void dot_rr_tile(const GraphDev& g, const Tile& tl, int inner, double* __restrict__ partial) {
  const PcgVecs<T> v = pcg_vecs<T>(g);
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  double s = 0.0;
  if (on) {
    const T r = v.r[6 * size_t(tl.row0 + rl) + q];
    s = double(r) * double(r);
  }
  finish_partials(s, 0.0, partial, tl.slot, false);

}
template <typename T>
__global__ void __launch_bounds__(kWG) k_dot_rr(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                int ntiles, int inner, double* __restrict__ partial) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  dot_rr_tile<T>(g, tl, inner, partial);
}

// y = A x over one matrix (no graph table): tests and the bandwidth probe
template <typename T>
__global__ void __launch_bounds__(kWG) k_spmv_plain(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                    const T* __restrict__ A, const T* __restrict__ x, T* __restrict__ y,
                                                    int n) {
  const int ntiles = (n + kTileRows - 1) / kTileRows;
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  int rl, q;
  lane_row(rl, q);
  const int row0 = t * kTileRows, nrows = min(kTileRows, n - row0);
  const bool on = rl >= 0 && rl < nrows;
  const int row = row0 + (on ? rl : 0);
  const T v = lpb_rows<T, false>(rowptr, col, A, x, row0, nrows, on, row, q);
  if (on) y[6 * size_t(row) + q] = v;
}

// ---- start of a solve ------------------------------------------------------------------------

// r = b - K x (FP64), partials b . b and r . r
__device__ __forceinline__ void begin_residual_tile(const GraphDev& g, const Tile& tl, double* __restrict__ partial) {
  const LevelDev& L = g.lv[0];
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int row = tl.row0 + (on ? rl : 0);
  const size_t e = 6 * size_t(row) + q;
  const double b = on ? g.b[e] : 0.0;
  const double kx = lpb_rows<double, false>(L.rowptr, L.col, L.f64.A, g.x, tl.row0, tl.nrows, on, row, q);
  double sb = 0.0, sr = 0.0;
  if (on) {
    const double r = b - kx;
    g.r[e] = r;
    sb = b * b;
    sr = r * r;
  }
  finish_partials(sb, sr, partial, tl.slot, true);

}
__global__ void __launch_bounds__(kWG) k_begin_residual(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                        int ntiles, double* __restrict__ partial) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, SERVE_NEW)) return;
  begin_residual_tile(g, tl, partial);
}

// the solve's start, per graph: |b|, a zero or non-finite load (x = 0, r = 0); a mixed-precision
// graph goes on to its first refinement step
__device__ __forceinline__ void finish_begin_g(const GraphDev& g, int k0, int k1, int mixed, const double* __restrict__ partial) {
  PcgState* s = g.st;
  if (!live(g, SERVE_NEW)) return;
  const double bb = graph_sum(k0, k1, partial, 0);
  if (threadIdx.x == 0) {
    s->bn = sqrt(bb);
    s->target = s->rtol * s->bn;
    s->zero_load = (bb > 0.0) ? 0 : 1;
    if (mixed) s->phase = PH_OUTER;
  }
}
__global__ void __launch_bounds__(kWG) k_finish_begin(const GraphDev* __restrict__ G, const int32_t* __restrict__ gtile,
                                                      int count, int mixed, const double* __restrict__ partial) {
  if (int(blockIdx.x) >= count || G[blockIdx.x].resident) return;
  finish_begin_g(G[blockIdx.x], gtile[blockIdx.x], gtile[blockIdx.x + 1], mixed, partial);
}

// x = 0, r = 0 where the load is zero or not finite
__device__ __forceinline__ void zero_load_tile(const GraphDev& g, const Tile& tl) {
  int rl, q;
  lane_row(rl, q);
  if (rl < 0 || rl >= tl.nrows) return;
  const size_t e = 6 * size_t(tl.row0 + rl) + q;
  g.x[e] = 0.0;
  g.r[e] = 0.0;

}
__global__ void __launch_bounds__(kWG) k_zero_load(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                   int ntiles) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || g.st->state != ST_RUNNING || !g.st->zero_load || (g.st->phase != PH_NEW && g.st->phase != PH_OUTER)) return;
  zero_load_tile(g, tl);
}

// p = z, partials r . z and r . r (the start of the conjugate directions)
template <typename T>
__device__ __forceinline__ void start_p_tile(const GraphDev& g, const Tile& tl, int inner, double* __restrict__ partial) {
  const PcgVecs<T> v = pcg_vecs<T>(g);
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  double s0 = 0.0, s1 = 0.0;
  if (on) {
    const size_t e = 6 * size_t(tl.row0 + rl) + q;
    const T z = z_at<T>(g, v, e, inner), r = v.r[e];
    v.p[e] = z;
    s0 = double(r) * double(z);
    s1 = double(r) * double(r);
  }
  finish_partials(s0, s1, partial, tl.slot, true);

}
template <typename T>
__global__ void __launch_bounds__(kWG) k_start_p(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                 int ntiles, int inner, double* __restrict__ partial) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, inner)) return;
  start_p_tile<T>(g, tl, inner, partial);
}

// ---- per-graph finishes (one workgroup a graph; partials summed in tile order) ----------------


// an inner solve ends (PH_APPLY); one with FP64 products updated x itself, so it closes at once: its
// iterations count towards the solve's and the graph goes on to its next refinement step (what
// inner_close_g does after the FP32-product corrections are applied)
__device__ __forceinline__ void inner_end(PcgState* s) {
  s->phase = PH_APPLY;
  if (!s->inner64) return;
  s->iters += s->inner_iters;
  if (s->inner_state == ST_BREAKDOWN && s->inner_iters == 0) s->state = ST_BREAKDOWN;
  s->inner_state = ST_IDLE;
  s->phase = PH_OUTER;
}

__device__ __forceinline__ PcgState* finish_state(const GraphDev& g, int mode, bool& go) {
  go = live(g, mode);
  return g.st;
}
__device__ __forceinline__ bool inner_mode(int mode) { return (mode & SERVE_INNER) != 0; }

// after p = z: rz, the first residual check (StressProblem::begin / the head of iterate); the
// graph goes on to iterate
__device__ __forceinline__ void finish_start_g(const GraphDev& g, int k0, int k1, int mode, const double* __restrict__ partial) {
  bool go;
  PcgState* s = finish_state(g, mode, go);
  if (!go) return;
  const double rz = graph_sum(k0, k1, partial, 0);
  const double rr = graph_sum(k0, k1, partial, 1);
  if (threadIdx.x != 0) return;
  s->rz = rz;
  const double rn = sqrt(rr);
  if (inner_mode(mode)) {
    s->inner_iters = 0;
    s->inner_rn = rn;
    s->phase = PH_INNER;
    if (!(rn > s->inner_target)) {
      s->inner_state = ST_CONVERGED;
      if (s->inner64) {
        s->rn = rn;
        s->rel_res = rn / s->bn;
        s->state = ST_CONVERGED;
      }
      inner_end(s);
    }
    return;
  }
  s->phase = PH_ITER;
  s->rn = rn;
  if (!isfinite(s->bn) || !isfinite(rn)) {
    s->rel_res = INFINITY;
    s->state = ST_BREAKDOWN;
    return;
  }
  s->rel_res = s->bn > 0 ? rn / s->bn : 0.0;
  if (rn <= s->target || !(s->bn > 0) || g.nfree == 0) s->state = ST_CONVERGED;
}
__global__ void __launch_bounds__(kWG) k_finish_start(const GraphDev* __restrict__ G, const int32_t* __restrict__ gtile,
                                         int count, int mode, const double* __restrict__ partial) {
  if (int(blockIdx.x) >= count || G[blockIdx.x].resident) return;
  finish_start_g(G[blockIdx.x], gtile[blockIdx.x], gtile[blockIdx.x + 1], mode, partial);
}

__device__ __forceinline__ void finish_alpha_g(const GraphDev& g, int k0, int k1, int mode, const double* __restrict__ partial) {
  bool go;
  PcgState* s = finish_state(g, mode, go);
  if (!go) return;
  const double pq = graph_sum(k0, k1, partial, 0);
  if (threadIdx.x != 0) return;
  // (the iteration budget is checked at an iteration's head: a stopped run resumes where it was)
  if (inner_mode(mode)) {
    if (s->inner_iters >= s->inner_maxit) {
      s->inner_state = ST_STOPPED;
      inner_end(s);
      return;
    }
    ++s->inner_iters;
    if (!(pq > 0)) {
      s->inner_state = ST_BREAKDOWN;
      inner_end(s);
      return;
    }
  } else {
    if (s->iters >= s->maxit) {
      s->state = ST_STOPPED;
      return;
    }
    ++s->iters;
    if (!(pq > 0)) {
      s->state = ST_BREAKDOWN;
      return;
    }
  }
  s->pq = pq;
  s->alpha = s->rz / pq;
}
__global__ void __launch_bounds__(kWG) k_finish_alpha(const GraphDev* __restrict__ G, const int32_t* __restrict__ gtile,
                                         int count, int mode, const double* __restrict__ partial) {
  if (int(blockIdx.x) >= count || G[blockIdx.x].resident) return;
  finish_alpha_g(G[blockIdx.x], gtile[blockIdx.x], gtile[blockIdx.x + 1], mode, partial);
}

__device__ __forceinline__ void finish_rn_g(const GraphDev& g, int k0, int k1, int mode, const double* __restrict__ partial) {
  bool go;
  PcgState* s = finish_state(g, mode, go);
  if (!go) return;
  const double rr = graph_sum(k0, k1, partial, 0);
  if (threadIdx.x != 0) return;
  const double rn = sqrt(rr);
  if (inner_mode(mode)) {
    s->inner_rn = rn;
    if (rn <= s->inner_target) {
      s->inner_state = ST_CONVERGED;
      if (s->inner64) {
        s->rn = rn;
        s->rel_res = rn / s->bn;
        s->state = ST_CONVERGED;
      }
      inner_end(s);
    }
    return;
  }
  s->rn = rn;
  s->rel_res = rn / s->bn;
  if (rn <= s->target) s->state = ST_CONVERGED;
}
__global__ void __launch_bounds__(kWG) k_finish_rn(const GraphDev* __restrict__ G, const int32_t* __restrict__ gtile,
                                         int count, int mode, const double* __restrict__ partial) {
  if (int(blockIdx.x) >= count || G[blockIdx.x].resident) return;
  finish_rn_g(G[blockIdx.x], gtile[blockIdx.x], gtile[blockIdx.x + 1], mode, partial);
}

__device__ __forceinline__ void finish_beta_g(const GraphDev& g, int k0, int k1, int mode, const double* __restrict__ partial) {
  bool go;
  PcgState* s = finish_state(g, mode, go);
  if (!go) return;
  const double rz1 = graph_sum(k0, k1, partial, 0);
  if (threadIdx.x != 0) return;
  if (!(rz1 > 0)) {
    if (inner_mode(mode)) {
      s->inner_state = ST_BREAKDOWN;
      inner_end(s);
    } else {
      s->state = ST_BREAKDOWN;
    }
    return;
  }
  s->beta = rz1 / s->rz;
  s->rz = rz1;
}
__global__ void __launch_bounds__(kWG) k_finish_beta(const GraphDev* __restrict__ G, const int32_t* __restrict__ gtile,
                                         int count, int mode, const double* __restrict__ partial) {
  if (int(blockIdx.x) >= count || G[blockIdx.x].resident) return;
  finish_beta_g(G[blockIdx.x], gtile[blockIdx.x], gtile[blockIdx.x + 1], mode, partial);
}

// ---- mixed precision: the refinement steps --------------------------------------------------

// r = b - K x in FP64, partial r . r (the refinement's true residual)
__device__ __forceinline__ void outer_residual_tile(const GraphDev& g, const Tile& tl, double* __restrict__ partial) {
  const LevelDev& L = g.lv[0];
  int rl, q;
  lane_row(rl, q);
  const bool on = rl >= 0 && rl < tl.nrows;
  const int row = tl.row0 + (on ? rl : 0);
  const size_t e = 6 * size_t(row) + q;
  const double b = on ? g.b[e] : 0.0;
  const double kx = lpb_rows<double, false>(L.rowptr, L.col, L.f64.A, g.x, tl.row0, tl.nrows, on, row, q);
  double sr = 0.0;
  if (on) {
    const double r = b - kx;
    g.r[e] = r;
    sr = r * r;
  }
  finish_partials(sr, 0.0, partial, tl.slot, false);

}
__global__ void __launch_bounds__(kWG) k_outer_residual(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                        int ntiles, double* __restrict__ partial) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, SERVE_REFINE)) return;
  outer_residual_tile(g, tl, partial);
}

// outer check: converged, or an inner solve of K d = r to max(eta |r|, target / 2)
__device__ __forceinline__ void finish_outer_g(const GraphDev& g, int k0, int k1, const double* __restrict__ partial) {
  PcgState* s = g.st;
  if (!live(g, SERVE_REFINE)) return;
  const double rr = graph_sum(k0, k1, partial, 0);
  if (threadIdx.x != 0) return;
  const double rn = sqrt(rr);
  s->rn = rn;
  s->outer_rn = rn;
  if (!isfinite(s->bn) || !isfinite(rn)) {
    s->rel_res = INFINITY;
    s->state = ST_BREAKDOWN;
    return;
  }
  s->rel_res = s->bn > 0 ? rn / s->bn : 0.0;
  if (rn <= s->target || !(s->bn > 0) || g.nfree == 0) {
    s->state = ST_CONVERGED;
    return;
  }
  if (s->iters >= s->maxit) {
    s->state = ST_STOPPED;
    return;
  }
  // a step that missed its goal tenfold: FP32 K's error times K's condition number stalls this graph;
  // its inner solves take FP64 K products from now on and stop, as the plain FP64 PCG does, on their
  // recurrence residual at the target (it waits until the host launches FP64 products)
  const bool to64 = s->outer > 0 && rn > 10.0 * s->inner_target && !s->inner64;
  if (to64) s->inner64 = 1;
  s->outer_rn_prev = rn;
  s->inner_target = s->inner64 ? s->target : fmax(s->eta * rn, 0.5 * s->target);
  s->inner_state = ST_RUNNING;
  s->inner_iters = 0;
  s->inner_maxit = s->maxit - s->iters;
  s->phase = to64 ? PH_WAIT64 : PH_START;
  ++s->outer;
}
__global__ void __launch_bounds__(kWG) k_finish_outer(const GraphDev* __restrict__ G, const int32_t* __restrict__ gtile,
                                                      int count, const double* __restrict__ partial) {
  if (int(blockIdx.x) >= count || G[blockIdx.x].resident) return;
  finish_outer_g(G[blockIdx.x], gtile[blockIdx.x], gtile[blockIdx.x + 1], partial);
}

// graphs waiting for FP64-product inner solves start them (the host launches those from now on)
__global__ void k_release_wait64(const GraphDev* __restrict__ G, int count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  PcgState* s = G[k].st;
  if (s->state == ST_RUNNING && s->phase == PH_WAIT64) s->phase = PH_START;
}

// inner start: r_i = (float) r, x_i = 0
__device__ __forceinline__ void inner_init_tile(const GraphDev& g, const Tile& tl) {
  int rl, q;
  lane_row(rl, q);
  if (rl < 0 || rl >= tl.nrows) return;
  const size_t e = 6 * size_t(tl.row0 + rl) + q;
  g.ir[e] = float(g.r[e]);
  g.ix[e] = 0.0f;

}
__global__ void __launch_bounds__(kWG) k_inner_init(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                    int ntiles) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || !live(g, SERVE_START)) return;
  inner_init_tile(g, tl);
}

// x += (double) x_i after an inner solve; the inner iterations count towards the solve's
__device__ __forceinline__ void inner_apply_tile(const GraphDev& g, const Tile& tl) {
  // (an FP64-product inner solve updated x itself)
  int rl, q;
  lane_row(rl, q);
  if (rl < 0 || rl >= tl.nrows) return;
  const size_t e = 6 * size_t(tl.row0 + rl) + q;
  g.x[e] += double(g.ix[e]);

}
__global__ void __launch_bounds__(kWG) k_inner_apply(const GraphDev* __restrict__ G, const Tile* __restrict__ tiles,
                                                     int ntiles) {
  const int t = xcd_tile(ntiles);
  if (t >= ntiles) return;
  const Tile tl = tiles[t];
  const GraphDev& g = G[tl.g];
  if (g.resident || g.st->state != ST_RUNNING || g.st->phase != PH_APPLY || g.st->inner64) return;
  inner_apply_tile(g, tl);
}

// the inner iterations count towards the solve's; the graph goes on to its next refinement step
__device__ __forceinline__ void inner_close_g(const GraphDev& g) {
  PcgState* s = g.st;
  if (s->phase != PH_APPLY) return;
  s->iters += s->inner_iters;
  if (s->inner_state == ST_BREAKDOWN && s->inner_iters == 0) s->state = ST_BREAKDOWN;
  s->inner_state = ST_IDLE;
  s->phase = PH_OUTER;
}
__global__ void k_inner_close(const GraphDev* __restrict__ G, int count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count || G[k].resident) return;
  inner_close_g(G[k]);
}

// ---- resident graphs: whole steps in one workgroup ------------------------------------------

// tile k of n rows, `per` rows a tile, partial slots from slot0
__device__ __forceinline__ Tile tile_of(int gi, int k, int n, int per, int slot0) {
  const int r0 = k * per;
  return Tile{gi, r0, min(per, n - r0), slot0 + k};
}

template <typename T, int S>
__device__ void smooth_g(const GraphDev& g, int gi, int l, int st, int in_buf, int out_buf, bool fresh, int mode) {
  const LevelDev& L = g.lv[l];
  const int per = tile_rows<S>(), nt = (L.n + per - 1) / per;
  for (int k = 0; k < nt; ++k) {
    const Tile tl = tile_of(gi, k, L.n, per, 0);
    if (fresh) cheb_step_tile<T, S, true>(g, tl, l, st, in_buf, out_buf, mode, nullptr);
    else cheb_step_tile<T, S, false>(g, tl, l, st, in_buf, out_buf, mode, nullptr);
  }
  __syncthreads();
}

template <typename T>
__device__ void smooth_any(const GraphDev& g, int gi, int l, int st, int in_buf, int out_buf, bool fresh, int mode) {
  if (g.lv[l].sub == 10) smooth_g<T, 10>(g, gi, l, st, in_buf, out_buf, fresh, mode);
  else smooth_g<T, 1>(g, gi, l, st, in_buf, out_buf, fresh, mode);
}

template <typename T, int S>
__device__ void residual_g(const GraphDev& g, int gi, int l, int buf, int mode) {
  const LevelDev& L = g.lv[l];
  const int per = tile_rows<S>(), nt = (L.n + per - 1) / per;
  for (int k = 0; k < nt; ++k) residual_tile<T, S>(g, tile_of(gi, k, L.n, per, 0), l, buf, mode);
  __syncthreads();
}

template <typename T, int S>
__device__ void restrict_g(const GraphDev& g, int gi, int l, int mode) {
  const int nc = g.lv[l].nc;
  const int per = tile_rows<S>(), nt = (nc + per - 1) / per;
  for (int k = 0; k < nt; ++k) restrict_tile<T, S>(g, tile_of(gi, k, nc, per, 0), l, mode);
  __syncthreads();
}

// the V-cycle of one graph (vcycle<T> of the host driver, restricted to this graph's levels;
// `degree` packs the smoother's degrees, level_degree)
template <typename T>
__device__ void vcycle_g(const GraphDev& g, int gi, int mode, int degree) {
  if (g.jacobi) {
    const int nt = (g.nfree + kTileRows - 1) / kTileRows;
    for (int k = 0; k < nt; ++k) jacobi_tile<T>(g, tile_of(gi, k, g.nfree, kTileRows, 0), mode);
    __syncthreads();
    return;
  }
  for (int l = 0; l < g.nlev; ++l) {
    const LevelDev& L = g.lv[l];
    const int dg = level_degree(degree, l);
    if (!L.terminal) {
      for (int st = 0; st < dg; ++st) smooth_any<T>(g, gi, l, st, (st + 1) & 1, st & 1, st == 0, mode);
      if (L.sub == 10) residual_g<T, 10>(g, gi, l, (dg - 1) & 1, mode);
      else residual_g<T, 1>(g, gi, l, (dg - 1) & 1, mode);
      if (L.rsub == 10 && !std::is_same<T, float>::value) restrict_g<T, 10>(g, gi, l, mode);
      else restrict_g<T, 1>(g, gi, l, mode);
    } else if (L.dense_m > 0) {
      dense_solve_g<T>(g, l);
      __syncthreads();
    } else {
      for (int st = 0; st < kStallDegree; ++st) smooth_any<T>(g, gi, l, st, (st + 1) & 1, st & 1, st == 0, mode);
    }
  }
  for (int l = g.nlev - 2; l >= 0; --l) {
    const LevelDev& L = g.lv[l];
    const int dg = level_degree(degree, l);
    const int nt = (L.n + kTileRows - 1) / kTileRows;
    for (int k = 0; k < nt; ++k) prolong_tile<T>(g, tile_of(gi, k, L.n, kTileRows, 0), l, (dg - 1) & 1, mode);
    __syncthreads();
    for (int j = 0; j < dg; ++j) {
      const int gs = dg + j;
      smooth_any<T>(g, gi, l, j, (gs + 1) & 1, gs & 1, false, mode);
    }
  }
}

// the graph's level-0 tiles (kTileRows rows; partial slots k0 ..)
#define SVXG_T0_LOOP(call)                                                          \
  do {                                                                               \
    for (int k = 0; k < nt0; ++k) {                                                  \
      const Tile tl = tile_of(gi, k, g.nfree, kTileRows, k0);                       \
      call;                                                                          \
    }                                                                                \
    __syncthreads();                                                                 \
  } while (0)

__global__ void __launch_bounds__(kWG) k_resident_steps(const GraphDev* __restrict__ G, const int32_t* __restrict__ gtile,
                                                        const int32_t* __restrict__ graphs, int count, int nsteps,
                                                        int degree, int mixed, int p32, int p64,
                                                        double* __restrict__ partial) {
  if (int(blockIdx.x) >= count) return;
  const int gi = graphs[blockIdx.x];
  const GraphDev& g = G[gi];
  const int k0 = gtile[gi], k1 = gtile[gi + 1], nt0 = k1 - k0;
  for (int step = 0; step < nsteps; ++step) {
    if (g.st->state != ST_RUNNING) return;
    // a solve's start (run_begin)
    if (live(g, SERVE_NEW)) {
      SVXG_T0_LOOP(begin_residual_tile(g, tl, partial));
      finish_begin_g(g, k0, k1, mixed, partial);
      __syncthreads();
      SVXG_T0_LOOP(if (g.st->state == ST_RUNNING && g.st->zero_load &&
                       (g.st->phase == PH_NEW || g.st->phase == PH_OUTER)) zero_load_tile(g, tl));
      if (!mixed) {
        vcycle_g<double>(g, gi, SERVE_NEW, degree);
        SVXG_T0_LOOP(if (live(g, SERVE_NEW)) start_p_tile<double>(g, tl, SERVE_NEW, partial));
        finish_start_g(g, k0, k1, SERVE_NEW, partial);
        __syncthreads();
      }
    }
    if (!mixed) {
      // fp64_step
      if (!live(g, SERVE_OUTER)) continue;
      SVXG_T0_LOOP(pcg_spmv_tile<double>(g, tl, SERVE_OUTER, partial));
      finish_alpha_g(g, k0, k1, SERVE_OUTER, partial);
      __syncthreads();
      if (live(g, SERVE_OUTER)) SVXG_T0_LOOP(axpy_rr_tile<double>(g, tl, SERVE_OUTER, partial));
      finish_rn_g(g, k0, k1, SERVE_OUTER, partial);
      __syncthreads();
      if (live(g, SERVE_OUTER)) vcycle_g<double>(g, gi, SERVE_OUTER, degree);
      if (live(g, SERVE_OUTER)) SVXG_T0_LOOP(dot_rz_tile<double>(g, tl, SERVE_OUTER, partial));
      finish_beta_g(g, k0, k1, SERVE_OUTER, partial);
      __syncthreads();
      if (live(g, SERVE_OUTER)) SVXG_T0_LOOP(update_p_tile<double>(g, tl, SERVE_OUTER));
      continue;
    }
    // mixed_step
    if (live(g, SERVE_REFINE)) {
      SVXG_T0_LOOP(outer_residual_tile(g, tl, partial));
      finish_outer_g(g, k0, k1, partial);
      __syncthreads();
    }
    if (live(g, SERVE_START)) SVXG_T0_LOOP(inner_init_tile(g, tl));
    if (p32 && live(g, SERVE_P32 | SERVE_IT)) SVXG_T0_LOOP(pcg_spmv_tile<float>(g, tl, SERVE_P32 | SERVE_IT, partial));
    if (p64 && live(g, SERVE_P64 | SERVE_IT)) SVXG_T0_LOOP(pcg_spmv_tile<double>(g, tl, SERVE_P64 | SERVE_IT, partial));
    finish_alpha_g(g, k0, k1, SERVE_ITER, partial);
    __syncthreads();
    if (p32 && live(g, SERVE_P32 | SERVE_IT)) SVXG_T0_LOOP(axpy_rr_tile<float>(g, tl, SERVE_P32 | SERVE_IT, partial));
    if (p64 && live(g, SERVE_P64 | SERVE_IT)) SVXG_T0_LOOP(axpy_rr_tile<double>(g, tl, SERVE_P64 | SERVE_IT, partial));
    finish_rn_g(g, k0, k1, SERVE_ITER, partial);
    __syncthreads();
    if (live(g, SERVE_INNER)) vcycle_g<float>(g, gi, SERVE_INNER, degree);
    if (p32 && live(g, SERVE_P32 | SERVE_IT)) SVXG_T0_LOOP(dot_rz_tile<float>(g, tl, SERVE_P32 | SERVE_IT, partial));
    if (p64 && live(g, SERVE_P64 | SERVE_IT)) SVXG_T0_LOOP(dot_rz_tile<double>(g, tl, SERVE_P64 | SERVE_IT, partial));
    finish_beta_g(g, k0, k1, SERVE_ITER, partial);
    __syncthreads();
    if (p32 && live(g, SERVE_P32 | SERVE_IT)) SVXG_T0_LOOP(update_p_tile<float>(g, tl, SERVE_P32 | SERVE_IT));
    if (p64 && live(g, SERVE_P64 | SERVE_IT)) SVXG_T0_LOOP(update_p_tile<double>(g, tl, SERVE_P64 | SERVE_IT));
    if (p32 && live(g, SERVE_P32 | SERVE_ST)) SVXG_T0_LOOP(start_p_tile<float>(g, tl, SERVE_P32 | SERVE_ST, partial));
    if (p64 && live(g, SERVE_P64 | SERVE_ST)) SVXG_T0_LOOP(start_p_tile<double>(g, tl, SERVE_P64 | SERVE_ST, partial));
    finish_start_g(g, k0, k1, SERVE_START, partial);
    __syncthreads();
    if (g.st->state == ST_RUNNING && g.st->phase == PH_APPLY && !g.st->inner64) SVXG_T0_LOOP(inner_apply_tile(g, tl));
    if (threadIdx.x == 0) inner_close_g(g);
    __syncthreads();
  }
}
#undef SVXG_T0_LOOP

// the current residual of each listed resident graph (the head of a resumed FP64 iterate)
__global__ void __launch_bounds__(kWG) k_resident_recheck(const GraphDev* __restrict__ G,
                                                          const int32_t* __restrict__ gtile,
                                                          const int32_t* __restrict__ graphs, int count,
                                                          double* __restrict__ partial) {
  if (int(blockIdx.x) >= count) return;
  const int gi = graphs[blockIdx.x];
  const GraphDev& g = G[gi];
  if (!live(g, SERVE_OUTER)) return;
  const int k0 = gtile[gi], k1 = gtile[gi + 1];
  for (int k = 0; k < k1 - k0; ++k) dot_rr_tile<double>(g, tile_of(gi, k, g.nfree, kTileRows, k0), SERVE_OUTER, partial);
  __syncthreads();
  finish_rn_g(g, k0, k1, SERVE_OUTER, partial);
}

// ---- host launchers ----------------------------------------------------------------------------

int grid_for(int ntiles) {
  if (ntiles < 64) return ntiles;
  return 8 * ((ntiles + 7) / 8);
}

template <typename T>
void launch_cheb_step(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int level, int step, int in_buf,
                      int out_buf, int fresh, int inner, double* partial, int S, ProfAcc pa, hipEvent_t t0,
                      hipEvent_t t1) {
  if (ntiles <= 0) return;
  const dim3 grid(grid_for(ntiles)), wg(kWG);
  if ((t0 || t1) && S == 1 && !fresh && !partial)
    hipExtLaunchKernelGGL((k_cheb_step<T, 1, false>), grid, wg, 0, s, t0, t1, 0, G, tiles, ntiles, level, step, in_buf,
                          out_buf, inner, (double*)nullptr, pa);
  else if (partial)
    hipLaunchKernelGGL(k_cheb_step_rz<T>, grid, wg, 0, s, G, tiles, ntiles, step, in_buf, out_buf, inner, partial);
  else if (S == 10 && fresh)
    hipLaunchKernelGGL((k_cheb_step<T, 10, true>), grid, wg, 0, s, G, tiles, ntiles, level, step, in_buf, out_buf, inner,
                       nullptr, pa);
  else if (S == 10)
    hipLaunchKernelGGL((k_cheb_step<T, 10, false>), grid, wg, 0, s, G, tiles, ntiles, level, step, in_buf, out_buf,
                       inner, nullptr, pa);
  else if (fresh)
    hipLaunchKernelGGL((k_cheb_step<T, 1, true>), grid, wg, 0, s, G, tiles, ntiles, level, step, in_buf, out_buf, inner,
                       nullptr, pa);
  else
    hipLaunchKernelGGL((k_cheb_step<T, 1, false>), grid, wg, 0, s, G, tiles, ntiles, level, step, in_buf, out_buf,
                       inner, nullptr, pa);
}
template <typename T>
void launch_residual(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int level, int buf, int inner,
                     int S) {
  if (ntiles <= 0) return;
  if (S == 10)
    hipLaunchKernelGGL((k_residual<T, 10>), dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, level, buf,
                       inner);
  else
    hipLaunchKernelGGL((k_residual<T, 1>), dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, level, buf, inner);
}
template <typename T>
void launch_restrict(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int level, int inner, int S) {
  if (ntiles <= 0) return;
  if (S == 10)
    hipLaunchKernelGGL((k_restrict<T, 10>), dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, level, inner);
  else
    hipLaunchKernelGGL((k_restrict<T, 1>), dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, level, inner);
}
template <typename T>
void launch_prolong(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int level, int buf, int inner) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_prolong<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, level, buf, inner);
}
template <typename T>
void launch_dense_solve(hipStream_t s, const GraphDev* G, const Tile* chunks, int count, int level, int inner,
                        double* part, int32_t* cnt) {
  if (count <= 0) return;
  if (part && cnt)
    hipLaunchKernelGGL(k_dense_solve_span<T>, dim3(count * kDenseSpans), dim3(64), 0, s, G, chunks, count, level, inner,
                       part, cnt);
  else
    hipLaunchKernelGGL(k_dense_solve<T>, dim3(count), dim3(kDenseSolveWG), 0, s, G, chunks, count, level, inner);
}
template <typename T>
void launch_jacobi(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_jacobi<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, inner);
}
template <typename T>
void launch_dot_rz(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_dot_rz<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, inner, partial);
}
template <typename T>
void launch_dot_rr(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_dot_rr<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, inner, partial);
}
template <typename T>
void launch_spmv_plain(hipStream_t s, const int32_t* rowptr, const int32_t* col, const T* A, const T* x, T* y, int n) {
  const int ntiles = (n + kTileRows - 1) / kTileRows;
  if (ntiles > 0)
    hipLaunchKernelGGL(k_spmv_plain<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, rowptr, col, A, x, y, n);
}
template <typename T>
void launch_pcg_spmv(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial,
                     ProfAcc pa, hipEvent_t t0, hipEvent_t t1) {
  if (ntiles <= 0) return;
  if (t0 || t1)
    hipExtLaunchKernelGGL(k_pcg_spmv<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, t0, t1, 0, G, tiles, ntiles, inner,
                          partial, pa);
  else
    hipLaunchKernelGGL(k_pcg_spmv<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, inner, partial, pa);
}
template <typename T>
void launch_axpy_rr(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_axpy_rr<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, inner, partial);
}
template <typename T>
void launch_update_p(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_update_p<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, inner);
}
template <typename T>
void launch_start_p(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_start_p<T>, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, inner, partial);
}

void launch_begin_residual(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, double* partial) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_begin_residual, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, partial);
}
void launch_release_wait64(hipStream_t s, const GraphDev* G, int count) {
  if (count > 0) hipLaunchKernelGGL(k_release_wait64, dim3((count + 255) / 256), dim3(256), 0, s, G, count);
}
void launch_finish_begin(hipStream_t s, const GraphDev* G, const int32_t* gtile, int count, int mixed,
                         const double* partial) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_finish_begin, dim3(count), dim3(kWG), 0, s, G, gtile, count, mixed, partial);
}
void launch_zero_load(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_zero_load, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles);
}
void launch_finish(hipStream_t s, FinishKind kind, const GraphDev* G, const int32_t* gtile, int count, int inner,
                   const double* partial) {
  if (count <= 0) return;
  switch (kind) {
    case FinishKind::Start:
      hipLaunchKernelGGL(k_finish_start, dim3(count), dim3(kWG), 0, s, G, gtile, count, inner, partial);
      break;
    case FinishKind::Alpha:
      hipLaunchKernelGGL(k_finish_alpha, dim3(count), dim3(kWG), 0, s, G, gtile, count, inner, partial);
      break;
    case FinishKind::Rn:
      hipLaunchKernelGGL(k_finish_rn, dim3(count), dim3(kWG), 0, s, G, gtile, count, inner, partial);
      break;
    case FinishKind::Beta:
      hipLaunchKernelGGL(k_finish_beta, dim3(count), dim3(kWG), 0, s, G, gtile, count, inner, partial);
      break;
    case FinishKind::Outer:
      hipLaunchKernelGGL(k_finish_outer, dim3(count), dim3(kWG), 0, s, G, gtile, count, partial);
      break;
  }
}
void launch_outer_residual(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, double* partial) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_outer_residual, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles, partial);
}
void launch_inner_init(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_inner_init, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles);
}
void launch_inner_apply(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_inner_apply, dim3(grid_for(ntiles)), dim3(kWG), 0, s, G, tiles, ntiles);
}
void launch_resident_steps(hipStream_t s, const GraphDev* G, const int32_t* gtile, const int32_t* graphs, int count,
                           int nsteps, int degree, int mixed, int p32, int p64, double* partial) {
  if (count <= 0 || nsteps <= 0) return;
  hipLaunchKernelGGL(k_resident_steps, dim3(count), dim3(kWG), 0, s, G, gtile, graphs, count, nsteps, degree, mixed,
                     p32, p64, partial);
}
void launch_resident_recheck(hipStream_t s, const GraphDev* G, const int32_t* gtile, const int32_t* graphs, int count,
                             double* partial) {
  if (count > 0) hipLaunchKernelGGL(k_resident_recheck, dim3(count), dim3(kWG), 0, s, G, gtile, graphs, count, partial);
}
void launch_inner_close(hipStream_t s, const GraphDev* G, int count) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_inner_close, dim3((count + 255) / 256), dim3(256), 0, s, G, count);
}

#define SVXG_INST(T)                                                                                                \
  template void launch_cheb_step<T>(hipStream_t, const GraphDev*, const Tile*, int, int, int, int, int, int, int,   \
                                    double*, int, ProfAcc, hipEvent_t, hipEvent_t);                                 \
  template void launch_residual<T>(hipStream_t, const GraphDev*, const Tile*, int, int, int, int, int);              \
  template void launch_restrict<T>(hipStream_t, const GraphDev*, const Tile*, int, int, int, int);                   \
  template void launch_prolong<T>(hipStream_t, const GraphDev*, const Tile*, int, int, int, int);                    \
  template void launch_dense_solve<T>(hipStream_t, const GraphDev*, const Tile*, int, int, int, double*, int32_t*); \
  template void launch_jacobi<T>(hipStream_t, const GraphDev*, const Tile*, int, int);                               \
  template void launch_dot_rz<T>(hipStream_t, const GraphDev*, const Tile*, int, int, double*);                      \
  template void launch_dot_rr<T>(hipStream_t, const GraphDev*, const Tile*, int, int, double*);                      \
  template void launch_spmv_plain<T>(hipStream_t, const int32_t*, const int32_t*, const T*, const T*, T*, int);      \
  template void launch_pcg_spmv<T>(hipStream_t, const GraphDev*, const Tile*, int, int, double*, ProfAcc, hipEvent_t, \
                                   hipEvent_t);                                                                     \
  template void launch_axpy_rr<T>(hipStream_t, const GraphDev*, const Tile*, int, int, double*);                     \
  template void launch_update_p<T>(hipStream_t, const GraphDev*, const Tile*, int, int);                             \
  template void launch_start_p<T>(hipStream_t, const GraphDev*, const Tile*, int, int, double*);
SVXG_INST(double)
SVXG_INST(float)

}  // namespace svxg
