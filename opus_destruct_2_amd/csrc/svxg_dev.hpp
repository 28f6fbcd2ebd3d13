// Device-side data layout of the svxg stress solve (shared by the kernels and the host driver).
//
// Every problem ("graph") owns its device arrays; a launch works on a batch of graphs through a
// table of per-graph descriptors (GraphDev) and a table of row tiles. A tile is up to kTileRows
// block rows of ONE graph, numbered from that graph's first row, so every per-tile partial sum,
// and the per-graph sums made from them in tile order, are the same whatever else is in the
// batch: a graph's answer is bitwise the same alone and in a batch.
//
// Lanes: a workgroup of 256 lanes is four 64-lane wavefronts; each wavefront serves 10 block rows
// with 6 lanes a row (lane q of a row owns DOF q of the row), lanes 60..63 idle. A row's 6 lanes
// never straddle a wavefront, so a row's values are exchanged with wavefront shuffles.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/svxg.h"

namespace svxg {

constexpr int kWG = 256;
constexpr int kRowsPerWave = 10;
constexpr int kTileRows = 40;  // 4 wavefronts x 10 rows
// the largest dense coarsest level (rows of its inverse): 6 x 192 nodes, the reference's limit at
// its default coarse_max (amg.cpp:535-541)
constexpr int kMaxDenseM = 1152;
constexpr int kMaxLevels = 16;

struct Tile {
  int32_t g;      // graph (index into the batch's GraphDev table)
  int32_t row0;   // first block row (the graph's own numbering)
  int32_t nrows;  // rows in this tile (<= kTileRows)
  int32_t slot;   // partial-sum slot (index of the tile within its table)
};

// per-graph solve state machine
enum : int32_t {
  ST_RUNNING = 0,
  ST_CONVERGED = 1,
  ST_BREAKDOWN = 2,
  ST_STOPPED = 3,   // maxit reached
  ST_IDLE = 4,      // not in this solve
};

struct PcgState {
  double rz, pq, alpha, beta, rn, bn, target, rtol;
  double rel_res;
  // mixed precision: outer residual norm and the inner solve's stopping norm
  double outer_rn, inner_target, eta, outer_rn_prev;
  double inner_rn;  // the inner solve's latest residual norm (for the host's poll schedule)
  // (the four fields every launch tests, together in 16 bytes: one load)
  alignas(16) int32_t state;
  int32_t phase;  // PH_*: where the graph is in its solve
  int32_t inner_state;
  // mixed precision: 1 when this graph's inner solves take FP64 K products (from the start, or
  // once a refinement step stalled: the FP32 K's error times K's condition number near 1)
  int32_t inner64;
  int32_t iters, maxit, outer;
  int32_t inner_iters, inner_maxit, zero_load;
};

// Every graph of a batch runs its own solve: each launch acts on the graphs in the phase (and, for
// inner solves, the product precision) it serves and leaves the rest alone, so graphs can enter a
// running batch, and leave it, between any two steps, each doing the same arithmetic as alone.
enum : int32_t {
  PH_NEW = 0,     // admitted: the start of the solve is next (r = b - K x, |b|; FP64: z, p, rz)
  PH_ITER = 1,    // FP64 PCG iterations
  PH_OUTER = 2,   // mixed: a refinement step's FP64 residual and check are next
  PH_START = 3,   // mixed: an inner solve (K d = r) starts
  PH_INNER = 4,   // mixed: the inner solve iterates
  PH_APPLY = 5,   // mixed: the inner solve ended; x += d and the step closes
  PH_WAIT64 = 6,  // mixed: its next inner solve takes FP64 products, once the host launches them
};

// who a launch serves (the graph must be running): SERVE_OUTER the FP64 PCG iterations,
// SERVE_NEW solve starts, SERVE_REFINE refinement residuals; otherwise inner solves, by bits: 1
// FP32 K products, 2 FP64 K products, 4 starting, 8 iterating
enum : int32_t {
  SERVE_OUTER = 0,
  SERVE_P32 = 1,
  SERVE_P64 = 2,
  SERVE_ST = 4,
  SERVE_IT = 8,
  SERVE_START = 1 | 2 | 4,
  SERVE_ITER = 1 | 2 | 8,
  SERVE_INNER32 = 1 | 4 | 8,
  SERVE_INNER64 = 2 | 4 | 8,
  SERVE_INNER = 1 | 2 | 4 | 8,
  SERVE_REFINE = 16,
  SERVE_NEW = 32,
};

// the smoother's degrees of a batch in one int: level 0's in the low byte, the coarser levels' in the
// next (0: the same as level 0's)
__host__ __device__ inline int level_degree(int packed, int level) {
  const int d0 = packed & 255, dc = (packed >> 8) & 255;
  return (level > 0 && dc > 0) ? dc : d0;
}

// an FP16 copy of a 6x6 block: the 36 entries over the block's largest magnitude (row-major), that
// magnitude as a float, 4 bytes of padding. (Scaling each 3x3 quarter on its own was measured: no
// fewer iterations, more bytes.)
constexpr int kHalfBlock = 80;

template <typename T>
struct LevelArrays {
  const T* A;      // 36 per block, row-major 6x6
  const T* Dinv;   // 36 per row
  const T* P;      // 36 per prolongation entry (rows of this level, columns of the next)
  T* x[2];         // ping-pong iterates (x[1] holds a V-cycle's result)
  T* b;            // right-hand side
  T* r;            // residual (restriction input)
  T* d;            // Chebyshev direction
};

struct LevelDev {
  int32_t n;          // block rows
  int32_t nc;         // rows of the next level (0 at the coarsest)
  int32_t terminal;   // 1: the graph's coarsest level
  int32_t dense_m;    // 6 n when a dense inverse exists at the coarsest, else 0
  int64_t nnzb, np;
  const int32_t* rowptr;
  const int32_t* col;
  const int32_t* Prow;   // n + 1
  const int32_t* Pcol;
  const int32_t* Rrow;   // nc + 1 (by coarse row)
  const int32_t* Rfine;  // the fine row of each restriction entry
  const int32_t* Rent;   // its prolongation entry (the restriction's block is that P block transposed)
  const double* Minv;    // dense inverse of the coarsest level (dense_m x dense_m, row-major)
  double lmax;           // lambda_max(D^-1 A) estimate (x 1.1)
  int32_t cheb_degree;
  int32_t sub;           // six-lane sub-groups a row in the tiled launches of this level (1 or 10)
  int32_t rsub;          // the same for its FP64 restriction (coarse rows of the next level; FP32 / FP16: a block a lane)
  int32_t pad_;
  const uint8_t* A16;    // A in FP16 blocks scaled by their largest entry (kHalfBlock bytes a block), or null
  const uint8_t* P16;    // P the same way, by prolongation entry (row-major blocks), or null
  const uint8_t* R16;    // P's blocks transposed, by restriction entry (the order of Rfine / Rent), or null
  const uint8_t* D16;    // the smoother's inverse diagonal blocks the same way, or null
  LevelArrays<double> f64;
  LevelArrays<float> f32;
};

struct GraphDev {
  int32_t nlev;
  int32_t nfree;
  int32_t jacobi;   // block-Jacobi preconditioner (one level, no hierarchy)
  int32_t resident; // small graph: its steps run in one workgroup (k_resident_steps), not in the tiled launches
  PcgState* st;
  // outer / FP64 PCG vectors, 6 nfree each
  double *x, *r, *z, *p, *q, *b;
  // inner FP32 PCG vectors (mixed precision)
  float *ix, *ir, *iz, *ip, *iq;
  LevelDev lv[kMaxLevels];
};

// ---- small device helpers --------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// the lane's (row within the tile, DOF) or row = -1 for the 4 idle lanes of each wavefront
__device__ __forceinline__ void lane_row(int& rl, int& q) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l >= 60) {
    rl = -1;
    q = 0;
    return;
  }
  rl = w * kRowsPerWave + l / 6;
  q = l % 6;
}

// Rows served by S six-lane sub-groups each (S in {1, 2, 5, 10}): a wavefront serves 10 / S rows;
// sub-group s of a row takes the row's blocks s, s + S, s + 2S, ... (long coarse rows spread their
// blocks over a wavefront instead of walking them with 6 lanes)
template <int S>
__device__ __forceinline__ void lane_row_s(int& rl, int& sub, int& q) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l >= 60) {
    rl = -1;
    sub = 0;
    q = 0;
    return;
  }
  rl = w * (kRowsPerWave / S) + l / (6 * S);
  sub = (l / 6) % S;
  q = l % 6;
}
template <int S>
constexpr int tile_rows() {
  return 4 * (kRowsPerWave / S);
}

// the sum over a row's S sub-groups of each DOF's partial, s ascending (every lane of the row gets it)
template <int S, typename T>
__device__ __forceinline__ T row_reduce(T v) {
  if (S == 1) return v;
  const int l = threadIdx.x & 63;
  const int base = (l / (6 * S)) * (6 * S) + (l % 6);
  T acc = __shfl(v, base, 64);
  for (int s = 1; s < S; ++s) acc += __shfl(v, base + 6 * s, 64);
  return acc;
}

// sum over a wavefront in a fixed butterfly order (every lane gets the sum)
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum over the workgroup in a fixed order: wavefront butterflies, then the four wavefront sums in
// order; valid in thread 0 (scratch: 4 doubles of LDS)
__device__ __forceinline__ double block_sum(double v, double* lds4) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = ((lds4[0] + lds4[1]) + lds4[2]) + lds4[3];
  __syncthreads();
  return s;
}

// the 6 values of this lane's row held by its 6 lanes (lane q holds v_q), read by every lane of
// the row
__device__ __forceinline__ void row_gather(double v, double out[6]) {
  const int base = (lane_id() / 6) * 6;
  for (int k = 0; k < 6; ++k) out[k] = __shfl(v, base + k, 64);
}
__device__ __forceinline__ void row_gather(float v, float out[6]) {
  const int base = (lane_id() / 6) * 6;
  for (int k = 0; k < 6; ++k) out[k] = __shfl(v, base + k, 64);
}

// the tile a workgroup serves: consecutive workgroups go round-robin over the 8 XCDs, so each
// XCD is given a contiguous range of tiles (its L2 holds the x values its rows gather)
__device__ __forceinline__ int xcd_tile(int ntiles) {
  const int b = blockIdx.x;
  if (ntiles < 64) return b;
  const int per = (ntiles + 7) / 8;
  const int t = (b % 8) * per + b / 8;
  return t;
}

}  // namespace svxg
