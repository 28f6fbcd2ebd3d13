// Host-side launchers of the svxg kernels (solve_kernels.hip, setup_kernels.hip, check_kernels.hip).
#pragma once

#include "svxg_dev.hpp"

namespace svxg {

// Chebyshev smoothing interval [lmax / kChebRatio, lmax]
constexpr double kChebRatio = 30.0;
constexpr int kStallDegree = 8;  // Chebyshev steps of a coarsest level left without a dense solve

enum class FinishKind { Start, Alpha, Rn, Beta, Outer };

int grid_for(int ntiles);

// ---- solve ---------------------------------------------------------------------------------------
// timing mode: the workgroup of each served graph's first tile adds that graph's algorithmic bytes
// (gb[g]) to *acc, so that launches left running after a graph stopped count time and no bytes
struct ProfAcc {
  const double* gb = nullptr;
  double* acc = nullptr;
};
// S: six-lane sub-groups a row (1: 10 rows a wavefront; 10: one row a wavefront, for long rows)
template <typename T>
// (t0, t1: events the kernel's own dispatch records at its start and end - kernel timing - or null)
void launch_cheb_step(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int level, int step, int in_buf,
                      int out_buf, int fresh, int inner, double* partial, int S, ProfAcc pa = {},
                      hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
template <typename T>
void launch_residual(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int level, int buf, int inner,
                     int S);
template <typename T>
void launch_restrict(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int level, int inner, int S);
template <typename T>
void launch_prolong(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int level, int buf, int inner);
// (part, cnt: scratch of count x kDenseSpans x 64 doubles and count zeroed counters for the launch
// a wavefront a (chunk, span); null: a 1024-lane workgroup a chunk)
constexpr int kDenseSpans = 16;
template <typename T>
void launch_dense_solve(hipStream_t s, const GraphDev* G, const Tile* chunks, int count, int level, int inner,
                        double* part = nullptr, int32_t* cnt = nullptr);
template <typename T>
void launch_jacobi(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner);
template <typename T>
void launch_pcg_spmv(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial,
                     ProfAcc pa = {}, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
template <typename T>
void launch_axpy_rr(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial);
template <typename T>
void launch_update_p(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner);
template <typename T>
void launch_start_p(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial);
template <typename T>
void launch_dot_rz(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial);
template <typename T>
void launch_dot_rr(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, int inner, double* partial);

void launch_begin_residual(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, double* partial);
// (mixed: the graphs go on to their first refinement step)
void launch_finish_begin(hipStream_t s, const GraphDev* G, const int32_t* gtile, int count, int mixed,
                         const double* partial);
// graphs waiting for FP64-product inner solves may start them
void launch_release_wait64(hipStream_t s, const GraphDev* G, int count);
void launch_zero_load(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles);
void launch_finish(hipStream_t s, FinishKind kind, const GraphDev* G, const int32_t* gtile, int count, int inner,
                   const double* partial);
void launch_outer_residual(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles, double* partial);
void launch_inner_init(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles);
// (kernel timing) acc += the bytes (gbytes per graph) of the graphs live for `inner`, read before the
// next launch changes their states
void launch_inner_apply(hipStream_t s, const GraphDev* G, const Tile* tiles, int ntiles);
void launch_inner_close(hipStream_t s, const GraphDev* G, int count);
// nsteps whole steps (a solve start first, for graphs just admitted) of each listed resident graph,
// one workgroup a graph: the same device functions, on the same tiles, in the same order as the
// tiled launches (mixed: mixed_step / fp64_step of the host driver), so the same bits
void launch_resident_steps(hipStream_t s, const GraphDev* G, const int32_t* gtile, const int32_t* graphs, int count,
                           int nsteps, int degree, int mixed, int p32, int p64, double* partial);
void launch_resident_recheck(hipStream_t s, const GraphDev* G, const int32_t* gtile, const int32_t* graphs, int count,
                             double* partial);
// FP64-product inner solves: r_i = (float) r (to32 = 1) or z = (double) z_i (to32 = 0)
// plain y = A x on one graph's level (FP64 or FP32 copy), tests and the roofline probe
template <typename T>
void launch_spmv_plain(hipStream_t s, const int32_t* rowptr, const int32_t* col, const T* A, const T* x, T* y, int n);

// ---- setup ---------------------------------------------------------------------------------------

// Assembly of one graph's K (stress.cpp:84-132): the bonds' four block products, their sums into
// the pattern's blocks in bond order, the identity rows.
struct AssemblyDev {
  int32_t nbonds, nfree, nslots;
  int32_t pad_;
  const svxg_bond* bonds;     // the graph's bonds
  const double* nodec;        // 3 per node
  const svxg_material* mats;  // the material table
  int32_t nmats;
  int32_t pad2_;
  const int32_t* bond_slot;   // 4 per bond: the (aa, bb, ab, ba) blocks, -1 absent
  const int32_t* cptr;        // nslots + 1
  const int32_t* cidx;        // 4 bond + product, in bond order
  const uint8_t* row_gone;    // per free row: its node is retired
  double* prod;               // 4 x 36 per bond (scratch)
  const int32_t* rowptr;
  double* K;                  // 36 per block
  double* id_scale;           // 1 (out)
  double* trace;              // per free row (scratch)
};

void launch_bond_products(hipStream_t s, const AssemblyDev* jobs, const Tile* tiles, int ntiles);
void launch_gather_blocks(hipStream_t s, const AssemblyDev* jobs, const Tile* tiles, int ntiles);
void launch_diag_trace(hipStream_t s, const AssemblyDev* jobs, const Tile* tiles, int ntiles);
void launch_id_scale(hipStream_t s, const AssemblyDev* jobs, int count);
void launch_identity_rows(hipStream_t s, const AssemblyDev* jobs, const Tile* tiles, int ntiles);
// in-place removal of bonds from K (stress.cpp:260-284): the listed bonds' products, subtracted
// (one workgroup a job: job j's bonds are ids[offs[j] .. offs[j + 1]))
// bonds leaving K in list order: each removed bond's four products (prod: 144 doubles a bond), then
// each touched block's products added in list order (groups: a job's block slot and its products'
// range in cont)
void launch_removal_products(hipStream_t s, const AssemblyDev* jobs, const int32_t* ids, const int32_t* tjob,
                             double* prod, int nids);
void launch_removal_apply(hipStream_t s, const AssemblyDev* jobs, const int32_t* gjob, const int32_t* gslot,
                          const int32_t* gstart, const int32_t* cont, const double* prod, int ngroups);
// rows of retired nodes become id_scale I (stress.cpp:299-308)
void launch_retire_rows(hipStream_t s, const AssemblyDev* job, const int32_t* rows, int count);

// Inverse diagonal blocks: mode 0 the multigrid's (amg.cpp:228-254: regularised by 1e-9 tr on a
// singular block; *fail set when that too is singular), mode 1 block-Jacobi's
// (stress.cpp:152-157: (1 / id_scale) I on a singular block)
struct DiagDev {
  int32_t n;
  int32_t mode;
  const int32_t* rowptr;
  const double* A;
  double* Dinv;
  const double* id_scale;
  int32_t* fail;
};
void launch_inv_diag(hipStream_t s, const DiagDev* jobs, const Tile* tiles, int ntiles);

// One coarsening step of one graph's level (amg.cpp:320-533)
struct CoarsenDev {
  // the fine level
  int32_t n;
  int32_t pad0_;
  int64_t nnzb;
  const int32_t* rowptr;
  const int32_t* col;
  const double* A;
  const double* Dinv;
  const double* pos;   // 3 per row
  double theta;
  double omega;        // prolongation smoothing weight (0: unsmoothed)
  // scratch over blocks / rows
  double* s;           // block strength (nnzb)
  uint8_t* strong;     // nnzb
  int32_t* state;      // n
  uint64_t* t1;        // n
  uint64_t* t2;        // n
  int32_t* rootof;     // n
  int32_t* rootof2;    // n
  int32_t* scan_a;     // n + 1 (anchors)
  int32_t* scan_b;     // n + 1 (lonely)
  int32_t* cursor;     // n + 1 (fill cursors)
  int32_t* counts;     // 8 ints: [0] undecided (cumulative), [1] anchors, [2] lonely, [3] rows over the
                       // distinct-column limit, [4] np, [5] nap, [6] nac, [7] candidates
  // results
  int32_t* agg;        // n
  int32_t na;
  int32_t pad1_;
  int32_t* acnt;       // na + 1
  int32_t* amem;       // n (members by aggregate)
  double* posc;        // 3 na
  double* off;         // 3 n
  int32_t* cand;       // candidate scratch
  int64_t ncand;
  int32_t* cptr;       // n + 1 (candidate offsets)
  int32_t* Prow;       // n + 1
  int32_t* Pcol;
  double* P;
  int64_t np;
  int32_t* Rrow;       // na + 1
  int32_t* Rfine;      // np
  int32_t* Rent;       // np (the P entry of each restriction entry)
  int32_t* AProw;      // n + 1
  int32_t* APcol;
  double* AP;
  int64_t nap;
  int32_t* Crow;       // na + 1
  int32_t* Ccol;
  double* C;
  int64_t nac;
  int32_t* rid;        // the row of each P / A P / C entry (scratch)
  double coarse_scale; // applied to C (1: none)
};

// MIS-2 aggregation
void launch_strength(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles);
void launch_strong(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles);
// one distance-2 MIS round; round r's undecided rows land in the job's counts[r & 1 ? 6 : 0], and a
// round after one that left a graph nothing undecided skips that graph
void launch_mis_round(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, int round);
void launch_aggregate(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, int count);
// per-graph exclusive scans: scan_a / scan_b (n + 1 each) -> totals in counts[1], counts[2]
void launch_scan_pair(hipStream_t s, const CoarsenDev* jobs, int count);
void launch_assign(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles);
void launch_coarse_positions(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles,
                             const Tile* ctiles, int nctiles, int count);
// prolongator: candidates (sorted, distinct) per row, their count in Prow; columns and each entry's
// row; the blocks one entry a thread (etiles: tiles over entries)
void launch_p_count(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles);
void launch_scan_rows(hipStream_t s, const CoarsenDev* jobs, int count, int which);
void launch_p_cols(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles);
void launch_p_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* etiles, int ntiles, bool lanes);
void launch_r_build(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, const Tile* ctiles,
                    int nctiles, int count);
// Galerkin product: candidate counts and candidates per row, distinct sorted columns (a wavefront
// a row, utiles of four rows; a workgroup for the rows too long for that), the pattern, the blocks.
// which: 0 A P, 1 P^T (A P)
void launch_ap_cand_count(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles);
void launch_ap_cand_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles);
void launch_c_cand_count(hipStream_t s, const CoarsenDev* jobs, const Tile* ctiles, int nctiles);
void launch_c_cand_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* ctiles, int nctiles);
void launch_unique_rows(hipStream_t s, const CoarsenDev* jobs, const Tile* utiles, int ntiles, int which);
void launch_pattern_cols(hipStream_t s, const CoarsenDev* jobs, const Tile* tiles, int ntiles, int which);
void launch_ap_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* etiles, int ntiles, bool lanes);
void launch_c_fill(hipStream_t s, const CoarsenDev* jobs, const Tile* etiles, int ntiles, bool lanes);

// power iteration for lambda_max(D^-1 A) (amg.cpp:256-275): 15 steps, per-graph norms
struct PowerDev {
  int32_t n;
  int32_t self;  // few tiles: its steps find |v| from the last step's partials themselves (no finish launch)
  const int32_t* rowptr;
  const int32_t* col;
  const double* A;
  const double* Dinv;
  double* v[2];
  double* lam;   // 2 doubles: [0] the estimate, [1] the last norm
};
void launch_power_init(hipStream_t s, const PowerDev* jobs, const Tile* tiles, int ntiles, int count);
// a graph normalises itself in the steps when it has at most this many tiles
constexpr int kPowerSelfTiles = 256;
// one step of every graph: tiles1 (S = 1) and tiles10 (S = 10) in one launch; partials by parity of step
void launch_power_steps(hipStream_t s, const PowerDev* jobs, const Tile* tiles1, int n1, const Tile* tiles10, int n10,
                        const int32_t* gtile, int step, double* partial);
// the norm of the partials of parity `which`: the graphs of many tiles after each step, or (close) the
// self-normalising ones after the last
void launch_power_finish(hipStream_t s, const PowerDev* jobs, const int32_t* gtile, int count, const double* partial,
                         int which, bool close);

// the coarsest level's dense inverse (amg.cpp:535-578, then inverted)
// the P, A P and P^T A P fills run a wavefront an entry (its lanes find the entry's terms together,
// 36 of them then add them up, one a block element; four entries a workgroup) when a setup has fewer
// than kFillLanesMax entries to fill (latency-bound), else one thread an entry (throughput-bound);
// both give the same bits
constexpr int kFillEntries = kWG / 64;
constexpr int64_t kFillLanesMax = 200000;

struct DenseDev {
  int32_t n;
  int32_t m;
  int32_t mp;     // m rounded up to a multiple of 16 (the padded scratch)
  int32_t pad_;
  const int32_t* rowptr;
  const int32_t* col;
  const double* A;
  double* L;      // mp x mp scratch (the factor)
  double* W;      // mp x mp scratch (its inverse)
  double* Minv;   // m x m
};
void launch_dense_factor(hipStream_t s, const DenseDev* jobs, int count);
// the same spread over many workgroups (graphs whose padded size is above kDenseOneBlocks tile rows;
// max_nb: the largest graph's tile rows; dmax: a double a graph of scratch). One workgroup runs a
// graph's tile products one memory round trip after another: 0.9 ms at 288 rows against about a
// third of that spread; the spread form takes every graph (kDenseOneBlocks = 0: the one-workgroup
// kernel stays for a caller that wants a graph's inverse in one launch)
constexpr int kDenseOneBlocks = 0;
void launch_dense_factor_wide(hipStream_t s, const DenseDev* jobs, int count, int max_nb, double* dmax);

// FP32 copies
void launch_to_f32(hipStream_t s, const double* in, float* out, int64_t count);

// ---- bulk copies of a batch setup: many small segments, one launch -----------------------------------
struct BulkSeg {
  void* dst;
  const void* src;  // device source (copies), or a double array (conversions), or null (fills)
  int64_t count;    // bytes (fills, copies) or elements (conversions)
};
// each segment's count bytes at dst set to zero (dst 16-byte aligned)
void launch_zero_many(hipStream_t s, const BulkSeg* segs, int n);
// each segment's count doubles at src written as floats at dst
void launch_f32_many(hipStream_t s, const BulkSeg* segs, int n);
// FP16 block copies (kHalfBlock bytes each) of nnzb FP32 6x6 blocks
void launch_half_blocks(hipStream_t s, const float* A, uint8_t* H, int64_t nnzb);
// the FP16 copy of the transposes of the blocks P[ent[e]], e < n, in the order of e
void launch_half_blocks_t(hipStream_t s, const float* P, const int32_t* ent, uint8_t* H, int64_t n);
// many FP16 copies in one launch: piece k converts `count` (at most kWG) blocks into dst, from src's
// consecutive blocks, or transposed from src's blocks ent[0 .. count) when ent is set
struct HalfSeg {
  const float* src;
  const int32_t* ent;
  uint8_t* dst;
  int64_t count;
};
void launch_half_many(hipStream_t s, const HalfSeg* segs, int n);
// each segment's count bytes copied from src to dst (both 16-byte aligned)
void launch_copy_many(hipStream_t s, const BulkSeg* segs, int n);

// ---- bond checks (stress.cpp:14-57, 477-502) ---------------------------------------------------
struct CheckDev {
  int32_t nbonds;
  int32_t nmats;
  const svxg_bond* bonds;
  const svxg_node* nodes;
  const svxg_material* mats;
  const double* u;   // 6 per node (all nodes)
  double fragility;
  svxg_bond_load* loads;
  double* phi;
  uint8_t* mode;
};
void launch_bond_checks(hipStream_t s, const CheckDev* jobs, const Tile* tiles, int ntiles);
// u over all nodes (6 per node) from the free-DOF solution
void launch_scatter_u(hipStream_t s, const double* x, const int32_t* dof, double* u, int nnodes);
// x (free DOFs) from u (6 per node, all nodes)
void launch_gather_u(hipStream_t s, const double* u, const int32_t* dof, double* x, int nnodes);
// v[i] = 0 where v[i] is not finite
void launch_zero_nonfinite(hipStream_t s, double* v, int64_t n);
// u (all nodes) from x (free DOFs) for many problems, one workgroup a problem at a time
struct ScatterJob {
  const double* x;
  const int32_t* dof;
  double* u;
  int32_t nnodes;
  int32_t pad_;
};
void launch_scatter_many(hipStream_t s, const ScatterJob* jobs, int n);

// ---- the judge (world.cpp:1162-1277, world_pieces.cpp:647-671) ----------------------------------
// a bond over strength: its phi's bits (phi >= 1, so the bits order as the values), index, mode
struct OverEntry {
  uint64_t key;
  int32_t bond;
  int32_t mode;
};
// a problem's judgement: filled by the judge pass (max_bits, over) and the selection (picked, hinges)
struct JudgeOut {
  unsigned long long max_bits;  // the largest judged phi's bits (phi > 0; 0 if none)
  int32_t over;
  int32_t picked;
  int32_t hinges;
  int32_t pad_;
};
constexpr int kMaxPicks = 4096;  // a round's picks at most (the selection sorts them in LDS)
struct JudgeDev {
  int32_t nbonds, nmats;
  int32_t shock, transient;
  int32_t rule;        // 0 structure quota, 1 piece quota
  int32_t max_breaks;  // 1..kMaxPicks
  int32_t round;       // written into the picks
  int32_t skip;        // judged, nothing picked (rounds >= max_rounds)
  int32_t hinges;      // plastic-hinge cuts of the picks
  int32_t pad_;
  double fragility, dif, band, quota_fraction, hinge_shape;
  uint64_t ductile[2];
  const svxg_bond* bonds;
  const svxg_node* nodes;
  const svxg_material* mats;
  const double* u;          // 6 per node (all nodes)
  const uint8_t* dead;      // per bond: 1 = not judged
  svxg_bond_load* base;     // baseline loads (null: no bond has one)
  uint8_t* has_base;
  double* phi;              // optional: each judged bond's phi, 0 for the others
  OverEntry* over;          // room for nbonds
  JudgeOut* out;
  svxg_break* picks;        // room for max_breaks
};
// phi of every judged bond, the baseline update and the list of bonds over strength
void launch_judge(hipStream_t s, const JudgeDev* jobs, const Tile* tiles, int ntiles);
// each problem's picks (one workgroup a problem): sorted, banded, capped, with hinge cuts
void launch_judge_select(hipStream_t s, const JudgeDev* jobs, int count);
// rows (6 values each) of x and b set to 0, a workgroup a job
struct ZeroRowsJob {
  double* x;
  double* b;
  const int32_t* rows;
  int32_t n;
  int32_t pad_;
};
void launch_zero_rows(hipStream_t s, const ZeroRowsJob* jobs, int count);

// ---- bond sections from voxel faces (world_internal.hpp:129-249, 374-400) --------------------------
struct SectionDev {
  const svxg_node* nodes;
  const svxg_section_head* heads;
  const svxg_section_face* faces;
  const uint8_t* vox;
  const uint8_t* damage;  // or null
  const svxg_material* mats;
  svxg_bond* out;
  int32_t lo[3], dim[3];
  int32_t nmats, m;
  double h;
};
void launch_bond_sections(hipStream_t s, const SectionDev* dev, int m);

// the general case: an element (face or junction sample) of a bond as a square in the bond's frame
struct SecElem {
  double x[3], nk[3];  // centre, normal from a to b
  double ar, sd;       // area (times the sample's weight), side
  double st[5];        // strength shares: ft, fb, fc, cohesion (w x condition x weaker), friction (w x weaker)
  double w;            // strength weight: 1 a face, 1 / S^2 a sample
};
struct SectionGenDev {
  const svxg_node* nodes;
  const svxg_section_head_general* heads;
  const svxg_section_face* faces;
  const svxg_junction_sample* samples;
  const svxg_lattice* lat;  // (device voxel pointers)
  const svxg_material* mats;
  const int64_t* eoff;      // m + 1: the bonds' element ranges (0 for a one-lattice bond)
  SecElem* elem;
  svxg_bond* out;
  int64_t E;
  int32_t m, nmats, S, pad_;
};
void launch_bond_sections_general(hipStream_t s, const SectionGenDev* dev, int m, int64_t elements);

// ---- one lattice's graph (world.cpp:425-865 for one lattice, a node a fragment) -----------------------
struct LatticeDev {
  const uint8_t* vox;
  const int32_t* frag;
  const int32_t* fnode;  // a fragment's node (clusters), or null: a node a fragment
  const svxg_material* mats;
  int32_t lo[3], dim[3];
  int64_t nvox;
  int32_t nfrag, nmats;
  double h;
};
// clusters of fragments (append_nodes, world.cpp:741-800): each fragment's first voxel and cell,
// rounds of min-label propagation over the links inside cells, the labels' ranks, the nodes
void launch_frag_first(hipStream_t s, const LatticeDev* L, uint64_t* nfirst, int64_t nvox);
void launch_frag_cells(hipStream_t s, const LatticeDev* L, const uint64_t* nfirst, int cell, int cubes,
                       uint64_t* cellkey, int32_t* label, int nfrag);
void launch_cluster_round(hipStream_t s, const LatticeDev* L, const uint64_t* cellkey, int32_t* label,
                          int32_t* changed, int64_t nslots, int nfrag);
void launch_cluster_ids(hipStream_t s, const int32_t* label, int32_t* root, int nfrag);
void launch_cluster_map(hipStream_t s, const int32_t* label, const int32_t* rank, int32_t* fnode, uint32_t* key,
                        int32_t* idx, int32_t* count, int nfrag);
void launch_cluster_nodes(hipStream_t s, const int32_t* member, const int32_t* start, const svxg_node* fnodes,
                          const uint8_t* fmat, svxg_node* nodes, uint8_t* nmat, int nnodes);
void launch_face_flags(hipStream_t s, const LatticeDev* L, int32_t* flag, int64_t nslots);
void launch_face_write(hipStream_t s, const LatticeDev* L, const int32_t* flag, const int32_t* at, uint64_t* key,
                       int32_t* idx, svxg_section_face* face, uint8_t* mb, int64_t nslots);
void launch_gather_faces(hipStream_t s, const int32_t* idx, const svxg_section_face* in, const uint8_t* mb_in,
                         svxg_section_face* out, uint8_t* mb_out, const uint64_t* key, int32_t* first, int n);
void launch_heads(hipStream_t s, const uint64_t* key, const int32_t* first, const int32_t* hid, const uint8_t* mb,
                  const uint8_t* nmat, int nfaces, svxg_section_head* heads, int nheads);
void launch_voxel_keys(hipStream_t s, const LatticeDev* L, uint32_t* key, int32_t* idx, int32_t* count, int64_t nvox);
void launch_node_sums(hipStream_t s, const LatticeDev* L, const int32_t* vidx, const int32_t* start, int64_t skip,
                      svxg_node* nodes, uint8_t* nmat, int nfrag);
// ---- a structure's extraction: the walk from a seed (world.cpp:425-620, 739-865) -------------------
struct WalkState {
  int32_t count;         // members so far
  int32_t level_begin;   // the level being walked: members [level_begin, level_end)
  int32_t level_end;
  int32_t eligible;      // fragments the level met first that may be taken
  int32_t truncated, any_support, touched_unloaded, bad;
};
struct WalkDev {
  const uint8_t* vox;
  const uint8_t* broken;  // or null
  const int32_t* frag;
  const svxg_fragment* frags;
  int32_t* bbox;          // 6 a fragment: its voxels' least and greatest box-local x, y, z
  int32_t* slot;          // a fragment's member index; -1 not met; -2 held fixed
  uint64_t* meet;         // the least (member, voxel, face) key that met a fragment
  int32_t* ecnt;          // a member's first-met fragments that may be taken
  int32_t* ebase;         // their first slot among the level's
  int32_t* members;
  WalkState* st;
  int32_t lo[3], dim[3];
  int64_t nvox;
  int32_t nfrag, seed, max_nodes, outside;
  double max_radius;
  int64_t fresh_from;
};
// the merge's arrays (k_walk_heads, k_walk_faces)
struct WalkMerge {
  const uint64_t* mkey;      // node-level keys, sorted (over the fine bonds in the order made)
  const int32_t* mstart;     // a merged bond's first entry in that list (nmerged + 1)
  const int32_t* mpos;       // its first fine bond's place in the order made
  const int32_t* mseg;       // merged bonds started before an entry (exclusive scan of mfirst)
  const int32_t* mfirst;     // an entry starts a merged bond
  const int32_t* kinv;       // a fine bond's (by order) entry in the list
  const int32_t* final_of;   // a merged bond's place among the bonds
  const int32_t* face0;      // a bond's first face (bonds + 1)
  const int32_t* S;          // faces before an entry of the list (exclusive scan)
  const int32_t* byfo;       // the fine bond made o-th (key order index)
  const int32_t* fo;         // a fine bond's place in the order made
  const uint64_t* fkey;      // a fine bond's key
  const int32_t* fstart;     // its first record in the sorted records (nfine + 1)
  const int32_t* fpos;       // its first record's place in the walk
  const int32_t* seg;        // fine bonds started before a sorted record (exclusive scan of ffirst)
  const int32_t* ffirst;     // a sorted record starts a fine bond
  const int32_t* sidx;       // a sorted record's place in the walk
  const svxg_section_face* rface;
  const uint8_t* rmb;
  const int32_t* fnode;      // a member's node
  const uint8_t* nmat;
  const double* nstr;
  const uint64_t* ident;
  uint8_t* swapped;          // a merged bond's ends swapped for their identities
};
// the inputs' rules checked (*bad: 1 a fragment index out of range, 2 a fragment on air or an
// anchored voxel, 4 a fragment's voxel outside its chunk); the walk's arrays set up
void launch_walk_check(hipStream_t s, const WalkDev* W, int64_t nvox, int32_t* bad);
void launch_walk_init(hipStream_t s, const WalkDev* W, int nfrag, int64_t nvox);
// one level of the walk (meet, own, scan, assign, next): no-ops once a level is empty
void launch_walk_level(hipStream_t s, const WalkDev* W);
void launch_walk_count(hipStream_t s, const WalkDev* W, int nm, int32_t* count);
void launch_walk_emit(hipStream_t s, const WalkDev* W, int nm, const int32_t* off, uint64_t* rkey,
                      svxg_section_face* rface, uint8_t* rmb);
void launch_seg_first(hipStream_t s, const uint64_t* key, int n, int32_t* first);
void launch_fine(hipStream_t s, const uint64_t* skey, const int32_t* sidx, const int32_t* first, const int32_t* seg,
                 int n, int32_t* fstart, int32_t* fpos, uint64_t* fkey, int32_t* flag);
void launch_fine_order(hipStream_t s, const int32_t* fpos, const int32_t* rank, int nfine, int32_t* fo, int32_t* byfo);
void launch_walk_cells(hipStream_t s, const WalkDev* W, int cell, int cubes, uint64_t* cellkey, int32_t* label, int nm);
void launch_walk_link_round(hipStream_t s, const uint64_t* fkey, int nfine, const uint64_t* cellkey, int32_t* label,
                            int32_t* changed, int nm);
void launch_walk_iota(hipStream_t s, int32_t* a, int32_t* b, int n);
void launch_walk_nodes(hipStream_t s, const WalkDev* W, const int32_t* mlist, const int32_t* start, int nn,
                       svxg_node* nodes, uint8_t* nmat, double* nstr, uint64_t* ident);
void launch_merge_keys(hipStream_t s, const uint64_t* fkey, const int32_t* byfo, const int32_t* fnode, int nfine,
                       uint64_t* mkey, int32_t* mval);
void launch_merge_counts(hipStream_t s, const uint64_t* mkey, const int32_t* mval, const int32_t* byfo,
                         const int32_t* fstart, int n, int32_t* cnt, int32_t* kinv);
void launch_merged(hipStream_t s, const uint64_t* mkey, const int32_t* mval, const int32_t* first, const int32_t* seg,
                   int n, int32_t* mstart, int32_t* mpos, int32_t* flag);
void launch_merged_order(hipStream_t s, const int32_t* mpos, const int32_t* rank, const int32_t* mstart,
                         const int32_t* S, int nmerged, int32_t* final_of, int32_t* fcount);
void launch_walk_heads(hipStream_t s, const WalkMerge& M, int nmerged, svxg_section_head* heads, uint64_t* bid);
void launch_walk_faces(hipStream_t s, const WalkMerge& M, int nrec, svxg_section_face* out);

hipError_t scan_exclusive_i32(void* tmp, size_t& bytes, const int32_t* in, int32_t* out, int64_t n, hipStream_t s);
hipError_t sort_pairs_u64(void* tmp, size_t& bytes, const uint64_t* kin, uint64_t* kout, const int32_t* vin,
                          int32_t* vout, int64_t n, int bits, hipStream_t s);
hipError_t sort_pairs_u32(void* tmp, size_t& bytes, const uint32_t* kin, uint32_t* kout, const int32_t* vin,
                          int32_t* vout, int64_t n, int bits, hipStream_t s);

}  // namespace svxg
