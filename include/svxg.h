/* svxg — the fragment-graph elastic-equilibrium stress solve on one MI355X (gfx950).
 *
 * C ABI of the device library libsvxg.so. Plain structs, pointers and sizes: no C++ types, no
 * torch types. Every call returns an int status (SVXG_OK = 0); nothing aborts.
 *
 * The data model mirrors the reference's (core/include/svx/stress/stress.hpp:21-51 SNode/SBond,
 * core/include/svx/material/material.hpp:47-86 Material) field for field, as plain structs. The
 * entry points replace the member functions of svx::StressProblem
 * (core/include/svx/stress/stress.hpp:74-160) and the free functions bond_strength /
 * bond_utilization (stress.hpp:57,66); each declaration cites the function it replaces.
 *
 * Threading: calls on distinct problems may come from distinct host threads (each problem
 * owns its stream); calls on one problem must not overlap. svxg_solve_batch runs many
 * problems in one set of launches (the preferred entry for many problems a tick).
 */
#ifndef SVXG_H
#define SVXG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
enum {
  SVXG_OK = 0,
  SVXG_BAD_INPUT = 1,    /* invalid argument (index out of range, size mismatch, null) */
  SVXG_OOM_BUDGET = 2,   /* the context's HBM budget would be exceeded */
  SVXG_DEVICE_ERROR = 3, /* a HIP call failed (the message: svxg_last_error) */
  SVXG_NOT_ASSEMBLED = 4,/* the problem has no operator (assemble failed or was never run) */
  SVXG_SINGULAR = 5      /* assemble: a diagonal block stayed singular after regularisation */
};

/* ---- precision of a solve ---------------------------------------------------------------- */
enum {
  /* the plain path: FP64 PCG (FP64 K products, dots and vector updates) around an FP64 V-cycle
     (FP64 smoother, residuals, transfers and dense coarsest solve) */
  SVXG_FP64 = 0,
  /* mixed precision, by default (svxg_stress_options.mixed_fp64_rows): FP64 PCG (FP64 K products,
     dots and updates) around a V-cycle computed in FP32 (its matrices, prolongators and inverse
     diagonal blocks stored in FP32, or as FP16 blocks scaled by their largest entry for graphs of
     vcycle_fp16_rows free nodes and more; the dense coarsest solve in FP64), checked by an FP64
     residual. Graphs below mixed_fp64_rows free nodes instead run FP64 iterative refinement (FP64
     residual recomputation and norms) around an inner PCG whose K products also run in FP32. */
  SVXG_MIXED = 1
};

/* ---- data model -------------------------------------------------------------------------- */

/* A fragment: rigid, 6 DOFs (u, theta about c). stress.hpp:21-26 (SNode). */
typedef struct svxg_node {
  double c[3];     /* centre (the solve's frame) */
  double mass;
  uint8_t fixed;   /* pinned: no DOFs */
  uint8_t gone;    /* retired: an identity row, no bonds */
  uint8_t pad_[6];
} svxg_node;       /* 40 bytes */

/* An interface between two fragments (b < 0: a fixed support). stress.hpp:28-51 (SBond). */
typedef struct svxg_bond {
  int32_t a, b;
  uint8_t broken;     /* severed */
  uint8_t in_k;       /* its stiffness is in the assembled K (kept by the library) */
  uint8_t ma, mb;     /* material ids of the two sides */
  uint8_t sectioned;  /* section strengths below are used (else the weaker material's) */
  uint8_t pad0_[3];
  double n[3], t1[3], t2[3]; /* local frame: n from a towards b */
  double area;               /* m^2 */
  double p[3];               /* section centroid */
  double s1, s2;             /* second moments of area about p along t1, t2 */
  double c1, c2;             /* extreme fibre distances along t1, t2 */
  double rmax;               /* largest distance of the section from p (torsion) */
  double la, lb;             /* lengths along n: node a to section, section to node b */
  double strength;           /* design strength multiplier */
  float ft, fb, fc, coh, mu; /* section strengths (Pa) when sectioned */
  int32_t faces, tag;
  int32_t pad1_;
} svxg_bond;                 /* 216 bytes */

/* Material properties a solve and the checks read. material.hpp:47-57 (Material). */
typedef struct svxg_material {
  double E, G, rho;
  double ft, fb, fc, cohesion, friction;
} svxg_material;             /* 64 bytes */

/* AmgOptions (core/include/svx/solve/amg.hpp:46-69). The device V-cycle smooths with Chebyshev
   polynomials of degree cheb_degree in D^-1 A (amg.hpp:53-55 names that smoother; it runs
   row-parallel and is symmetric), not with the reference's hybrid Gauss-Seidel. */
typedef struct svxg_amg_options {
  double strength;           /* 0.1: aggregate across ||A_ij|| >= s sqrt(||A_ii|| ||A_jj||) */
  double coarse_scale;       /* 0.6: scaling of unsmoothed Galerkin coarse operators */
  int32_t coarse_max;        /* 48: dense solve at or below this many nodes */
  int32_t smoothed;          /* 1: smoothed aggregation P = (I - w D^-1 A) P_rigid */
  int32_t max_levels;        /* 16 */
  int32_t sweeps;            /* 1: smoother applications before and after the coarse correction */
  int32_t cheb_degree;       /* device smoother degree (default 2) */
  int32_t coarsen_dense;     /* 1 */
  int32_t reaggregate;       /* 1 */
  int32_t smooth_levels;     /* device: levels whose prolongator is smoothed, from the finest (0: all,
                                as the reference); coarser ones are plain aggregation (scaled by
                                coarse_scale), which keeps coarse operators sparse */
  int32_t coarse_cheb_degree;/* device: the smoother's degree below level 0 (default 1; 0: cheb_degree).
                                A batch runs its first problem's degrees */
  int32_t lean_rows;         /* device: graphs with at least this many free nodes (default 16384; 0: none)
                                take a lean hierarchy: unless smooth_levels says otherwise the two
                                finest prolongators are smoothed, coarser ones not, aggregation uses
                                lean_strength, and coarsening stops at the first level the dense
                                solve takes (at most 4 coarse_max nodes). Their coarse operators stay
                                sparse and the cycle short */
  double lean_strength;      /* device: the strength threshold of graphs that take the lean hierarchy
                                (default 0.05; 0: strength). On the 64x64x48 block the coarse levels
                                aggregate about twelve to one instead of two to four: four levels
                                [167129, 18396, 1234, 145] instead of six, 32 PCG iterations instead
                                of 38 (with the two smoothed prolongators) */
} svxg_amg_options;

/* StressOptions (stress.hpp:68-72). */
typedef struct svxg_stress_options {
  double rtol;               /* 2e-3: relative residual of a converged solve */
  int32_t amg_min_nodes;     /* 160: smaller graphs are block-Jacobi preconditioned */
  int32_t precision;         /* SVXG_FP64 or SVXG_MIXED */
  svxg_amg_options amg;
  int32_t mixed_fp64_rows;   /* SVXG_MIXED, device: graphs with fewer free nodes (default 2^30: all)
                                run one PCG with FP64 K products and the FP32 V-cycle (the
                                reference's own precision split, no refinement restarts); larger
                                ones take FP32 K products inside FP64 iterative refinement (0: all) */
  int32_t vcycle_fp16_rows;  /* SVXG_MIXED, device: graphs with at least this many free nodes
                                (default 16384; 0: all, 2^30: none) read their V-cycle's matrices
                                (smoothing and residual passes) as FP16 blocks, each 6x6 block over
                                its largest magnitude (a symmetric matrix stays symmetric),
                                computing in FP32; smaller graphs, latency-bound, read FP32 copies */
} svxg_stress_options;

/* PcgResult (amg.hpp:120-125). */
typedef struct svxg_pcg_result {
  int32_t iters;
  int32_t converged;
  int32_t breakdown;
  int32_t refinements;       /* SVXG_MIXED: outer refinement steps (0 for SVXG_FP64) */
  double rel_res;
} svxg_pcg_result;

/* BondLoad (stress.hpp:61-63): N (tension > 0), shear V1, V2, torsion T, bending M1, M2. */
typedef struct svxg_bond_load {
  double N, V1, V2, T, M1, M2;
} svxg_bond_load;

typedef struct svxg_ctx svxg_ctx;
typedef struct svxg_problem svxg_problem;

/* ---- context ------------------------------------------------------------------------------ */

/* A device context: one GPU, an HBM budget (bytes; 0 = 200 GB) that every problem's device
   memory is admitted against. No reference counterpart (the reference has no device). */
int svxg_ctx_create(int device, int64_t hbm_budget, svxg_ctx** out);
int svxg_ctx_destroy(svxg_ctx* ctx);
/* Bytes of device memory held by the context's problems. */
int64_t svxg_ctx_bytes_in_use(const svxg_ctx* ctx);
/* The message of the last failing call on this thread. */
const char* svxg_last_error(void);
/* Default options: StressOptions{} / AmgOptions{} (stress.hpp:68-72, amg.hpp:46-69). */
void svxg_default_options(svxg_stress_options* out);

/* ---- problems ----------------------------------------------------------------------------- */

/* A problem holds the graph (copied: nodes, bonds, the material table) and, once assembled,
   its device operator and preconditioner. StressProblem's public members nodes / bonds / mats
   (stress.hpp:76-79) become these arrays; nmats <= 127 (material.hpp:45). */
int svxg_problem_create(svxg_ctx* ctx, const svxg_node* nodes, int32_t n, const svxg_bond* bonds, int32_t m,
                        const svxg_material* mats, int32_t nmats, svxg_problem** out);
int svxg_problem_destroy(svxg_problem* p);

/* StressProblem::assemble (stress.hpp:83; stress.cpp:134-176): K from the intact bonds on the
   device, then block-Jacobi (nfree < amg_min_nodes) or the multigrid hierarchy, built on the
   device. Returns SVXG_SINGULAR where the reference's assemble returns false. A problem set up
   for SVXG_MIXED keeps its hierarchy in FP32 only (its FP64 steps read K and the coarsest
   level's FP64 inverse); one set up for SVXG_FP64 keeps both copies. */
int svxg_problem_assemble(svxg_problem* p, const svxg_stress_options* opt);
/* The options of the next assemble (StressProblem::opt_, stress.hpp:153), without assembling:
   svxg_assemble_batch / svxg_solve_batch assemble unassembled problems with them. Switching a
   problem set up for SVXG_MIXED to SVXG_FP64 makes it unassembled (it needs an FP64 hierarchy). */
int svxg_problem_set_options(svxg_problem* p, const svxg_stress_options* opt);
/* StressProblem::release (stress.hpp:99; stress.cpp:326-339). */
int svxg_problem_release(svxg_problem* p);
/* StressProblem::invalidate (stress.hpp:85-88). */
int svxg_problem_invalidate(svxg_problem* p);

/* StressProblem::solve (stress.hpp:122; stress.cpp:348-360). f, u: 6 per node over all nodes
   (host memory); u is the warm start when warm != 0, the solution on return. maxit < 0: no cap.
   The precision is the options' (assemble). */
int svxg_problem_solve(svxg_problem* p, const double* f, double* u, double rtol, int32_t maxit, int32_t warm,
                       svxg_pcg_result* res);
/* StressProblem::begin / iterate / current (stress.hpp:126-130; stress.cpp:362-475): the same
   solve spread over calls. u0 may be null (cold start). */
int svxg_problem_begin(svxg_problem* p, const double* f, const double* u0);
int svxg_problem_iterate(svxg_problem* p, int32_t maxit, double rtol, svxg_pcg_result* res);
int svxg_problem_current(svxg_problem* p, double* u);

/* StressProblem::remove_bond (stress.hpp:108; stress.cpp:260-284): the bonds break and their
   stiffness leaves K in place (never twice). */
int svxg_problem_remove_bonds(svxg_problem* p, const int32_t* ids, int32_t count);
/* StressProblem::retire_nodes (stress.hpp:109; stress.cpp:286-310). */
int svxg_problem_retire_nodes(svxg_problem* p, const int32_t* ids, int32_t count);
/* Restores bond b's `broken` flag to 0 without putting it back into K (the caller-kept topology of
   a retired chip, as test_stress.cpp:248-249 does through the public SBond member). */
int svxg_problem_set_bond_broken(svxg_problem* p, int32_t b, int32_t broken);

/* StressProblem::bond_load for every bond (stress.hpp:132; stress.cpp:477-502), from u (host,
   6 per node) or, when u is null, from the device-resident solution of the last solve. */
int svxg_problem_bond_loads(svxg_problem* p, const double* u, svxg_bond_load* out);
/* bond_utilization of every bond (stress.hpp:57,66; stress.cpp:14-57) after its bond_load:
   phi (1 = at strength) and the FailMode (0 None, 1 Tension, 2 Crush, 3 Shear). u as above. */
int svxg_problem_bond_checks(svxg_problem* p, const double* u, double fragility, double* phi, uint8_t* mode);
/* svxg_problem_bond_checks over device arrays, allocating no device memory. No counterpart in the
   reference: for callers that cannot allocate (code under HIP graph capture, tight serving loops).
   u: device, 6 doubles a node over all nodes, read in place (null: the resident solution of the last
   solve, which this call leaves as it is); phi: device, one double a bond; mode: device, one byte a
   bond, or null. The kernels' scratch (the job descriptor and the bond tile table) is a fixed array
   the library holds on each GPU until the process ends, for at most 16,777,216 bonds (SVXG_BAD_INPUT
   beyond that); the problem must be assembled (its graph's device copy is used). Calls on one GPU
   share that array: calls that overlap (through two contexts on one device) race. Returns when phi
   and mode are written. */
int svxg_problem_bond_checks_noalloc(svxg_problem* p, const double* u, double fragility, double* phi, uint8_t* mode);

/* StressProblem::add_force (stress.hpp:134; stress.cpp:504-513), on a host load vector. */
void svxg_add_force(double* f, int32_t node, const double c[3], const double F[3], const double at[3]);
/* StressProblem::extend_warm_start (stress.hpp:115; stress.cpp:213-250) over a graph's nodes and
   bonds: u (6 per node) of nodes first_new.. that have none yet, the mean of their bonded older
   neighbours' motion carried rigidly to their centres, in up to 4 passes. */
void svxg_extend_warm_start(const svxg_node* nodes, int32_t n, const svxg_bond* bonds, int32_t m, double* u,
                            int32_t first_new);

/* Problem facts: free nodes, K blocks, levels, per-level sizes and blocks (up to 16 each),
   StressProblem::work_per_iteration / assembly_work (stress.hpp:95,119), device bytes. */
typedef struct svxg_problem_info {
  int32_t nodes, bonds, nfree, levels;
  int32_t block_jacobi, assembled, preconditioner_current;
  int32_t vcycle_fp16;       /* the V-cycle reads FP16 block copies (svxg_stress_options.vcycle_fp16_rows) */
  int64_t blocks;
  int64_t work_per_iteration, assembly_work;
  int64_t device_bytes;
  int32_t level_rows[16];
  int64_t level_blocks[16];
  int64_t level_pblocks[16];
  int32_t lean;              /* the lean hierarchy (svxg_amg_options.lean_rows): the two finest prolongators
                                smoothed, aggregation at lean_strength, coarsening ended at the first
                                level the dense solve takes */
  int32_t dense_rows;        /* rows of the coarsest level's dense inverse (6 x its nodes), 0 if none */
} svxg_problem_info;
int svxg_problem_info_get(const svxg_problem* p, svxg_problem_info* out);

/* The graph as the library holds it (broken / in_k flags move with remove_bonds / retire_nodes). */
int svxg_problem_get_bonds(const svxg_problem* p, svxg_bond* out);
int svxg_problem_get_nodes(const svxg_problem* p, svxg_node* out);

/* ---- batches ------------------------------------------------------------------------------ */

/* Solves `count` assembled problems in one set of device launches (a batch of the tick's
   problems; the reference runs them one by one, world_pieces.cpp:1452-1463, or in turn,
   world.cpp:986-1129). f[i], u[i]: host vectors of problem i (6 per node); a null f[i] reuses the
   problem's resident load. Each problem's answer is the same alone and in a batch (bitwise:
   every reduction runs in a fixed order over the problem's own rows). */
int svxg_solve_batch(svxg_ctx* ctx, svxg_problem* const* problems, int32_t count, const double* const* f,
                     double* const* u, double rtol, int32_t maxit, int32_t warm, svxg_pcg_result* results);

/* StressProblem::assemble of many problems in one set of device launches (their own options).
   status[i]: SVXG_OK, or SVXG_SINGULAR where the reference's assemble returns false. */
int svxg_assemble_batch(svxg_ctx* ctx, svxg_problem* const* problems, int32_t count, int32_t* status);

/* ---- device-resident views (tests and benchmarks) ----------------------------------------- */

/* K as assembled, in the reference's Bsr6 layout (amg.hpp:20-27): rowptr[nfree+1], col[blocks],
   val[36 blocks] row-major 6x6, diagonal block first in each row. Null pointers: skipped. */
int svxg_problem_get_matrix(const svxg_problem* p, int32_t* rowptr, int32_t* col, double* val);
/* A level's inverse diagonal blocks (Amg::finalize, amg.cpp:227-244: inv6 of each diagonal block,
   regularised by 1e-9 x its |trace| where singular), 36 a row, row-major: the FP64 blocks where the
   level holds them (*fp64 = 1), else its FP32 copies widened (*fp64 = 0). */
int svxg_problem_get_diag_inverse(const svxg_problem* p, int32_t level, double* out, int32_t* fp64);
/* y = K x on the device (Bsr6::apply, amg.cpp:174-183) over free DOFs (6 nfree each), FP64 or,
   with fp32 != 0, through the FP32 copy of K. */
int svxg_problem_apply_matrix(svxg_problem* p, const double* x, double* y, int32_t fp32);
/* z = M^-1 r: one application of the preconditioner (Amg::apply, amg.cpp:718-731; block-Jacobi,
   stress.cpp:312-324) over free DOFs, in the options' precision. */
int svxg_problem_apply_preconditioner(svxg_problem* p, const double* r, double* z);
/* Milliseconds of device time of the last solve (HIP events on the context's stream around the
   solve) and wall milliseconds of the last assemble. */
int svxg_problem_last_timing(const svxg_problem* p, double* solve_ms, double* setup_ms);

/* Kernel timing (no reference counterpart; for benchmarks): with profiling on, the solve drivers
   bracket every launch of the profiled kernels with HIP events on the context's stream.
   kind 0 / 1: the outer K product, FP64 / FP32; kind 2 / 3: a level-0 Chebyshev smoothing step
   (an A product fused with the smoother update), FP64 / FP32. Stats: launches, summed device
   milliseconds and summed algorithmic bytes since the last call (which resets them). */
int svxg_ctx_set_profiling(svxg_ctx* ctx, int32_t on);
/* Graphs with up to `rows` free nodes (0 by default: none) run their solve steps in one
   workgroup each (one launch a window of steps), the rest in tiled launches. The arithmetic, and
   so every bit of a result, is the same either way; this only moves time. No counterpart in the
   reference. */
int svxg_ctx_set_resident_rows(svxg_ctx* ctx, int32_t rows);
/* Tiled solve steps recorded once per batch and precision mix as a HIP graph and replayed, or
   launched kernel by kernel: on = 0 never recorded, 1 (the default) recorded from a batch's third
   window of steps for batches of 8192 or more level-0 tiles (below that, measured, the launches are
   faster), 2 always. Same kernels, same arguments, same bits; the graph saves the host's launches.
   (Profiled solves launch kernel by kernel, and a context created in a process with a rocprofiler
   tool attached - ROCP_TOOL_LIBRARIES set, as rocprofv3 does - starts at 0.) No counterpart in
   the reference. */
int svxg_ctx_set_step_graphs(svxg_ctx* ctx, int32_t on);
int svxg_ctx_kernel_stats(svxg_ctx* ctx, int32_t kind, int64_t* launches, double* total_ms, double* bytes);

/* The autotuner (no counterpart in the reference). Per level shape - what a level's launches serve:
   its rows over the batch (log2), its entries a row, its precision and block storage - it picks the
   tiling of the block-a-lane launches (blocks a workgroup pass: 64, 128, 256 or 512, which sets the
   rows a workgroup takes) of the smoothing and residual, the prolongation and the FP16 restriction,
   and whether a dense coarsest solve of few chunks runs a wavefront a span or a workgroup a chunk.
   A shape met for the first time is tuned online on the batch's first V-cycles: each candidate's
   launches are timed with events in turn (the same kernels on the same rows: a row's sums do not
   depend on its tile, so the answer's bits do not change), and the fastest mean is kept for the
   context's later batches of that shape. on = 0 uses the untuned defaults (256 blocks; split dense
   solves for up to 64 chunks). */
typedef struct svxg_tune_record {
  int32_t kind;              /* 0 smoothing + residual, 1 prolongation, 2 FP16 restriction, 3 dense coarsest solve */
  int32_t fp64;              /* the V-cycle's precision: 1 FP64, 0 FP32 */
  int32_t half;              /* the level's blocks stored as FP16 */
  int32_t rows_log2;         /* the level's rows over the batch, log2 (kind 3: its dense rows) */
  int32_t per_row;           /* entries a row (kind 3: 64-row chunks) */
  int32_t ncand;             /* candidates timed */
  int32_t choice;            /* the index of the one kept */
  int32_t samples;           /* launches timed a candidate */
  int32_t cand[4];           /* kinds 0-2: blocks a pass; kind 3: 0 a workgroup a chunk, 1 a wavefront a span */
  double us[4];              /* each candidate's mean launch time (microseconds) */
  int32_t graphs_log2;       /* the graphs the level's launch spans, log2 */
  int32_t pad_;
} svxg_tune_record;
int svxg_ctx_set_autotune(svxg_ctx* ctx, int32_t on);
/* The shapes tuned so far: min(count, cap) records to out, *count the number tuned. */
int svxg_ctx_tuned(svxg_ctx* ctx, svxg_tune_record* out, int32_t cap, int32_t* count);

/* ---- the judge and break rounds ------------------------------------------------------------
   The World's judgement of a solved structure (World::Impl::judge, world.cpp:1162-1277) and of a
   piece's check (world_pieces.cpp:647-671): every intact bond's loads from the solution, the DIF
   lerp of a sudden change (lerp_load, world_internal.hpp:452-461), bond_utilization, and the
   round's picks - the bonds over strength sorted by phi (descending, ties by bond index), those at
   or above max(1, break_band x max phi) and at least the rule's quota, at most max_breaks - with
   the plastic hinges among them (hinge_of, world.cpp:1286-1334). svxg_break_rounds_batch runs the
   cascade: solve, judge, break the picks in K, detach what lost its supports, solve again, with K,
   the multigrid and the solution resident on the device throughout. */
enum {
  SVXG_JUDGE_STRUCTURE = 0, /* quota max(1, (over + 3) / 4) (world.cpp:1245); broken bonds skipped */
  SVXG_JUDGE_PIECE = 1      /* quota ceil(quota_fraction x over) (world_pieces.cpp:665); bonds that are
                               broken or whose node a is retired skipped (world_pieces.cpp:653) */
};

/* WorldConfig's judge settings (world.hpp:114-183) and Material::ductile (material.hpp:66). */
typedef struct svxg_judge_options {
  double fragility;        /* 1.0: divides every bond strength */
  double dif;              /* 1.5: a shock's loads are L0 + dif (L - L0), L0 the baseline */
  double break_band;       /* 0.85 */
  double quota_fraction;   /* 0.25 (impact_round_fraction; SVXG_JUDGE_PIECE), clamped to [0, 1] */
  double hinge_shape;      /* 1.3: a hinge holds hinge_shape x the section's elastic moment */
  uint64_t ductile[2];     /* bit k of word k / 64: material k is ductile (hinges form in it) */
  int32_t rule;            /* SVXG_JUDGE_STRUCTURE or SVXG_JUDGE_PIECE */
  int32_t max_breaks;      /* 256: picks of a round at most (1..4096) */
  int32_t max_rounds;      /* 400: a structure past this many rounds is judged, nothing breaks */
  int32_t plastic_hinges;  /* 1: hinge cuts of the picks */
  int32_t restart_iters;   /* cascade: a solve on a stale multigrid that has not converged after
                              this many iterations (120) gets it rebuilt and goes on (world.cpp:1068-1093) */
  int32_t rebuild_iters;   /* cascade: one that converged after more (60) rebuilds it for the next round */
} svxg_judge_options;      /* 80 bytes */
void svxg_default_judge_options(svxg_judge_options* out);

/* A structure's cascade state (Structure::shock, transient, rounds). */
typedef struct svxg_judge_state {
  int32_t shock;      /* the loads changed suddenly: bonds with a baseline are judged on the DIF lerp */
  int32_t transient;  /* an impact load case: judged, the baseline is left as it was */
  int32_t rounds;     /* break rounds of the running cascade so far */
  int32_t pad_;
} svxg_judge_state;

/* HingeCut (world_impl.hpp:372-378): the pivot p, the axis the sides turn about, the bond's
   normal n, the plastic moment mp (N m) and the force that tears it (N). */
typedef struct svxg_hinge_cut {
  int32_t a, b;
  double p[3], axis[3], n[3];
  double mp, pull;
} svxg_hinge_cut;          /* 96 bytes */

typedef struct svxg_break {
  int32_t bond;
  int32_t round;      /* the cascade round it broke in (svxg_judge_state.rounds then) */
  double phi;
  uint8_t mode;       /* FailMode: 1 Tension, 2 Crush, 3 Shear */
  uint8_t hinge;      /* 1: cut is a plastic hinge */
  uint8_t pad_[6];
  svxg_hinge_cut cut;
} svxg_break;              /* 120 bytes */

typedef struct svxg_judge_result {
  double max_phi;     /* over the judged bonds (WorldStats::max_utilization) */
  int32_t over;       /* judged bonds with phi >= 1 */
  int32_t picked;     /* bonds this round breaks (0: none over, or rounds >= max_rounds) */
  int32_t hinges;     /* picks with a hinge cut */
  int32_t pad_;
} svxg_judge_result;       /* 24 bytes */

/* The judged loads a shock is measured from (World::Impl::judged_, world.cpp:1176-1190), per bond
   (svxg_bond_load), has[b] = 0 where a bond has none. loads == NULL clears them (none has one). */
int svxg_problem_set_judge_baseline(svxg_problem* p, const svxg_bond_load* loads, const uint8_t* has);
int svxg_problem_get_judge_baseline(svxg_problem* p, svxg_bond_load* loads, uint8_t* has);

/* One judgement of `count` problems (World::Impl::judge without its side effects): the loads
   from u[i] (host, 6 per node) or, where u or u[i] is null, from the problem's device-resident
   solution; the baseline becomes these loads unless state[i].transient. picks[i] (room for
   max_breaks) receives the round's picks in order, round = state[i].rounds. phi[i] (optional, one
   per bond): each judged bond's utilisation, 0 for the others. Nothing breaks: the caller removes
   the picks (svxg_problem_remove_bonds), as the World does. */
int svxg_judge_batch(svxg_ctx* ctx, svxg_problem* const* problems, int32_t count, const double* const* u,
                     const svxg_judge_options* opt, const svxg_judge_state* state, svxg_judge_result* res,
                     svxg_break* const* picks, double* const* phi);

enum {
  SVXG_CASCADE_SETTLED = 0, /* a judgement found no bond over strength */
  SVXG_CASCADE_ROUNDS = 1,  /* state.rounds reached max_rounds */
  SVXG_CASCADE_UNSOLVED = 2,/* a solve did not converge (after a rebuild) or broke down */
  SVXG_CASCADE_GONE = 3,    /* every node retired */
  SVXG_CASCADE_RESEED = 4,  /* more than max(64, active / 2) nodes retired: the World extracts the
                               structure again (world.cpp:1488-1493) */
  SVXG_CASCADE_LIMIT = 5    /* the call's round limit */
};

typedef struct svxg_cascade_result {
  svxg_pcg_result last;     /* the last solve */
  double max_phi;           /* of the last judgement */
  int32_t end;              /* SVXG_CASCADE_* */
  int32_t rounds;           /* rounds that broke bonds in this call */
  int32_t breaks;           /* bonds broken by the judge (records written: min(breaks, capacity)) */
  int32_t retired;          /* nodes retired by detaching (ids written: min(retired, n)) */
  int32_t solves, iters;    /* solves and PCG iterations */
  int32_t rebuilds;         /* multigrid rebuilds (stale after breaks) */
  int32_t pad_;
} svxg_cascade_result;     /* 64 bytes */

/* The break-round cascade of `count` structures (World::Impl::step_structures' solve, judge and
   detach_unsupported, world.cpp:986-1129, 1162-1277, 1371-1444, one round after another). Each
   round: solve K u = f warm from the last u (maxit iterations, rtol; a stale multigrid is rebuilt
   as opt.restart_iters / rebuild_iters say), judge (as svxg_judge_batch, with state[i], which the
   call advances: shock after a round that broke, rounds counted), break the picks in K, then
   retire the nodes no longer connected to a support bond (graph_components,
   world_internal.hpp:660-706; their loads and motion become 0, as those of nodes retired before
   the call, world.cpp:1472-1480). Ends per structure as
   svxg_cascade_result.end says, or after max_call_rounds rounds (0: no limit). f[i], u[i]: host,
   6 per node; u[i] is the warm start and receives the last solution. breaks[i] (capacity[i]
   records) and retired[i] (room for every node; may be null) receive what broke and what left.
   Every problem must be assembled with one precision; the judge runs in FP64. */
int svxg_break_rounds_batch(svxg_ctx* ctx, svxg_problem* const* problems, int32_t count, const double* const* f,
                            double* const* u, double rtol, int32_t maxit, const svxg_judge_options* opt,
                            int32_t max_call_rounds, svxg_judge_state* state, svxg_cascade_result* res,
                            svxg_break* const* breaks, const int32_t* capacity, int32_t* const* retired);

/* ---- bond sections from voxel faces ----------------------------------------------------------
   The geometry and strengths of a structure's bonds from the voxel faces between its fragments
   (SecAcc::finish_lattice, world_internal.hpp:129-249, and section_strengths, :374-400, called by
   World::Impl::append_nodes, world.cpp:835-850): per bond, the interface normal (the faces' mean
   normal from a to b; a support's face axis), the section's centroid, principal axes, area, second
   moments, extreme fibres and torsion radius, the spring lengths la / lb from the two centres, and
   the section strengths (each face as strong as the weaker of its two voxels' materials, times the
   condition of the more damaged; the faces' mean). One lattice, unrotated (the reference's
   finish_lattice case); svxg_bond_sections_general below takes placed lattices and junctions. */

/* Vox (core/include/svx/world/grid.hpp:28-36): 0 air; else material id + 1 in the low 7 bits,
   0x80 for an anchored (ground) voxel. A box of voxels lo .. lo + dim - 1, x fastest; voxels
   outside it are air. */
typedef struct svxg_voxel_box {
  int32_t lo[3];
  int32_t dim[3];
  double h;                  /* voxel size (m) */
  const uint8_t* vox;        /* dim[0] * dim[1] * dim[2] values */
  const uint8_t* damage;     /* the same count, 0 intact .. 255 no strength left; or null */
} svxg_voxel_box;

/* One bond's accumulated faces (SecHead / SecAcc, world_internal.hpp:105-125): nodes a, b (b < 0:
   a support, its face axis and a's side sign), faces [face0, face0 + nfaces) of the face list, the
   two sides' materials and strength multipliers (class_mult; a support's strength_b 1e9). */
typedef struct svxg_section_head {
  int32_t a, b;
  int32_t axis, sign;
  int32_t face0, nfaces;
  uint8_t ma, mb;
  uint8_t pad_[6];
  double strength_a, strength_b;
} svxg_section_head;       /* 48 bytes */

/* A face: its lower voxel p (the face joins p and p + e_axis) and whether a is on that lower side
   (+1: the face's normal from a to b is +e_axis). */
typedef struct svxg_section_face {
  int32_t p[3];
  int8_t axis;
  int8_t sign;
  uint8_t pad_[2];
} svxg_section_face;       /* 16 bytes */

/* out[k] (m bonds): bond k of heads[k], as append_nodes makes it: a, b, ma, mb, faces, the frame,
   area, p, s1, s2, c1, c2, rmax, la, lb, strength = min(strength_a, strength_b), the section
   strengths (sectioned = 1) and tag = k; broken / in_k 0. nodes: the graph's node centres. */
int svxg_bond_sections(svxg_ctx* ctx, const svxg_node* nodes, int32_t n, const svxg_section_head* heads, int32_t m,
                       const svxg_section_face* faces, int32_t nfaces, const svxg_voxel_box* box,
                       const svxg_material* mats, int32_t nmats, svxg_bond* out);

/* ---- bond sections across lattices: placed grids and junction samples -----------------------
   The general case of the same (SecAcc::finish, world_internal.hpp:120-127, finish_general,
   :250-362, and section_strengths_general, :402-441; called by append_nodes, world.cpp:824-852, and
   the piece graphs, world_pieces.cpp:429-440): a bond's faces lie in one lattice (its `grid`),
   which may be placed in the bond's frame by a rotation and an offset, and it may hold junction
   samples - S x S points of a voxel face of one lattice that land in a voxel of another - each a
   small square of side h / S. A bond with no samples on an unplaced lattice takes the one-lattice
   formulas (finish_lattice) exactly as svxg_bond_sections. */

/* A lattice (LatticeXf, core/include/svx/phys/rigid.hpp:85-95, with its voxels): to(s) = off + R s
   in the frame (R row-major, as LatticeXf::make leaves it), identity = 1 for an unplaced lattice
   (off and R unused); voxel size h; its voxels lo .. lo + dim - 1 (x fastest; outside: air). */
typedef struct svxg_lattice {
  double off[3];
  double R[9];
  double h;
  int32_t identity;
  int32_t lo[3];
  int32_t dim[3];
  int32_t pad_;
  const uint8_t* vox;        /* dim[0] * dim[1] * dim[2] Vox codes */
  const uint8_t* damage;     /* the same count, or null */
} svxg_lattice;              /* 152 bytes */

/* A bond's accumulator in the general case (SecHead / SecAcc, world_internal.hpp:105-125): the
   svxg_section_head fields, its faces' lattice `grid`, and samples [sample0, sample0 + nsamples)
   of the sample list. A support with samples (b < 0, a junction support) takes its normal from
   them. */
typedef struct svxg_section_head_general {
  int32_t a, b;
  int32_t axis, sign;
  int32_t face0, nfaces;
  int32_t sample0, nsamples;
  uint16_t grid;
  uint8_t ma, mb;
  uint8_t pad_[4];
  double strength_a, strength_b;
} svxg_section_head_general; /* 56 bytes */

/* A junction sample (JSample with its accumulator's jsg and jw, world_internal.hpp:78-84, 114-117):
   sample `sub` (0 .. S*S-1) of face `face` (axis * 2 + 1 for the +axis face) of voxel v of lattice
   vg lands in voxel o of lattice og; kind 1: in the unknown world (its strength is its own voxel's);
   sign +1 when the face's normal points from a to b (or out of a); weight: its share of the area. */
typedef struct svxg_junction_sample {
  int32_t v[3], o[3];
  uint16_t vg, og;
  uint8_t face, sub, kind;
  int8_t sign;
  float weight;
} svxg_junction_sample;      /* 36 bytes */

/* out[k] (m bonds): bond k as SecAcc::finish then section_strengths(_general) make it (the fields
   of svxg_bond_sections; faces = faces + samples). S: samples per face edge (the caller's
   junction_samples clamped to 1..7, world.cpp:825). nodes: the graph's node centres in the frame. */
int svxg_bond_sections_general(svxg_ctx* ctx, const svxg_node* nodes, int32_t n,
                               const svxg_section_head_general* heads, int32_t m, const svxg_section_face* faces,
                               int32_t nfaces, const svxg_junction_sample* samples, int32_t nsamples,
                               const svxg_lattice* lattices, int32_t nlattices, int32_t S, const svxg_material* mats,
                               int32_t nmats, svxg_bond* out);

/* A structure's graph from one voxel lattice, all on the device (the walk of World::Impl::extract
   and append_nodes, world.cpp:425-865, for one unrotated lattice and a node a fragment): frag holds
   each voxel's fragment (0 .. nfrag-1 for solid unanchored voxels, -1 elsewhere), box the Vox codes
   (and damage). nodes[f] (nfrag): the fragment's centre of mass and mass (the materials' densities,
   voxel centres at h times their integer coordinates, fragments.cpp:315). The bonds: one a pair of
   touching fragments and one a (fragment, axis, side) resting on an anchored voxel, ordered by
   (node a, pairs before supports, node b / axis and side), their faces in the box's scan order
   (x fastest, then the axis); heads / faces: the svxg_bond_sections inputs, bonds: its outputs.
   *nbonds, *nfaces: in, the room of heads and bonds / of faces; out, the counts. With heads, faces
   and bonds all null only the counts (and nodes) are made; SVXG_BAD_INPUT when the room is short
   (the counts are set). */
int svxg_lattice_graph(svxg_ctx* ctx, const svxg_voxel_box* box, const int32_t* frag, int32_t nfrag,
                       const svxg_material* mats, int32_t nmats, svxg_node* nodes, int32_t* nbonds, int32_t* nfaces,
                       svxg_section_head* heads, svxg_section_face* faces, svxg_bond* bonds);

/* The same with fragments clustered into nodes (append_nodes, world.cpp:741-800, for one lattice):
   touching fragments whose first voxels (the least (x, y, z) of their voxels, x first) share a cell
   make one node, clusters numbered in order of their least fragment (cluster_items,
   world_internal.hpp:466-491). The cell: with cubes != 0 a cube of `cell` voxels; else the
   reference's default key, whose packing drops the cell along x (WorldConfig::cluster_cubes off,
   world.hpp:234-241): `cell` voxels in y and z and a chunk (32 voxels) in x. cell: 0 (a node a
   fragment, as svxg_lattice_graph), or a divisor of 32 (World::Impl::cluster_cell gives 8 or 16).
   A node's mass is its fragments' masses summed in index order, its centre their centres weighted
   by their masses (summed, then times 1 / mass), its material the heaviest fragment's; faces between
   fragments of one node drop out. frag_node (nfrag, or null) receives each fragment's node and
   *nnodes the node count; nodes then has room for *nnodes. The two-call protocol of
   svxg_lattice_graph. */
int svxg_lattice_graph_clustered(svxg_ctx* ctx, const svxg_voxel_box* box, const int32_t* frag, int32_t nfrag,
                                 int32_t cell, int32_t cubes, const svxg_material* mats, int32_t nmats,
                                 int32_t* frag_node, int32_t* nnodes, svxg_node* nodes, int32_t* nbonds,
                                 int32_t* nfaces, svxg_section_head* heads, svxg_section_face* faces,
                                 svxg_bond* bonds);

/* ---- a structure's extraction: the walk from a seed fragment ----------------------------------
   World::Impl::extract's walk and append_nodes (world.cpp:425-620, 739-865) for a structure of the
   world grid (grid 0, no junctions), on the device, in the reference's order: the members in the
   order the walk takes them (breadth first from the seed; a member's voxels in its chunk's scan
   order, x slowest; each voxel's faces -x, +x, -y, +y, -z, +z), its bonds in the order their first
   faces are met and their faces in the order met (SectionLog, world_internal.hpp:533-620), nodes
   clustered as append_nodes clusters them, a pair bond's ends in the order of their identities.
   A fragment is a FragInfo of one chunk (fragments.cpp:283-345): chunk-local, 32^3 voxels a
   chunk. */
typedef struct svxg_fragment {
  double com[3];             /* FragInfo::com (world, m) */
  double mass;               /* FragInfo::mass */
  double strength;           /* class_mult(frag_class(f)) (world.cpp:408-421; 1 undesigned) */
  int64_t owner;             /* owner_of(f): the registered structure that holds it, 0 none */
  int32_t first[3];          /* FragInfo::first as world voxel coordinates (its chunk is the fragment's) */
  uint8_t mat;               /* FragInfo::mat */
  uint8_t pad_[3];
} svxg_fragment;             /* 64 bytes */

typedef struct svxg_extract_options {
  int32_t max_nodes;         /* WorldConfig::structure_max_nodes (60000) */
  int32_t cluster_nodes;     /* WorldConfig::cluster_nodes (2500), cluster_cell's limit */
  double max_radius;         /* WorldConfig::structure_max_radius (60 m) from the seed's centre */
  int64_t fresh_from;        /* owners >= fresh_from are structures of this refresh: held fixed
                                (World::Impl::fresh_from_; INT64_MAX: none) */
  int32_t cluster_cubes;     /* WorldConfig::cluster_cubes (0) */
  int32_t outside;           /* voxels outside the box: 0 air; 1 chunks not resident, which hold
                                what touches them (the walk's `unloaded`) */
} svxg_extract_options;      /* 32 bytes */

typedef struct svxg_extract_result {
  int32_t members;           /* fragments taken, in walk order */
  int32_t nodes;             /* after clustering */
  int32_t bonds;
  int32_t faces;
  int32_t cell;              /* cluster_cell(members): 0, 8 or 16 */
  int32_t truncated;         /* a fragment met was held fixed (radius, node limit, a fresh owner) */
  int32_t any_support;       /* 0: nothing holds it - extract makes no structure (a free piece) */
  int32_t touched_unloaded;  /* it reached outside the box with outside = 1 */
} svxg_extract_result;

void svxg_extract_default_options(svxg_extract_options* out);

/* The walk from fragment `seed` over box (Vox codes, damage), broken (per voxel, bit a: the face to
   its +a neighbour is broken; or null) and frag (each voxel's fragment, -1 none; fragments only on
   solid unanchored voxels, each inside the chunk of its `first`). Outputs, each with room for what
   res holds on entry (members, nodes, bonds, faces):
     members[k]      the fragment of member k (walk order)
     member_node[k]  its node
     nodes           centre and mass (append_nodes: masses summed in member order, centres weighted
                     by them, then times 1 / mass)
     node_ident      frag_ident of each node's first fragment (world_internal.hpp:24)
     heads, faces    each bond's section head (ma, mb and strengths as append_nodes sets them) and
                     faces, the svxg_bond_sections inputs
     bonds           the bonds, sections computed (SecAcc::finish_lattice + section_strengths)
     bond_ident      bond_ident of each bond (world_internal.hpp:25-27)
   With every output null only the counts are made (res); SVXG_BAD_INPUT when a room is short (the
   counts are set). Any output may be null. */
int svxg_extract_structure(svxg_ctx* ctx, const svxg_voxel_box* box, const uint8_t* broken, const int32_t* frag,
                           const svxg_fragment* frags, int32_t nfrag, int32_t seed, const svxg_extract_options* opt,
                           const svxg_material* mats, int32_t nmats, svxg_extract_result* res, int32_t* members,
                           int32_t* member_node, svxg_node* nodes, uint64_t* node_ident, svxg_section_head* heads,
                           svxg_section_face* faces, svxg_bond* bonds, uint64_t* bond_ident);

/* ---- serving: a per-tick request loop with continuous batching ------------------------------
   Replaces the reference's per-tick structure loop (step_structures, world.cpp:986-1129: the
   structures of a tick solved one after another within a block-operation budget) and its piece
   checks (world_pieces.cpp:1452-1463). A server owns two worker threads and two streams of the
   context: one assembles what arrives (bond removals, assembly and multigrid setup, loads) while
   the other steps the running batch; requests join the running batch, and finished ones leave it,
   at its next look (every few PCG steps), so work that arrives while others are on the GPU
   overlaps with them. Every graph does the same arithmetic as alone: a result is bitwise the same
   whatever else is in the batch. Admission is by the context's HBM budget: a request whose problem
   must be assembled waits until its estimated bytes fit. A problem submitted to a server belongs
   to it until its request completes: no other call may use it meanwhile. */
typedef struct svxg_server svxg_server;

typedef struct svxg_server_options {
  int32_t precision;      /* SVXG_MIXED (default) or SVXG_FP64 */
  int32_t max_active;     /* graphs in the running batch at most (4096) */
  double rtol;            /* relative residual of every solve (1e-8) */
  int32_t maxit;          /* iterations of a solve at most (5000) */
  int32_t amg_min_nodes;  /* assembly option of the problems a server assembles (160) */
  int32_t window;         /* PCG steps between looks at most (8) */
  /* the reference's rebuild policy for stale preconditioners (world.cpp:1068-1110): a solve on a
     stale multigrid or a block-Jacobi one stops after restart_iters (120) iterations, gets its
     preconditioner built again (multigrid) and goes on from where it is (at most twice); one that
     converged after more than rebuild_iters (60) on one rebuilds it for its next request. 0
     turns either off. */
  int32_t restart_iters;
  int32_t rebuild_iters;
  int32_t setup_workers;  /* setup workers (2), each with its own stream: groups of arrivals set up side by side */
  /* step_structures' failure policy (world.cpp:1056-1100). Every run of a solve has a budget of
     restart_iters iterations (0: none); at the budget, unconverged, it restarts from its current
     iterate - on a rebuilt preconditioner when its own is stale (bonds broke since it was built) or
     block-Jacobi, else on the same one. With restart_diverging (1), a run that ends at its budget with
     a residual of 10 x |b| or more, on a preconditioner it may rebuild, restarts on a rebuilt one from
     the problem's last solution (0 without one). A breakdown (a non-finite value, a direction of no
     curvature) rebuilds the preconditioner, zeroes what is not finite in the load and the last
     solution, retires the nodes that lost every path to a support bond (detach_unsupported,
     world.cpp:1371-1380) and restarts from that solution. After solve_restarts (60) restarts in a
     row the solve is abandoned: the request completes unconverged with `abandoned` set (0: no
     restarts at all - a breakdown or a divergence is final, a run at its budget is abandoned). */
  int32_t solve_restarts;
  int32_t restart_diverging;
  /* 1: each converged request is judged on its solution (World::Impl::judge, world.cpp:1111) with
     judge_opt and state {shock: the request removed bonds, not transient, round 0}: the counters in
     svxg_serve_done.judge, the picks through svxg_server_picks. 0 (default): not judged. */
  int32_t judge;
  int32_t pad_;
  svxg_judge_options judge_opt;
} svxg_server_options;

typedef struct svxg_serve_done {
  uint64_t ticket;
  svxg_problem* problem;
  svxg_pcg_result result;
  int32_t status;         /* SVXG_OK, or what kept the request from solving (SVXG_SINGULAR, ...) */
  int32_t assembled;      /* 1: this request assembled its problem (a first solve or a re-setup) */
  double t_submit_ms;     /* on the server's clock (svxg_server_now_ms) */
  double t_ready_ms;      /* setup and loads done */
  double t_admit_ms;      /* joined the running batch */
  double t_done_ms;       /* solution written */
  int32_t restarts;       /* runs restarted (budget, divergence, breakdown) before this result */
  int32_t abandoned;      /* 1: left unconverged after solve_restarts restarts */
  int32_t breakdowns;     /* runs that broke down and were restarted */
  int32_t detached;       /* nodes retired because they lost every path to a support */
  svxg_judge_result judge;/* the judgement of a converged request (svxg_server_options.judge) */
} svxg_serve_done;

typedef struct svxg_server_stats {
  int64_t submitted, completed, failed;
  int64_t steps;          /* PCG steps of the running batch */
  int64_t looks;          /* synchronisations of the running batch (admissions, completions) */
  int64_t assembled;      /* problems assembled */
  int64_t max_active;     /* largest running batch */
  double step_ms;         /* engine time stepping and looking (rebuilds included) */
  double rebuild_ms;      /* engine time rebuilding the running batch and starting new solves */
  double setup_ms;        /* setup-worker time: removals, assemblies, loads */
  double assemble_ms;     /* of which assemblies */
  int64_t setup_groups;   /* groups of requests the setup worker took at once */
  int64_t rebuilds;
  int64_t restarts;       /* runs restarted (world.cpp:1068-1110) */
  int64_t solves_abandoned; /* requests left unconverged after solve_restarts restarts (WorldStats::solves_abandoned) */
  int64_t breakdowns;     /* runs that broke down */
  int64_t diverged;       /* runs restarted for a growing residual */
  int64_t judged;         /* requests judged */
} svxg_server_stats;

void svxg_server_default_options(svxg_server_options* out);
int svxg_server_create(svxg_ctx* ctx, const svxg_server_options* opt, svxg_server** out);
/* Waits for every submitted request to complete, then stops the workers. */
int svxg_server_destroy(svxg_server* srv);
/* One solve of problem p: its bonds `remove` (count nremove, may be 0) leave K first (in place, as
   svxg_problem_remove_bonds; an unassembled problem is assembled without them), then K u = f is
   solved to the server's rtol. f: 6 per node (all nodes), or null to keep the problem's resident
   load. warm: start from the problem's resident solution (else from 0). u: where the solution
   (6 per node, all nodes) is written when the request completes, or null (it stays resident).
   f is copied; u must stay valid until completion. Returns SVXG_BAD_INPUT if p is in flight. */
int svxg_server_submit(svxg_server* srv, svxg_problem* p, const double* f, double* u, int32_t warm,
                       const int32_t* remove, int32_t nremove, uint64_t* ticket);
/* count requests at once (arrays of count; f[i], u[i], remove[i] may be null; warm, nremove may be
   null for 0): as many svxg_server_submit calls, all or none (SVXG_BAD_INPUT names the first bad one
   in svxg_last_error). */
int svxg_server_submit_many(svxg_server* srv, int32_t count, svxg_problem* const* problems, const double* const* f,
                            double* const* u, const int32_t* warm, const int32_t* const* remove, const int32_t* nremove,
                            uint64_t* tickets);
/* The picks of problem p's last judgement by the server (svxg_server_options.judge), in order:
   min(count, cap) records to out, *count the number picked. Valid once its request completed. */
int svxg_server_picks(svxg_server* srv, const svxg_problem* p, svxg_break* out, int32_t cap, int32_t* count);
/* Completed requests, oldest first (up to cap); waits up to wait_ms for at least one. */
int svxg_server_poll(svxg_server* srv, svxg_serve_done* out, int32_t cap, int32_t* count, int32_t wait_ms);
/* Destroys p (its device memory returns to the budget) once no request of it is in flight. */
int svxg_server_release(svxg_server* srv, svxg_problem* p);
double svxg_server_now_ms(const svxg_server* srv);
int svxg_server_stats_get(svxg_server* srv, svxg_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* SVXG_H */
