/* oracle — CPU restatement of the reference's fragment-graph stress solve.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the device library is compared with. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product path
 * (libsvxg.so and the opus_destruct_2_amd package) never does.
 *
 * The same entry points, prefixed ref_ instead of orc_, are exported by oracle/_ref/libsvxref.so:
 * the reference's own sources (core/src/stress/stress.cpp, core/src/solve/amg.cpp,
 * core/src/base/parallel.cpp, core/src/material/material.cpp) compiled by oracle/Makefile behind
 * oracle/ref_shim.cpp.
 */
#ifndef SVX_ORACLE_H
#define SVX_ORACLE_H

#include "svxg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

orc_problem* orc_create(const svxg_node* nodes, int32_t n, const svxg_bond* bonds, int32_t m, const svxg_material* mats,
                        int32_t nmats);
void orc_destroy(orc_problem* h);
int orc_assemble(orc_problem* h, const svxg_stress_options* opt);
void orc_release(orc_problem* h);
int orc_solve(orc_problem* h, const double* f, double* u, double rtol, int32_t maxit, int32_t warm, svxg_pcg_result* res);
void orc_begin(orc_problem* h, const double* f, const double* u0);
void orc_iterate(orc_problem* h, int32_t maxit, double rtol, svxg_pcg_result* res);
void orc_current(orc_problem* h, double* u);
void orc_remove_bond(orc_problem* h, int32_t b);
void orc_retire_nodes(orc_problem* h, const int32_t* ids, int32_t count);
void orc_set_bond_broken(orc_problem* h, int32_t b, int32_t broken);
void orc_bond_load(orc_problem* h, int32_t b, const double* u, svxg_bond_load* out);
/* every bond's loads (loads, may be null) and, with phi, its utilisation and FailMode (mode may be null) */
void orc_bond_checks(orc_problem* h, const double* u, double fragility, svxg_bond_load* loads, double* phi,
                     uint8_t* mode);
void orc_extend_warm_start(const svxg_node* nodes, int32_t n, const svxg_bond* bonds, int32_t m, double* u,
                           int32_t first_new);
double orc_bond_utilization(const svxg_bond* b, const svxg_bond_load* L, double fragility, const svxg_material* mats,
                            int32_t nmats, uint8_t* mode);
void orc_bond_strength(const svxg_bond* b, double fragility, const svxg_material* mats, int32_t nmats, double out[5]);
void orc_info(orc_problem* h, svxg_problem_info* out);
void orc_get_matrix(orc_problem* h, int32_t* rowptr, int32_t* col, double* val);
void orc_apply_matrix(orc_problem* h, const double* x, double* y);
void orc_apply_preconditioner(orc_problem* h, const double* r, double* z);
void orc_get_bonds(orc_problem* h, svxg_bond* out);
void orc_get_nodes(orc_problem* h, svxg_node* out);
void orc_set_threads(int32_t n);
/* World::Impl::judge without its side effects (world.cpp:1162-1277; the piece rule of
   world_pieces.cpp:647-671): the judged bonds' loads from u (6 per node), the DIF lerp against
   base/has (updated unless st->transient), phi (phi_out: per bond, 0 where not judged), the round's
   picks (room for opt->max_breaks) with their hinge cuts. Returns the number of picks. */
int32_t orc_judge(orc_problem* h, const double* u, const svxg_judge_options* opt, const svxg_judge_state* st,
                  svxg_bond_load* base, uint8_t* has, double* phi_out, svxg_judge_result* res, svxg_break* picks);
/* hinge_of (world.cpp:1286-1334); ductile as svxg_judge_options.ductile. Returns 1 with *out set. */
int32_t orc_hinge_of(const svxg_bond* b, const svxg_bond_load* L, const double rot_a[3], const double rot_b[3],
                     const svxg_material* mats, int32_t nmats, const uint64_t ductile[2], double fragility,
                     double hinge_shape, svxg_hinge_cut* out);
/* lerp_load (world_internal.hpp:452-461) */
void orc_lerp_load(const svxg_bond_load* prev, const svxg_bond_load* now, double k, svxg_bond_load* out);
/* SecAcc::finish_lattice + section_strengths (world_internal.hpp:129-249, 374-400) of every head, as
   World::Impl::append_nodes makes the bonds (world.cpp:835-850): the restatement of
   svxg_bond_sections. */
void orc_bond_sections(const svxg_node* nodes, int32_t n, const svxg_section_head* heads, int32_t m,
                       const svxg_section_face* faces, int32_t nfaces, const svxg_voxel_box* box,
                       const svxg_material* mats, int32_t nmats, svxg_bond* out);
/* graph_components (world_internal.hpp:660-706): comp[n]; returns the number of components */
void orc_bond_sections_general(const svxg_node* nodes, int32_t n, const svxg_section_head_general* heads, int32_t m,
                               const svxg_section_face* faces, int32_t nfaces, const svxg_junction_sample* samples,
                               int32_t nsamples, const svxg_lattice* lattices, int32_t nlattices, int32_t S,
                               const svxg_material* mats, int32_t nmats, svxg_bond* out);
int32_t orc_graph_components(int32_t n, const svxg_bond* bonds, int32_t m, const uint8_t* seed, int32_t* comp);
/* World::Impl::extract's walk and append_nodes (world.cpp:425-620, 739-865) for the world grid over a
   box: the restatement of svxg_extract_structure, taken sequentially as the reference takes it */
int32_t orc_extract_structure(const svxg_voxel_box* box, const uint8_t* broken, const int32_t* frag,
                              const svxg_fragment* frags, int32_t nfrag, int32_t seed, const svxg_extract_options* opt,
                              const svxg_material* mats, int32_t nmats, svxg_extract_result* res, int32_t* members,
                              int32_t* member_node, svxg_node* nodes, uint64_t* node_ident, svxg_section_head* heads,
                              svxg_section_face* faces, svxg_bond* bonds, uint64_t* bond_ident);
/* the SplitMix64 output function the identities use (world_internal.hpp:17-23) */
uint64_t orc_mix64_value(uint64_t x);

void orc_default_options(svxg_stress_options* out);
/* the 6x6 block inverse of a level's diagonal (amg.cpp:105-135 blk6::inv6): 1 and out = A^-1, or 0
   (zero or singular at the reference's pivot tolerance) */
int32_t orc_inv6(const double* A, double* out);

#ifdef __cplusplus
}
#endif

#endif
