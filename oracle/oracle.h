/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the javagg/rmetis multilevel k-way greedy path
 * (SHEM/RM matching, contraction, BFS growing, greedy refinement, edge-cut
 * and part-weight reductions).  It is the checker the HIP product is compared
 * against; nothing in the product (rmetis_amd/, include/) may link or call it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference crate root).  All index/weight data is int32 (types.rs:4);
 * int32 sums wrap as in the reference's release build (Cargo.toml:42-44).
 *
 * Parity status of each function is recorded in oracle.cpp's header.
 */
#ifndef RMETIS_ORACLE_H
#define RMETIS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* rand 0.8.5 SmallRng (= Xoshiro256PlusPlus on 64-bit targets). */
typedef struct { uint64_t s[4]; } orc_rng;
void     orc_rng_seed_from_u64(orc_rng* r, uint64_t seed);
void     orc_rng_from_words(orc_rng* r, const uint64_t w[4]);
uint64_t orc_rng_next_u64(orc_rng* r);
uint32_t orc_rng_next_u32(orc_rng* r);
uint64_t orc_rng_range_incl(orc_rng* r, uint64_t low, uint64_t high);
uint64_t orc_splitmix64_next(uint64_t* state);

/* graph/mod.rs reductions */
int32_t orc_edge_cut(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                     const int32_t* adjwgt, const int32_t* part);
void    orc_partition_weights(int32_t n, int32_t ncon, const int32_t* vwgt,
                              const int32_t* part, int32_t nparts, int32_t* pw);
void    orc_total_weight(int32_t n, int32_t ncon, const int32_t* vwgt, int32_t* tw);
int32_t orc_boundary_vertices(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                              const int32_t* part, int32_t* out);
void    orc_compute_target_weights(int32_t n, int32_t ncon, const int32_t* vwgt,
                                   int32_t nparts, const float* tpwgts, double* out);
int32_t orc_validate_csr(int32_t n, int32_t ncon, const int32_t* xadj, int64_t xadj_len,
                         const int32_t* adjncy, int64_t adjncy_len,
                         const int32_t* vwgt, int64_t vwgt_len,
                         const int32_t* adjwgt, int64_t adjwgt_len);

/* refine/greedy.rs */
int32_t orc_compute_move_gain(const int32_t* xadj, const int32_t* adjncy,
                              const int32_t* adjwgt, int32_t v, const int32_t* part,
                              int32_t from, int32_t to);
/* One synchronous greedy pass: mode 0 = balance repair (greedy.rs:60-123),
 * mode 1 = one cut-optimisation round (greedy.rs:127-182).  pw [nparts*ncon]
 * is the running weight vector (in/out), part is updated with the recorded
 * moves; moves_out receives (v, p) pairs in admission order.  Returns #moves. */
int32_t orc_greedy_pass(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                        const int32_t* vwgt, const int32_t* adjwgt, int32_t* part,
                        int32_t nparts, const double* tpwgts, const double* ubvec,
                        int32_t mode, int32_t* pw, int32_t* moves_out);
int32_t orc_greedy_refine(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                          const int32_t* vwgt, const int32_t* adjwgt, int32_t* part,
                          int32_t nparts, const double* tpwgts, const double* ubvec,
                          int32_t niter);

/* refine/fm_kway.rs:12-30 (fm_kway_refine) and one pass (33-158).  The
 * queue pops equal gains smallest vertex first (the reference's order among
 * equal gains is internal to the priority-queue crate).  The pass returns
 * the reference's boolean and *nmoves the moves it kept; pw (in/out) is the
 * running part-weight vector. */
int32_t orc_fm_kway_refine(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                           const int32_t* vwgt, const int32_t* adjwgt, int32_t* part, int32_t nparts,
                           const double* tpwgts, const double* ubvec, int32_t niter);
int32_t orc_fm_kway_pass(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                         const int32_t* vwgt, const int32_t* adjwgt, int32_t* part, int32_t nparts,
                         const double* tpwgts, const double* ubvec, int32_t* pw, int32_t* nmoves);

/* coarsen/shem.rs, coarsen/rm.rs: matching[v] = partner (or v). */
void orc_shem_matching(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                       const int32_t* adjwgt, orc_rng* rng, int32_t* matching);
void orc_random_matching(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                         orc_rng* rng, int32_t* matching);

/* graph/mod.rs:284-384.  Output buffers sized by the caller: cxadj n+1,
 * cadjncy/cadjwgt xadj[n], cvwgt n*ncon, cvsize n, cmap n.  Returns cn;
 * *cm receives the coarse entry count; *flags bit0 = coarse vwgt present,
 * bit1 = coarse vsize present, bit2 = coarse adjwgt present (not None). */
int32_t orc_build_coarse_graph(int32_t n, int32_t ncon, const int32_t* xadj,
                               const int32_t* adjncy, const int32_t* vwgt,
                               const int32_t* adjwgt, const int32_t* vsize,
                               const int32_t* matching, int32_t* cxadj, int32_t* cadjncy,
                               int32_t* cadjwgt, int32_t* cvwgt, int32_t* cvsize,
                               int32_t* cmap, int32_t* cm, int32_t* flags);

/* initial/grow.rs:12-103 and initial/random.rs:8-10.  Return 0, or -1 when the
 * graph has fewer vertices than nparts (the reference's seed loop never ends). */
int32_t orc_grow_partition(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                           const int32_t* vwgt, int32_t nparts, const double* tpwgts,
                           orc_rng* rng, int32_t* part);

/* refine/fm.rs:12-145 (2-way FM); tpw[2*ncon] and ub[ncon] as f64; equal
 * gains pop the smallest vertex first (the crate's order is unpinned). */
int32_t orc_fm2_refine(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy, const int32_t* vwgt,
                       const int32_t* adjwgt, int32_t* part, const double* tpw, const double* ub, int32_t niter);
int32_t orc_fm2_pass(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy, const int32_t* vwgt,
                     const int32_t* adjwgt, int32_t* part, const double* tpw, const double* ub);
/* partition/recursive.rs:11-187 (part_graph_recursive with one SmallRng
 * seeded from seed through the whole recursion). */
int32_t orc_part_graph_recursive(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                                 const int32_t* vwgt, const int32_t* adjwgt, int32_t nparts, const float* tpwgts,
                                 const float* ubvec, int32_t ctype_shem, int32_t iptype, int32_t rtype_greedy,
                                 int32_t niter, uint64_t seed, int32_t ufactor, int32_t* objval, int32_t* part);
/* partition/nd.rs:17-120 with separator/mod.rs:10-61. */
int32_t orc_node_nd(int32_t n, const int32_t* xadj, const int32_t* adjncy, const int32_t* vwgt, int32_t ctype_shem,
                    int32_t iptype, int32_t rtype_greedy, int32_t niter, uint64_t seed, int32_t ufactor,
                    int32_t* perm, int32_t* iperm);

/* Timing/size counters filled by orc_part_graph_kway when non-NULL. */
typedef struct {
    double  t_coarsen_s;     /* matching + contraction, all levels, all attempts */
    double  t_initial_s;
    double  t_refine_s;      /* all greedy calls */
    int64_t coarsen_edges;   /* sum of fine-level xadj[n] over coarsen passes */
    int64_t refine_edges;    /* sum of level xadj[n] over executed greedy passes */
    int32_t nlevels;         /* levels of the last attempt */
    int32_t coarsest_n;
} orc_stats;

/* partition/kway.rs:20-141 behind the c_api.rs:140-209 argument conventions
 * (NULL = None).  Options are the already-parsed fields; rtype_greedy 0
 * refines with k-way FM (kway.rs:158-160).  Returns METIS codes
 * (1 ok, -2 input, -4 other). */
int32_t orc_part_graph_kway(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                            const int32_t* vwgt, const int32_t* vsize, const int32_t* adjwgt,
                            int32_t nparts, const float* tpwgts, const float* ubvec,
                            int32_t ctype_shem, int32_t iptype, int32_t rtype_greedy,
                            int32_t ncuts, int32_t niter, uint64_t seed, int32_t ufactor,
                            int32_t* objval, int32_t* part, orc_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
