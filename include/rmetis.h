/*
 * rmetis.h -- C ABI of the MI355X (gfx950) k-way partitioner, librmetis_hip.so.
 *
 * The first two entry points are drop-ins for the reference crate's C ABI
 * (javagg/rmetis, src/ffi/c_api.rs) with identical signatures, argument
 * meaning, option-array parsing and return codes.  The rmetis_hip_* entry
 * points expose the device-resident path and the individual hot-path
 * functions (each cites the reference function it reproduces) so that callers
 * and parity tests can drive them through plain pointers and sizes.
 *
 * Index type idx_t = int32_t (src/types.rs:4), real_t = float (src/types.rs:6).
 * Return codes (src/types.rs:164-167): METIS_OK 1, METIS_ERROR_INPUT -2,
 * METIS_ERROR_MEMORY -3, METIS_ERROR -4.
 */
#ifndef RMETIS_HIP_H
#define RMETIS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t idx_t;
typedef float real_t;

#define RMETIS_NOPTIONS 40
#define METIS_OK 1
#define METIS_ERROR_INPUT (-2)
#define METIS_ERROR_MEMORY (-3)
#define METIS_ERROR (-4)

/* Option-array indices (src/types.rs:173-186). */
#define METIS_OPTION_PTYPE 0
#define METIS_OPTION_OBJTYPE 1
#define METIS_OPTION_CTYPE 2
#define METIS_OPTION_IPTYPE 3
#define METIS_OPTION_RTYPE 4
#define METIS_OPTION_DBGLVL 5
#define METIS_OPTION_NITER 6
#define METIS_OPTION_NCUTS 7
#define METIS_OPTION_SEED 8
#define METIS_OPTION_NO2HOP 9
#define METIS_OPTION_MINCONN 10
#define METIS_OPTION_CONTIG 11
#define METIS_OPTION_UFACTOR 16
#define METIS_OPTION_NUMBERING 17

/* Replaces RMETIS_SetDefaultOptions, src/ffi/c_api.rs:14-23.  Writes the
 * defaults of src/types.rs:77-95 via to_raw (types.rs:227-245); index 0 stays
 * -1, so an array filled here parses as "all defaults" until options[0] is
 * set.  Returns METIS_OK, or METIS_ERROR_INPUT for NULL. */
int32_t RMETIS_SetDefaultOptions(idx_t* options);

/* Replaces RMETIS_PartGraphKway, src/ffi/c_api.rs:55-81 (+ c_partition_impl
 * 140-209).  Host pointers.  Required: nvtxs, ncon, xadj, adjncy, nparts,
 * part; optional (NULL = None): vwgt[n*ncon], vsize[n], adjwgt[xadj[n]],
 * tpwgts[nparts*ncon], ubvec[ncon], options[40], objval.  Inputs are copied;
 * part/objval are written only on success.  The multilevel path runs on the
 * current HIP device; RTYPE Greedy (1) refines with the greedy passes of
 * refine/greedy.rs, anything else (the default, Fm) with the k-way FM of
 * refine/fm_kway.rs (kway.rs:154-161). */
int32_t RMETIS_PartGraphKway(const idx_t* nvtxs, const idx_t* ncon, const idx_t* xadj, const idx_t* adjncy,
                             const idx_t* vwgt, const idx_t* vsize, const idx_t* adjwgt, const idx_t* nparts,
                             const real_t* tpwgts, const real_t* ubvec, const idx_t* options, idx_t* objval,
                             idx_t* part);

/* Replaces RMETIS_PartGraphRecursive, src/ffi/c_api.rs:25-48 (+
 * c_partition_impl 140-209 with kway = false, i.e.
 * partition/recursive.rs:11-57).  Same arguments, copies, checks (nparts
 * outside 2..nvtxs: METIS_ERROR_INPUT), option parsing and return codes as
 * RMETIS_PartGraphKway.  Multilevel recursive bisection on the current HIP
 * device: each depth's subgraphs are bisected together (RTYPE Greedy: greedy
 * passes; otherwise the 2-way FM of refine/fm.rs); objval is the edge cut. */
int32_t RMETIS_PartGraphRecursive(const idx_t* nvtxs, const idx_t* ncon, const idx_t* xadj, const idx_t* adjncy,
                                  const idx_t* vwgt, const idx_t* vsize, const idx_t* adjwgt, const idx_t* nparts,
                                  const real_t* tpwgts, const real_t* ubvec, const idx_t* options, idx_t* objval,
                                  idx_t* part);

/* Replaces RMETIS_NodeND, src/ffi/c_api.rs:83-134 (partition/nd.rs:16-44,
 * separator/mod.rs:10-61).  Required: nvtxs, xadj, adjncy, perm, iperm
 * (NULL or nvtxs == 0: METIS_ERROR_INPUT); vwgt[n] and options optional.  The
 * graph is validated as Graph::new(n, 1, ...) does.  perm[v] = the position
 * of vertex v in the fill-reducing order, iperm its inverse. */
int32_t RMETIS_NodeND(const idx_t* nvtxs, const idx_t* xadj, const idx_t* adjncy, const idx_t* vwgt,
                      const idx_t* options, idx_t* perm, idx_t* iperm);

/* RMETIS_PartGraphRecursive / RMETIS_NodeND on a graph already in device
 * memory (read in place; d_part / d_perm / d_iperm are device buffers of
 * nvtxs entries).  No reference counterpart: the same computations without
 * the copies, for callers whose CSR lives in HBM. */
int32_t rmetis_hip_recursive_device(idx_t nvtxs, idx_t ncon, const idx_t* d_xadj, const idx_t* d_adjncy,
                                    const idx_t* d_vwgt, const idx_t* d_adjwgt, idx_t nparts, const real_t* tpwgts,
                                    const real_t* ubvec, const idx_t* options, idx_t* objval, idx_t* d_part);
int32_t rmetis_hip_nodend_device(idx_t nvtxs, const idx_t* d_xadj, const idx_t* d_adjncy, const idx_t* d_vwgt,
                                 const idx_t* options, idx_t* d_perm, idx_t* d_iperm);

/* METIS .graph files (README.md:208-218, DEVELOPMENT_GUIDE.md:846-874):
 * header "nvtxs nedges [fmt] [ncon]", one line per vertex with its size
 * (fmt 100), ncon weights (fmt 010) and 1-based neighbours each followed by
 * its edge weight (fmt 001); '%' starts a comment line.  The reference names
 * the format but ships no reader.  rmetis_read_graph fills 0-based CSR
 * arrays allocated with malloc (NULL for absent weights / sizes; free them
 * with rmetis_free) and returns METIS_OK or METIS_ERROR_INPUT for a missing
 * or malformed file (message in rmetis_hip_last_error()).
 * rmetis_write_graph writes the same format (NULL weights are left out). */
int32_t rmetis_read_graph(const char* path, idx_t* nvtxs, idx_t* ncon, idx_t** xadj, idx_t** adjncy, idx_t** vwgt,
                          idx_t** vsize, idx_t** adjwgt);
int32_t rmetis_write_graph(const char* path, idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                           const idx_t* vwgt, const idx_t* vsize, const idx_t* adjwgt);
void rmetis_free(void* p);

/* The tuning table: launch shapes and thresholds measured by the autotuner
 * (rmetis_amd/autotune.py) and read on first use from tuning.json beside
 * the library (or the file named by the environment variable
 * RMETIS_HIP_TUNING); compiled-in defaults otherwise.  No reference
 * counterpart: every entry is exact (results never depend on it).
 * values[0..24) suitor matching shape per level bucket, size-major: size
 *               bucket (n in [0,2^19), [2^19,2^21), [2^21,2^23), [2^23,inf))
 *               times 6 plus mean-degree bucket (m/n in [0,6), [6,10),
 *               [10,16), [16,24), [24,48), [48,inf)); shapes 0 = 1 lane x 4
 *               neighbours per round, 1 = 2 x 4, 2 = 4 x 2, 3 = 8 x 1;
 * values[24]    single-graph levels of at least this many vertices run
 *               incremental cut passes;
 * values[25]    admission speculates on levels with at least this many
 *               vertices per part.
 * get returns the number of entries written (at most 26); set installs the
 * first count entries for the whole process, or returns METIS_ERROR_INPUT
 * for an out-of-range value. */
int32_t rmetis_hip_get_tuning(int32_t* values, int32_t count);
int32_t rmetis_hip_set_tuning(const int32_t* values, int32_t count);
const char* rmetis_hip_tuning_source(void);

/* Per-level figures of the last attempt of a k-way call (finest first; for
 * a batched call, of its last pass over the union of the attempts it ran
 * together). */
#define RMETIS_HIP_MAX_LEVELS 64
typedef struct {
    int32_t nlevels;
    int32_t coarsest_n;
    float initial_ms;
    int32_t level_n[RMETIS_HIP_MAX_LEVELS];
    int32_t level_m[RMETIS_HIP_MAX_LEVELS];
    float coarsen_ms[RMETIS_HIP_MAX_LEVELS];
    float refine_ms[RMETIS_HIP_MAX_LEVELS];
    int32_t passes[RMETIS_HIP_MAX_LEVELS];
    int32_t attempts_per_call;   /* batched calls: attempts run together in one pass (levels: the last pass's) */
} rmetis_hip_stats;

/* Same computation as RMETIS_PartGraphKway (src/partition/kway.rs:20-141) on
 * a graph already resident in device memory: d_xadj, d_adjncy, d_vwgt,
 * d_adjwgt and the output d_part are device pointers (vwgt/adjwgt may be
 * NULL); tpwgts, ubvec, options, objval and stats are host pointers (stats
 * may be NULL).  The CSR is not validated (the caller provides a graph that
 * src/graph/validate.rs:16-142 accepts). */
int32_t rmetis_hip_kway_device(idx_t nvtxs, idx_t ncon, const idx_t* d_xadj, const idx_t* d_adjncy,
                               const idx_t* d_vwgt, const idx_t* d_adjwgt, idx_t nparts, const real_t* tpwgts,
                               const real_t* ubvec, const idx_t* options, idx_t* objval, idx_t* d_part,
                               rmetis_hip_stats* stats);

/* No counterpart in the reference: rmetis_hip_kway_device for callers that
 * cannot allocate (code that forbids hipMalloc, such as tight serving loops
 * and harnesses that pre-plan all device memory).  Same arguments, same
 * result, but no device or pinned memory is allocated and no stream is
 * created: every device buffer of the call comes from a fixed workspace of
 * 128 MiB per GPU (a static device array of the library), the kernels run
 * on the null stream and host staging uses heap memory.  A graph whose work
 * buffers and first coarse level do not fit is rejected with
 * METIS_ERROR_MEMORY before any kernel runs (part and objval unwritten); a
 * hierarchy that outgrows the workspace at a deeper level also ends with
 * METIS_ERROR_MEMORY, after the call's queued work has finished (objval
 * unwritten; d_part may then hold an earlier attempt's partition when
 * NCUTS > 1).  Calls on one GPU share the workspace and are serialised by a
 * lock; the workspace keeps the last call's contents. */
int32_t rmetis_hip_kway_device_noalloc(idx_t nvtxs, idx_t ncon, const idx_t* d_xadj, const idx_t* d_adjncy,
                                       const idx_t* d_vwgt, const idx_t* d_adjwgt, idx_t nparts,
                                       const real_t* tpwgts, const real_t* ubvec, const idx_t* options,
                                       idx_t* objval, idx_t* d_part);

/* Many graphs in one call (no single counterpart in the reference: each
 * graph gets what RMETIS_PartGraphKway / partition_kway, kway.rs:20-141,
 * returns for it alone).  The graphs are laid out one after another in one
 * CSR with global vertex ids: graph i owns vertices [gvtx[i], gvtx[i+1]) and
 * their rows, and its edges stay inside that range.  gvtx[ngraphs+1],
 * nparts[ngraphs], seeds[ngraphs] and objval[ngraphs] are host arrays;
 * d_xadj[n+1], d_adjncy, d_vwgt[n*ncon], d_adjwgt and d_part[n] are device
 * pointers, n = gvtx[ngraphs]; d_vwgt and d_adjwgt may be NULL (for every
 * graph: absent vertex weights are not the same as ones, since the coarse
 * vertices of an unweighted graph keep weight 1, graph/mod.rs:300-314).  options are shared
 * (RTYPE must be Greedy, so options[0] must not be -1; NCUTS attempts per graph); attempt a of graph i uses
 * seed seeds[i] + a (options' SEED + a when seeds is NULL; a zero seed draws
 * entropy) and graph i keeps its strictly best cut, so every graph's part
 * and objval equal those of a separate call with that seed.  ubvec[ncon]
 * (NULL = 1 + ufactor/1000) applies to every graph; target weights are
 * uniform.  The CSR is not validated.  The attempts run several at a time:
 * one pass over a union holding one copy of the batch per attempt (as many
 * as the int32 offsets and the free device memory allow, balanced over
 * the fewest passes; stats->attempts_per_call).  stats (may be NULL)
 * describes the union levels of the last pass. */
int32_t rmetis_hip_kway_batch_device(idx_t ngraphs, const idx_t* gvtx, idx_t ncon, const idx_t* d_xadj,
                                     const idx_t* d_adjncy, const idx_t* d_vwgt, const idx_t* d_adjwgt,
                                     const idx_t* nparts, const real_t* ubvec, const idx_t* options,
                                     const uint64_t* seeds, idx_t* objval, idx_t* d_part, rmetis_hip_stats* stats);

/* rmetis_hip_kway_batch_device on host arrays, validated per graph as
 * src/graph/validate.rs:16-142 validates one graph (rows inside their graph,
 * no self-loops, symmetric entry counts, positive vertex weights). */
int32_t rmetis_hip_kway_batch(idx_t ngraphs, const idx_t* gvtx, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                              const idx_t* vwgt, const idx_t* adjwgt, const idx_t* nparts, const real_t* ubvec,
                              const idx_t* options, const uint64_t* seeds, idx_t* objval, idx_t* part);

/* Serving (no counterpart in the reference): a request loop with continuous
 * batching.  Each request is one graph with its own nparts and seed; a
 * server runs up to `slots` (1..8) batched calls at once (one HIP stream each),
 * every idle slot taking the oldest pending request plus the pending
 * requests that can share its batch (same ncon, ubvec and options but the
 * seed) within the device-memory budget (<= 0: 80 % of the free memory at
 * creation) and max_batch_vertices (<= 0: 2^25).  Device memory is taken in
 * segments of at least 1 GiB as batches need it, up to the budget, and kept
 * until the server is destroyed.  A request's result equals that of
 * RMETIS_PartGraphKway with its seed, whatever it was batched with.
 * Returns NULL on error (rmetis_hip_last_error()). */
typedef struct rmetis_hip_server rmetis_hip_server;
rmetis_hip_server* rmetis_hip_server_create(int32_t slots, int64_t memory_budget, int64_t max_batch_vertices);

/* Queues one request; host arrays are validated (src/graph/validate.rs:16-142,
 * nparts in [2, nvtxs]) and copied.  seed 0 means options' SEED.  Returns the
 * request id (> 0) or a METIS error code (< 0). */
int64_t rmetis_hip_server_submit(rmetis_hip_server* server, idx_t nvtxs, idx_t ncon, const idx_t* xadj,
                                 const idx_t* adjncy, const idx_t* vwgt, const idx_t* adjwgt, idx_t nparts,
                                 const real_t* ubvec, const idx_t* options, uint64_t seed);

/* rmetis_hip_server_submit for a graph already in device memory: d_xadj,
 * d_adjncy, d_vwgt, d_adjwgt (the weights may be NULL) are read in place and
 * must stay valid until the result is collected, which also means the part
 * has been written to d_part[nvtxs] (device).  Not validated, like
 * rmetis_hip_kway_device. */
int64_t rmetis_hip_server_submit_device(rmetis_hip_server* server, idx_t nvtxs, idx_t ncon, const idx_t* d_xadj,
                                        const idx_t* d_adjncy, const idx_t* d_vwgt, const idx_t* d_adjwgt,
                                        idx_t nparts, const real_t* ubvec, const idx_t* options, uint64_t seed,
                                        idx_t* d_part);

/* Result of request id: METIS_OK with part[nvtxs] (host requests; may be
 * NULL, and is not written for device requests) and objval written (and
 * the request forgotten), 0 while it is pending, or its METIS error code.
 * wait = 0 returns at once, 1 waits until the request is done, and a larger
 * value waits at most that many milliseconds, then returns METIS_ERROR with
 * every slot's phase in rmetis_hip_last_error() (the request stays queued).
 * times_ms[3] (may be NULL) = queued, running, total milliseconds;
 * *batch_graphs = graphs of the batched call that ran it. */
int32_t rmetis_hip_server_result(rmetis_hip_server* server, int64_t id, int32_t wait, idx_t* part, idx_t* objval,
                                 double* times_ms, int32_t* batch_graphs);

/* out[5] = batched calls run, requests served, batches in flight, most
 * batches in flight at once, most graphs in one batch. */
int32_t rmetis_hip_server_stats(rmetis_hip_server* server, int64_t* out);

/* Stops the slots after their current batch; pending requests fail. */
void rmetis_hip_server_destroy(rmetis_hip_server* server);

/* Graph::edge_cut, src/graph/mod.rs:123-137 (host pointers; adjwgt may be NULL). */
int32_t rmetis_hip_edge_cut(idx_t nvtxs, const idx_t* xadj, const idx_t* adjncy, const idx_t* adjwgt,
                            const idx_t* part, idx_t* cut);

/* Graph::partition_weights, src/graph/mod.rs:140-149; pw[nparts*ncon]. */
int32_t rmetis_hip_partition_weights(idx_t nvtxs, idx_t ncon, const idx_t* vwgt, const idx_t* part,
                                     idx_t nparts, idx_t* pw);

/* Boundary set (Graph::boundary_vertices, src/graph/mod.rs:164-171) and gain
 * table (compute_move_gain, src/refine/greedy.rs:197-218) of a partition.
 * rec[4*v] = (v if boundary else -1, part[v], ncand, nbefore | zflag<<31);
 * g0[v] = gain of moving v to any part it has no edge weight to;
 * cand_part / cand_gain [xadj[n]] hold, at offset xadj[v], the ncand parts
 * with a nonzero edge-weight sum, ordered by gain descending then part. */
int32_t rmetis_hip_gain_table(idx_t nvtxs, const idx_t* xadj, const idx_t* adjncy, const idx_t* adjwgt,
                              const idx_t* part, idx_t nparts, idx_t* rec, idx_t* g0, idx_t* cand_part,
                              idx_t* cand_gain);

/* One synchronous greedy pass of src/refine/greedy.rs:36-193 (nprocs = 1):
 * mode 0 = balance repair (60-123), 1 = one cut-optimisation round (127-182).
 * tpwgts[nparts*ncon] are the f64 target weights and ubvec[ncon] the f64
 * imbalance factors the reference's greedy_refine receives.  pw (in/out) is
 * the running part-weight vector, part (in/out) is updated with the moves;
 * moves[2*nmoves] receives (v, p) in admission order. */
int32_t rmetis_hip_greedy_pass(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                               const idx_t* vwgt, const idx_t* adjwgt, idx_t* part, idx_t nparts,
                               const double* tpwgts, const double* ubvec, idx_t mode, idx_t* pw, idx_t* moves,
                               idx_t* nmoves);

/* greedy_refine, src/refine/greedy.rs:12-22 / 36-193: repair pass plus up to
 * niter cut rounds on part (in/out); *cut receives the final edge cut. */
int32_t rmetis_hip_greedy_refine(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                                 const idx_t* vwgt, const idx_t* adjwgt, idx_t* part, idx_t nparts,
                                 const double* tpwgts, const double* ubvec, idx_t niter, idx_t* cut);

/* fm_kway_refine, src/refine/fm_kway.rs:12-30: up to niter FM passes on part
 * (in/out), stopping after a pass that keeps no move; tpwgts[nparts*ncon]
 * f64 target weights and ubvec[ncon] f64 imbalance factors as the
 * reference's function receives them; *cut = the final edge cut.  The queue
 * pops equal gains smallest vertex first (the reference's order among equal
 * gains is internal to its priority-queue crate). */
int32_t rmetis_hip_fm_refine(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy, const idx_t* vwgt,
                             const idx_t* adjwgt, idx_t* part, idx_t nparts, const double* tpwgts,
                             const double* ubvec, idx_t niter, idx_t* cut);

/* fm_refine, src/refine/fm.rs:12-29: up to niter 2-way FM passes on part
 * (0/1 per vertex, in/out), stopping after a pass that does not improve;
 * tpwgts[2*ncon] f64 target weights and ubvec[ncon] f64 factors as the
 * reference's function receives them; *cut = the final edge cut.  The
 * queue pops equal gains smallest vertex first (the reference's order among
 * equal gains is internal to its priority-queue crate). */
int32_t rmetis_hip_fm2_refine(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy, const idx_t* vwgt,
                              const idx_t* adjwgt, idx_t* part, const double* tpwgts, const double* ubvec, idx_t niter,
                              idx_t* cut);

/* The GPU matching used in place of shem_matching / random_matching
 * (src/coarsen/shem.rs:13-56, src/coarsen/rm.rs:13-43): seeded, deterministic,
 * symmetric (matching[matching[v]] == v); ctype 1 = heavy edge, 0 = random. */
int32_t rmetis_hip_match(idx_t nvtxs, const idx_t* xadj, const idx_t* adjncy, const idx_t* adjwgt, idx_t ctype,
                         uint64_t seed, idx_t* matching);

/* The same matching with a weight cap on the matched pairs, as the k-way
 * driver runs it on weighted levels: a pair (u, v) is matched only when
 * vwgt[u][c] + vwgt[v][c] <= cap[c] for every constraint c (cap[ncon]; each
 * 0 <= cap[c] < INT64_MAX).  The result is the greedy matching of the same
 * edge order restricted to the pairs that fit.  The reference's matching has
 * no cap (src/coarsen/shem.rs:13-56): this entry exists to test the GPU's. */
int32_t rmetis_hip_match_capped(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy, const idx_t* vwgt,
                                const idx_t* adjwgt, idx_t ctype, uint64_t seed, const int64_t* cap,
                                idx_t* matching);

/* Graph::build_coarse_graph, src/graph/mod.rs:284-384, for a symmetric
 * matching.  Output capacities: cxadj n+1, cadjncy/cadjwgt xadj[n],
 * cvwgt n*ncon, cvsize n, cmap n.  *cn / *cm receive the coarse sizes;
 * *flags bit0 = cvwgt present, bit1 = cvsize present, bit2 = coarse adjwgt
 * present (the reference's Some; cadjwgt is always filled). */
int32_t rmetis_hip_build_coarse_graph(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                                      const idx_t* vwgt, const idx_t* adjwgt, const idx_t* vsize,
                                      const idx_t* matching, idx_t* cxadj, idx_t* cadjncy, idx_t* cadjwgt,
                                      idx_t* cvwgt, idx_t* cvsize, idx_t* cmap, idx_t* cn, idx_t* cm,
                                      idx_t* flags);

/* validate_csr, src/graph/validate.rs:16-142 (what Graph::new checks, 38-50),
 * with the array lengths given explicitly (vwgt/adjwgt may be NULL = None).
 * Host only; returns METIS_OK or METIS_ERROR_INPUT. */
int32_t rmetis_hip_validate_csr(idx_t nvtxs, idx_t ncon, const idx_t* xadj, int64_t xadj_len,
                                const idx_t* adjncy, int64_t adjncy_len, const idx_t* vwgt, int64_t vwgt_len,
                                const idx_t* adjwgt, int64_t adjwgt_len);

/* Options parsing exactly as RMETIS_PartGraphKway applies it (parse_options,
 * src/ffi/c_api.rs:211-222, Options::from_raw, src/types.rs:190-224).  Host
 * only.  out[13] = ctype (0 Rm, 1 Shem), iptype, rtype (0 Fm, 1 Greedy),
 * objtype, no2hop, ncuts, niter, seed, minconn, contig, ufactor, numbering,
 * dbglvl. */
int32_t rmetis_hip_parse_options(const idx_t* options, int64_t* out);

/* grow_partition, src/initial/grow.rs:12-103, with the attempt RNG
 * SmallRng::seed_from_u64(seed) (src/partition/kway.rs:62): the BFS growing
 * the k-way path runs on its coarsest level.  tpwgts[nparts*ncon] are the
 * f64 target weights (already scaled to the total, as kway.rs:67-69 passes
 * them).  Host arrays; the seeds are drawn on the host and the growing runs
 * on the device (the kernel the k-way path uses); returns METIS_OK or
 * METIS_ERROR_INPUT (also for nparts > nvtxs, where the reference's seed loop
 * would not end). */
int32_t rmetis_hip_grow_partition(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                                  const idx_t* vwgt, idx_t nparts, const double* tpwgts, uint64_t seed,
                                  idx_t* part);

/* The integer thresholds the greedy kernels compare with: for target
 * t = (tpwgts[p*ncon+c] as f64, or 1/nparts) * total[c] (Graph::
 * compute_target_weights, src/graph/mod.rs:392-410) and ub = ubvec[c] as f64
 * or 1 + ufactor/1000 (src/partition/kway.rs:44-49), fmax = floor(t*ub) with
 * w <= fmax <=> (w as f64 <= t*ub) and omax with w > omax <=> (w as f64 >
 * t*ub) (src/refine/greedy.rs:70,87,144).  Host only. */
int32_t rmetis_hip_level_thresholds(idx_t nparts, idx_t ncon, const real_t* tpwgts, const idx_t* total,
                                    const real_t* ubvec, idx_t ufactor, int64_t* fmax, int64_t* omax);

/* Test toggles (no counterpart in the reference): bit 0 makes the k-way path
 * recompute the part-weight table after each projection instead of carrying
 * the coarser level's table over; results must not change. */
void rmetis_hip_debug_flags(int32_t flags);

/* Device-memory plumbing for callers that hold no other HIP runtime handle
 * (no counterpart in the reference): plain hipMalloc/hipFree/hipMemcpy on the
 * current device; kind 1 = host->device, 2 = device->host, 3 = device->device. */
void* rmetis_hip_device_alloc(int64_t bytes);
void rmetis_hip_device_free(void* p);
int32_t rmetis_hip_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind);
int32_t rmetis_hip_device_synchronize(void);

/* Kernel timing (no counterpart in the reference; used by the benchmark).
 * While enabled, launches of each kernel family are bracketed by HIP events
 * on their own stream.  Tags: 0 gain tables/candidates (launches over every
 * tile), 1 candidate compaction, 2 ordered admission, 3 matching,
 * 4 contraction, 5 other, 6 incremental gain launches (only the tiles near
 * the previous pass's moves; bytes not modelled).
 * enable(on) clears the counters; read returns launches, total ms and the
 * algorithmic HBM bytes of the timed launches (0 where not modelled), over
 * all levels (size < 0) or the levels with exactly `size` vertices.
 * filter(tag, size) restricts the bracketing to launches of one tag (and,
 * with size >= 0, to levels of that many vertices), so that the events of
 * the other launches stay out of the stream; tag < 0 brackets every launch
 * (the default).  Each bracket adds two event records to the stream. */
void rmetis_hip_profile_enable(int32_t on);
void rmetis_hip_profile_filter(int32_t tag, int64_t size);
int32_t rmetis_hip_profile_read(int32_t tag, int64_t size, int64_t* launches, double* ms, double* bytes);

/* Message of the last failed call on this thread ("" if none). */
const char* rmetis_hip_last_error(void);

/* Number of visible HIP devices (0 when none; never throws). */
int32_t rmetis_hip_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
