// kernels.hpp -- host-side launchers of the gfx950 kernels.
#pragma once
#include "common.hpp"

namespace rk {

// ---- scan.hip ----
size_t scan_temp_bytes(int64_t n);
// out[i] = sum(in[0..i)); *total_dev (device) = sum(in) as int32.
void exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* total_dev, void* temp,
                        hipStream_t s);

// ---- reduce.hip ----
// Batched levels (graphs laid out one after another; gseg[v] = graph of v,
// per-graph k*ncon tables at offset gkc[g]):
// The validation of graph/validate.rs:16-142 that remains once the host has
// checked xadj[0] == 0 and xadj non-decreasing, on the uploaded graph (cnt:
// n ints of scratch; bad_host: pinned); true when the graph is valid.
bool validate_csr_device(const DevGraph& g, int32_t* cnt, int32_t* bad_dev, int32_t* bad_host, hipStream_t s);
// gmask (null: every graph): only the graphs g with gmask[g] set (of
// ngraphs) get their tables cleared and recomputed; the others keep theirs.
void part_weights_batch(const DevGraph& g, const int32_t* part, const int32_t* gseg, const int32_t* gkc,
                        const int32_t* gk, int64_t kc_total, int32_t* pw, hipStream_t s,
                        const uint8_t* gmask = nullptr, int32_t ngraphs = 0);
// cut_out[g] = edge cut of graph g (i32 wrap, halved); acc is ngraphs u64 scratch.
void edge_cut_batch(const DevGraph& g, const int32_t* part, const int32_t* gseg, int32_t ngraphs,
                    unsigned long long* acc, int32_t* cut_out, hipStream_t s);
// part[v] = packed[goff[g] + v - gv0[g]] for the graphs g = gseg[v] with goff[g] >= 0.
void scatter_parts(int32_t n, const int32_t* gseg, const int32_t* gv0, const int32_t* goff, const int32_t* packed,
                   int32_t* part, hipStream_t s);
// desc: 5 ints per listed graph {v0, n, e0, m, out}; packed[out...] = its
// xadj (n+1, rebased), adjncy (m, graph-local ids), vwgt (n*ncon).
void gather_csr(const DevGraph& g, const int32_t* desc, int32_t ndesc, int32_t* packed, hipStream_t s);
void copy_improved(int32_t n, const int32_t* gseg, const uint8_t* improved, const int32_t* part, int32_t* best,
                   hipStream_t s);
void fill_gseg(int32_t ngraphs, const int32_t* gv0, const int32_t* gn, int32_t* gseg, hipStream_t s);

// One graph of a level for grow_partitions (initial/grow.rs:12-103).
struct GrowArgs {
    const int32_t* xadj = nullptr;    // the graph's level (union CSR, vertex weights or null)
    const int32_t* adjncy = nullptr;
    const int32_t* vwgt = nullptr;
    int32_t* out = nullptr;           // [n] the grown part, graph-local
    int64_t v0 = 0, e0 = 0;           // the graph's first vertex and first entry in the level
    int32_t n = 0, m = 0, k = 0, ncon = 0;   // its vertices, entries, parts (< 2^24), constraints
    int32_t lds = 0;                  // 1: CSR and tables in LDS (grow_lds_bytes fits), the scratch below unused
    int32_t norep = 0;                // 1: no row repeats an entry (the long-row commits skip first-position checks)
    const double* tpw = nullptr;      // [k*ncon] target weights
    const int32_t* seeds = nullptr;   // [k] distinct graph-local seed vertices (grow.rs:24-32)
    int32_t* qnext = nullptr;         // [n] scratch: BFS queue links
    unsigned long long* owner = nullptr;   // [n] scratch: claim stamps
    int32_t* leftover = nullptr;      // [n] scratch: unassigned vertices in order
    int32_t* st = nullptr;            // [7k] scratch: queue heads, tails, proposals, flags, long-row lists
    long long* pw = nullptr;          // [k*ncon] scratch: running part weights
    long long* w0 = nullptr;          // [k] scratch for k > 4096 (constraint-0 weights)
};
// LDS bytes of a graph grown in the workgroup's LDS, and of a graph grown
// in global memory (its leftover weights).
size_t grow_lds_bytes(int32_t n, int64_t m, int32_t k, int32_t ncon);
size_t grow_global_lds_bytes(int32_t k);
constexpr size_t kGrowLdsBudget = 64 * 1024;   // LDS of one growing workgroup at most
// Grows the part of every listed graph (one workgroup each, lds bytes of
// dynamic LDS: the largest need among them) into its out array; the graphs
// may come from different levels.
void grow_partitions(const GrowArgs* table, int32_t ngraphs, size_t lds, hipStream_t s);
// best[v] = part[a*n + v] for a = choice[gseg[v]] >= 0 (part: a union of n
// vertices replicated once per attempt, copy a first at a*n).
void copy_attempt(int32_t n, const int32_t* gseg, const int32_t* choice, const int32_t* part, int32_t* best,
                  hipStream_t s);
// One graph's CSR (n vertices, m entries) into a union at vertex v0, entry e0.
void pack_csr(int32_t n, int64_t m, const int32_t* xadj, const int32_t* adjncy, int32_t v0, int32_t e0,
              int32_t* uxadj, int32_t* uadjncy, hipStream_t s);
// out[g*ncon + c] = int64 column sums of vwgt over graph g's vertices (g.vwgt
// required); outmax (may be null) the column maxima.
void graph_column_sums(const DevGraph& g, const int32_t* gvtx, int32_t ngraphs, long long* out, hipStream_t s,
                       long long* outmax = nullptr);
// out[4g..4g+3] = (first vertex, end vertex, first entry, end entry) of each
// graph g < ngraphs_total in the level (zeros for graphs not present).
void graph_ranges(int32_t n, const int32_t* gseg, const int32_t* xadj, int32_t ngraphs_total, int32_t* out,
                  hipStream_t s);
// Tile tables of a batched level; desc[i] = {graph, first vertex, n, first tile}.
void fill_tiles(const int4* desc, int32_t ndesc, int32_t* tile_v0, int32_t* tile_nv, int32_t* tile_g,
                hipStream_t s);

// Lanes per vertex for a level with n vertices and m directed entries.
int pick_group(int32_t n, int32_t m);
void edge_cut(const DevGraph& g, const int32_t* part, unsigned long long* acc, int32_t* cut_out, int G,
              hipStream_t s);
// zeroed: pw already holds zeros (a clear folded into an earlier kernel)
void part_weights(const DevGraph& g, const int32_t* part, int32_t k, int32_t* pw, hipStream_t s,
                  bool zeroed = false);
// part_f[v] = part_c[cmap[v]]; also clears clear_a[0..na) and clear_b[0..nb)
void project(int32_t n_fine, const int32_t* cmap, const int32_t* part_c, int32_t* part_f, hipStream_t s,
             int32_t* clear_a = nullptr, int64_t na = 0, int32_t* clear_b = nullptr, int32_t nb = 0);
// out[c] = sum over v of vwgt[v*ncon+c] as int64 (g.vwgt must be present);
// outmax (may be null) the column maxima.
void column_sums(const DevGraph& g, long long* out, hipStream_t s, long long* outmax = nullptr);

// ---- refine.hip ----
constexpr int kHeavyGlobalBlocks = 256;
constexpr int kMarkMaxDeg = 1024;   // a cut-pass mover of higher degree makes the next cut pass a full one
constexpr int kGainHeavy = 64;   // vertices of higher degree get a block each in the gain kernels

// Work buffers of the greedy rounds, sized for the finest level of a call.
struct RefineBufs {
    int4* tile_rec = nullptr;      // [ntiles * kTile]
    int32_t* tile_cnt = nullptr;   // [ntiles]
    int32_t* tile_off = nullptr;   // [ntiles] ordered ids of the tiles with movers (admission scratch)
    uint32_t* tile_bits = nullptr; // [ceil(ntiles/32)] bit per tile with movers, kept exact by the gain kernel
    uint32_t* tile_dirty = nullptr; // [ceil(ntiles/32)] tiles with a vertex or neighbour moved by the last
                                    // cut pass (set by admission, cleared by the gain kernel)
    int4* mov = nullptr;           // [n] dense candidate records, vertex order
    int32_t* pool = nullptr;       // [m] candidate parts
    int4* heavy_rec = nullptr;     // [n]
    int32_t* heavy_scratch = nullptr;  // [kHeavyGlobalBlocks * (3k+2)] when k > LDS capacity
    int2* moves = nullptr;         // [n] admitted (v, p) in admission order
    int32_t* pw = nullptr;         // [k*ncon]
    long long* fmax = nullptr;     // [k*ncon] feasibility thresholds floor(tpw*ub)
    long long* omax = nullptr;     // [k*ncon] overweight thresholds floor(tpw*ub)
    int32_t* adm_scratch = nullptr;  // [4*k*ncon] admission tables when k*ncon exceeds LDS
    int32_t* heavy_cnt = nullptr;  // device counters
    int32_t* nmov = nullptr;
    int32_t* nmoves = nullptr;
    int32_t* done = nullptr;
    int32_t* passes = nullptr;     // count of executed passes (statistics)
    int32_t* full = nullptr;       // "the next cut pass recomputes every tile" (a hub moved)
    int32_t* host_small = nullptr; // pinned host scratch for counter reads
    void* scan_tmp = nullptr;
};

// Arguments of one ordered admission pass (one graph).
struct AdmitArgs {
    int32_t mode, k, ncon, ntiles;
    const int32_t* tile_cnt;
    uint32_t* tile_bits;         // bit per tile with movers (cleared by the pass)
    int32_t* tile_list;          // [ntiles] ordered ids of those tiles
    const int4* tile_rec;
    int4* mov;
    const int32_t* pool;
    const int32_t* xadj;
    const int32_t* vwgt;
    const long long* fmax;
    const long long* omax;
    int32_t* pw;
    int32_t* part;
    int2* moves;
    int32_t* nmoves;
    int32_t* done;
    int32_t* passes;
    int32_t* gscratch;   // 3*k*ncon ints when k*ncon exceeds the LDS tables
    // dirty-tile marking for the next, incremental cut pass (null: none)
    const int32_t* adjncy;
    uint32_t* tile_dirty;
    int32_t* full;
    int32_t v0;          // first vertex of the graph (tiles count from it)
    int32_t spec_min;    // speculate on levels with at least this many vertices per part (set at launch)
};

// Heavy-vertex list (degree > kGainHeavy) of a level with n = *n_dev (or n_max
// when n_dev is null) vertices; the count lands in *count (no host sync).
void heavy_list_build(int32_t n_max, const int32_t* n_dev, const int32_t* xadj, int32_t* list, int32_t* count,
                      hipStream_t s, bool count_zeroed = false);
// heavy_list_build for a level whose size is known on the host, returning the
// count (synchronises).
int heavy_list_sync(const DevGraph& g, int32_t* list, RefineBufs& b, hipStream_t s);
// One greedy pass: mode 0 balance repair, mode 1 cut optimisation.  Reads
// b.pw (running weights) and updates it, part, b.moves/b.nmoves; a
// cut-optimisation pass with no move sets *b.done, which turns every later
// pass of the level into a no-op.  incremental: recompute only the tiles the
// previous pass marked; mark_next: mark the tiles around this pass's moves
// for an incremental next pass.
void refine_pass(const DevGraph& g, int32_t k, int mode, int nheavy, const int32_t* heavy_list, RefineBufs& b,
                 int32_t* part, hipStream_t s, bool incremental = false, bool mark_next = false);
// ---- admit.hip ----
// Ordered admission of the candidates that the gain kernels left in
// b.tile_rec / b.tile_cnt (one workgroup; updates b.pw, part, b.moves).
void admit_pass(const DevGraph& g, int32_t k, int mode, int ntiles, RefineBufs& b, int32_t* part, hipStream_t s);
// One admission pass for every graph of a batched level: workgroup i runs
// the pass described by table[i] (device memory); kmax and ncon size the
// LDS tables of all of them.
void admit_pass_batch(int mode, int ngraphs, const AdmitArgs* table, int32_t kmax, int32_t ncon, hipStream_t s);

// Device tables of one batched level (graphs laid out one after another):
// tiles never straddle graphs and each graph's tiles start at a multiple of
// 32, so that its words of tile_bits are its own.
struct BatchLevel {
    int32_t ntiles = 0;               // tiles of the level, padding included
    const int32_t* tile_v0 = nullptr; // [ntiles] first vertex (-1: padding)
    const int32_t* tile_nv = nullptr; // [ntiles] vertices in the tile
    const int32_t* tile_g = nullptr;  // [ntiles] graph (-1: padding)
    const int32_t* gseg = nullptr;    // [n] graph of each vertex
    const int32_t* gkc = nullptr;     // [ngraphs_total] offset of the graph's k*ncon tables
    const int32_t* gk = nullptr;      // [ngraphs_total] part count of each graph
    int32_t* gdone = nullptr;         // [ngraphs_total] "a cut pass of this level moved nothing"
    int32_t* gfull = nullptr;         // [ngraphs_total] "the next cut pass recomputes every tile"
    const AdmitArgs* admit = nullptr; // [ngraphs] admission arguments of the graphs present
    int32_t ngraphs = 0;              // graphs present at this level
    int32_t kmax = 0;                 // largest part count among them
};
// refine_pass over every graph of a batched level (b.pw / b.omax hold the
// per-graph tables at the offsets gkc).
void refine_pass_batch(const DevGraph& g, const BatchLevel& bl, int mode, int nheavy, const int32_t* heavy_list,
                       RefineBufs& b, int32_t* part, hipStream_t s, bool incremental = false);

// Gain table of every vertex (repair-mode candidate lists with every part
// treated as overweight): per-vertex record (v or -1, pv, ncand, nbefore |
// zero-group flag << 31), gain of the implicit zero group, and candidate
// parts / gains in admission order at xadj[v].
void gain_table(const DevGraph& g, int32_t k, int nheavy, const int32_t* heavy_list, RefineBufs& b,
                const int32_t* part, int4* vrec, int32_t* g0, int32_t* gains, hipStream_t s);

// Candidate lists of every boundary vertex for k-way FM: the gain table of
// gain_table() (pw0 all zeros and omin all INT64_MIN, so every part counts as
// overweight), without a host synchronisation; a set *done skips the launch.
void fm_candidates(const DevGraph& g, int32_t k, int nheavy, const int32_t* heavy_list, RefineBufs& b,
                   const int32_t* part, const int32_t* pw0, const long long* omin, const int32_t* done, int4* vrec,
                   int32_t* g0, int32_t* gains, hipStream_t s);

// ---- fm.hip ----
// Work buffers of the k-way FM passes of one call (sized for its finest level).
struct FmBufs {
    uint64_t* tree = nullptr;     // queue keys: n leaves, then the 64-ary max tree above them
    int32_t* tp = nullptr;        // [n] best target part of each queued vertex
    uint8_t* locked = nullptr;    // [n]
    int4* vrec = nullptr;         // [n] candidate records
    int32_t* g0 = nullptr;        // [n]
    int32_t* gains = nullptr;     // [m] candidate gains beside RefineBufs::pool
    int32_t* pw0 = nullptr;       // [k*ncon] zeros
    long long* omin = nullptr;    // [k*ncon] INT64_MIN
    int32_t* ovf = nullptr;       // [k] dense part table for recomputes when k exceeds the LDS table
    int32_t* err = nullptr;       // pop budget exhausted (never expected)
};
// One launch: pw0 = 0, omin = INT64_MIN (kc entries each).
void fm_fill_tables(int32_t* pw0, long long* omin, int64_t kc, hipStream_t s);
// fm_kway_refine (refine/fm_kway.rs:12-30) on one level: b.pw holds (or, when
// !pw_valid, receives) partition_weights(part), *b.done is cleared (or
// cleared already: cleared), niter passes run, each a no-op once a pass kept
// no move.  No host synchronisation.
void fm_refine_level(const DevGraph& g, int32_t k, const long long* d_fmax, int niter, int32_t* part, int nheavy,
                     const int32_t* heavy_list, RefineBufs& b, FmBufs& f, hipStream_t s, bool cleared,
                     bool pw_valid);

// ---- fm2.hip ----
// One graph of a batched level refined by 2-way FM: its id (index of its
// per-graph tables), first vertex, size and the start of its tree region.
struct Fm2Graph {
    int32_t g = 0, v0 = 0, n = 0, pad = 0;
    int64_t tbase = 0;
};
constexpr int kFm2MaxLvl = 12;  // tree levels, leaves included (8^11 > 2^31)
// A batched level for fm2_refine_batch.
struct Fm2Level {
    const int32_t* gseg = nullptr;      // [n] graph of each vertex (device)
    const int32_t* gdesc_of = nullptr;  // [gtotal] index into graphs of each graph id, -1 when absent (device)
    const Fm2Graph* graphs = nullptr;   // [ngraphs] the graphs present (device)
    int32_t ngraphs = 0;
    int32_t gtotal = 0;                 // graph ids of the batch
    int nlvl = 0;                       // tree levels of the largest graph present
    int32_t max_n = 0;                  // vertices of the largest graph present
    const int64_t* nodes = nullptr;     // [kFm2MaxLvl][ngraphs + 1] prefix node counts per level (device)
    int64_t level_nodes[kFm2MaxLvl] = {};   // total nodes per level (host)
};
// Work buffers of the 2-way FM passes (sized for the finest level; per-graph
// arrays for the batch's graph ids).
struct Fm2Bufs {
    uint64_t* leaf = nullptr;     // [n] queue keys
    uint64_t* tree = nullptr;     // [sum of fm2_tree_nodes] tree nodes above the leaves
    uint8_t* state = nullptr;     // [n] side, locked and queued bits
    int32_t* gain = nullptr;      // [n] gain of a move to the other side
    int32_t* ext = nullptr;       // [n] row entries on the other side
    const int32_t* sym = nullptr; // device flag from fm2_symmetry (of the call's input graph)
    int32_t* moves = nullptr;     // [n]
    int32_t* cut0 = nullptr;      // [gtotal] cut at a pass start
    int32_t* cut1 = nullptr;      // [gtotal] cut after the pops
    int32_t* best_len = nullptr;  // [gtotal]
    int32_t* best_cut = nullptr;  // [gtotal]
    int32_t* nmoves = nullptr;    // [gtotal]
    int32_t* err = nullptr;       // pop budget exhausted (never expected)
};
// *flag (device) = 1 when the graph's rows are pair-symmetric (every entry
// (v, u, w) matched by as many (u, v, w) in u's row), else 0.  Subgraphs and
// coarse graphs of a pair-symmetric graph are pair-symmetric too.
void fm2_symmetry(const DevGraph& g, int32_t* flag, hipStream_t s);
// The queue tree of a graph of n vertices (64-ary, leaves at level 0): its
// nodes above the leaves, its levels (leaves included) and its node count at
// level l (0 beyond its top).
int64_t fm2_tree_nodes(int32_t n);
int32_t fm2_tree_levels(int32_t n);
int64_t fm2_level_nodes(int32_t n, int l);
// fm_refine (refine/fm.rs:12-29) on every graph of a batched level with two
// parts: pw holds partition_weights(part) per graph at gkc[g], fmax the
// floor(target * ub) thresholds; gdone[g] is set when a pass does not
// improve (and must be clear on entry for the graphs to refine).  niter
// passes, each a no-op for the graphs done.  cutacc: gtotal u64 scratch.
void fm2_refine_batch(const DevGraph& g, const Fm2Level& lv, Fm2Bufs& f, const int32_t* gkc, const long long* fmax,
                      int32_t* pw, int32_t* part, int32_t* gdone, int niter, unsigned long long* cutacc,
                      hipStream_t s);

// ---- split.hip ----
// pre[0..n] = exclusive prefix counts of label == l (flag: n scratch ints).
void label_prefix(int32_t n, const int32_t* label, int32_t l, int32_t* flag, int32_t* pre, void* scan_tmp,
                  hipStream_t s);
// counts[g*L + l] = vertices of graph g with label l (pre: device array of L
// device pointers to the label prefixes; gv0: ngraphs + 1 first vertices).
void split_counts(int32_t ngraphs, int32_t L, const int32_t* gv0, const int32_t* const* pre, int32_t* counts,
                  hipStream_t s);
// rank[v] = v's position among the vertices of its child (g, label[v]);
// nid[v] = cbase[g*L + l] + rank[v] when cbase >= 0, else -1.
void split_rank(int32_t n, int32_t L, const int32_t* label, const int32_t* gseg, const int32_t* gv0,
                const int32_t* const* pre, const int32_t* cbase, int32_t* nid, int32_t* rank, hipStream_t s);
// The child union (nn vertices): out.xadj/adjncy/adjwgt/vwgt (allocated by
// the caller, adjwgt/vwgt null when g has none) and norig[nid[v]] = orig[v]
// (orig null: the identity); ndeg: nn + 1 scratch ints.
void split_graph(const DevGraph& g, const int32_t* orig, const int32_t* label, const int32_t* nid, int32_t nn,
                 DevGraph& out, int32_t* norig, int32_t* ndeg, void* scan_tmp, hipStream_t s);
// out[orig[v]] = pa[c] + min(rank[v], pcap[c]) for v of a child c = g*L +
// label[v] with pa[c] >= 0 (orig null: the identity).
void place_children(int32_t n, int32_t L, const int32_t* label, const int32_t* gseg, const int32_t* rank,
                    const int32_t* orig, const int32_t* pa, const int32_t* pcap, int32_t* out, hipStream_t s);
// label[v] = l for the vertices of the graphs g with flag[g] != 0.
void relabel_graphs(int32_t n, const int32_t* gseg, const uint8_t* flag, int32_t l, int32_t* label, hipStream_t s);
// iperm[perm[v]] = v.
void inverse_permutation(int32_t n, const int32_t* perm, int32_t* iperm, hipStream_t s);
// Vertex separator labels of bisected graphs (separator/mod.rs:29-58): 0 the
// lighter side, 1 the other, 2 the lighter side's vertices with a neighbour
// across; pw holds each graph's two side weights at 2*ncon*g.
void separator_labels(const DevGraph& g, const int32_t* side, const int32_t* gseg, const int32_t* pw,
                      int32_t* label, hipStream_t s);

// ---- coarsen.hip ----
struct CoarsenBufs {
    int32_t* match = nullptr;      // [n]
    int32_t* cand = nullptr;       // [n]
    int32_t* leader = nullptr;     // [n] flags, reused as deg_sum
    int32_t* cid = nullptr;        // [n] coarse id of leaders (scan)
    int32_t* members = nullptr;    // [2n] (a, b) per coarse vertex
    int32_t* degsum = nullptr;     // [n]
    int32_t* tmp_off = nullptr;    // [n]
    int32_t* tmp_adj = nullptr;    // [m]
    int32_t* tmp_w = nullptr;      // [m]
    int32_t* cdeg = nullptr;       // [n]
    int32_t* big_list = nullptr;   // [n] coarse vertices handled by the global-hash path
    int32_t* hash_keys = nullptr;  // [2m] global hash tables of the big path
    int32_t* hash_first = nullptr;
    int32_t* hash_sum = nullptr;
    int32_t* counters = nullptr;   // [16]: 0 cn, 1 cm, 2 nbig, 4 weight-not-one flag, 6-7 hash bump (u64), 8 nheavy
    int32_t* host_small = nullptr;   // pinned host scratch for the per-level counter read
    uint64_t* suitor = nullptr;      // [n] matching: suitor id and the pair's key inputs (suitor_word)
    void* scan_tmp = nullptr;
};

constexpr int kContractBig = 256;  // coarse vertices of larger degree sum use the global-hash path

// Deterministic seeded matching (ctype 1 heavy edge, 0 random): the greedy
// matching in edge-key order, computed by the suitor algorithm.
// The graphs of a batched level, laid out one after another in one CSR
// (null for a single graph).
struct GraphSegs {
    const int32_t* gseg = nullptr;     // [n] graph of each vertex
    const int32_t* gv0 = nullptr;      // [ngraphs] first vertex of each graph at this level
    const uint64_t* gseed = nullptr;   // [ngraphs] matching seed of each graph at this level
    const uint8_t* gon = nullptr;      // [ngraphs] graph is matched and contracted at this level
    int32_t* cgseg = nullptr;          // [n] out: graph of each coarse vertex
};

// cap (device, may be null; used only when g.vwgt is present): per
// constraint the largest weight a matched pair may have (per graph and
// constraint on batched levels).
void match_graph(const DevGraph& g, int ctype_shem, uint64_t seed, CoarsenBufs& b, hipStream_t s,
                 const GraphSegs* gs = nullptr, const long long* cap = nullptr);
// The cap of a graph with n vertices and total weight `total` of one
// constraint partitioned into k parts: 1.5 x total / max(20k, 20) (the
// coarsening limit), and at least twice the mean vertex weight.
inline long long match_weight_cap(long long total, long long n, int32_t k) {
    const long long limit = k * 20LL > 20 ? k * 20LL : 20;
    const long long share = 3 * total / (2 * limit);
    const long long mean2 = 2 * ((total + n - 1) / (n > 0 ? n : 1));
    return share > mean2 ? share : mean2;
}
struct CoarseAlloc {
    virtual int32_t* alloc_i32(int64_t count) = 0;
    virtual ~CoarseAlloc() = default;
};
struct ContractResult {
    int32_t cn = 0, cm = 0;
    bool adjwgt_not_one = false;   // some merged edge weight differs from 1
    int32_t nheavy = 0;            // coarse vertices of degree > kGainHeavy (when listed)
};
// Coarse graph of a symmetric matching (graph/mod.rs:284-384) into cg, with
// cmap[n]; arrays come from `al` sized by the fine graph.  When heavy_list is
// given, the coarse heavy-vertex list is built as well.  One synchronisation.
ContractResult contract(const DevGraph& g, const int32_t* match, int32_t* cmap, DevGraph& cg, bool aggregate_vsize,
                        CoarsenBufs& b, CoarseAlloc& al, int32_t* heavy_list, hipStream_t s,
                        const GraphSegs* gs = nullptr);

}  // namespace rk
