// driver.hpp -- multilevel k-way scheduler (host C++) over the gfx950 kernels.
#pragma once
#include <vector>

#include "context.hpp"
#include "kernels.hpp"
#include "rng.hpp"

namespace rk {

// Parsed options (types.rs:52-95).
struct Options {
    int ctype_shem = 1;     // CType::Shem
    int iptype = 0;         // 0 Grow, 1 Random, 4 MetisRb (grows, initial/mod.rs:17)
    int rtype_greedy = 0;   // RType::Fm is the reference default
    int objtype = 0;
    int no2hop = 0;
    int ncuts = 1;
    int niter = 10;
    uint64_t seed = 0;      // 0 = entropy
    int minconn = 0;
    int contig = 0;
    int ufactor = 30;
    int numbering = 0;
    int dbglvl = 0;
};

constexpr int kNOptions = 40;  // types.rs:170

struct ArenaAlloc : CoarseAlloc {
    Arena& a;
    explicit ArenaAlloc(Arena& ar) : a(ar) {}
    int32_t* alloc_i32(int64_t count) override { return a.get<int32_t>(count); }
};

// c_api.rs:211-222 + types.rs:190-224 (NULL or options[0] == -1 => defaults).
Options parse_options(const int32_t* options);
// types.rs:227-245 applied to Options::default() (index 0 stays -1).
void default_options_raw(int32_t* out);

// floor(target*ub) thresholds of one level (graph/mod.rs:392-410, kway.rs:44-49).
void level_thresholds(int32_t k, int32_t ncon, const float* tpwgts, const int32_t* total, const double* ub,
                      long long* fmax, long long* omax);
void thresholds_from_targets(int32_t k, int32_t ncon, const double* tpw, const double* ub, long long* fmax,
                             long long* omax);

struct LevelStat {
    int32_t n = 0, m = 0;
    float coarsen_ms = 0.f;    // matching + contraction producing the next level
    float refine_ms = 0.f;     // all greedy passes of this level
    int32_t passes = 0;        // greedy passes executed (repair + cut rounds)
};

struct KwayStats {
    std::vector<LevelStat> levels;   // finest first (last attempt)
    int32_t coarsest_n = 0;
    float initial_ms = 0.f;
    int32_t attempts = 0;
    int32_t attempts_per_call = 1;   // batched form: attempts run together in one pass
};

struct Buffers {
    RefineBufs rb;
    CoarsenBufs cb;
    unsigned long long* cut_acc = nullptr;
    int32_t* cut_out = nullptr;
    long long* wsum = nullptr;
};

// Allocates the per-call work buffers for a finest level of n vertices,
// m entries, k parts, ncon constraints (the tile bitmap is cleared on the
// context's stream).  A batch passes its tile count (padding included) and
// the total of its graphs' k*ncon tables; k is then the largest part count.
Buffers alloc_buffers(Context& cx, int32_t n, int32_t m, int32_t k, int32_t ncon, int64_t tiles = -1,
                      int64_t kc_total = -1);

// FM work buffers of a call whose finest level has n vertices and m entries
// (the constant tables are filled on the context's stream).
FmBufs alloc_fm_buffers(Context& cx, int32_t n, int32_t m, int32_t k, int32_t ncon);

// Device bytes kway_device takes from its arena before the second
// coarsening level: the work buffers, the finest heavy-vertex list and the
// first coarse level (deeper levels take more, which the arena reports).
size_t kway_first_level_bytes(int32_t n, int32_t m, int32_t k, int32_t ncon);

// Column sums of vwgt (or n per constraint when vwgt is absent), checked
// against the int32 range the reference's sums live in.
// maxw (may be null) receives the heaviest vertex weight per constraint (1
// without vertex weights).
std::vector<int32_t> total_weight_checked(Context& cx, const DevGraph& g, Buffers& B,
                                          std::vector<int32_t>* maxw = nullptr);

// One level's greedy refinement (greedy.rs:36-193) on the device partition.
// cleared: B.rb.pw and B.rb.done were already zeroed (by the projection).
// pw_valid: B.rb.pw already holds partition_weights(d_part) (the coarser
// level's final table carried over a projection with explicit vertex
// weights), so it is not recomputed.
void refine_level(Context& cx, const DevGraph& g, int32_t k, const long long* d_fmax, const long long* d_omax,
                  int niter, int32_t* d_part, int nheavy, const int32_t* heavy_list, Buffers& B, LevelStat* st,
                  bool cleared = false, bool pw_valid = false);

// Full k-way partition of a device-resident graph (partition/kway.rs:20-141);
// writes d_part and returns objval.
int32_t kway_device(Context& cx, const DevGraph& g, int32_t k, const float* tpwgts, const float* ubvec,
                    const Options& o, int32_t* d_part, KwayStats* st);

// Many graphs in one call (partition/kway.rs:20-141 per graph).  g is the
// union of the graphs laid out one after another (global vertex ids; graph i
// owns vertices [gvtx[i], gvtx[i+1]) and their adjacency rows, with no edge
// between graphs; vwgt required).  Attempt a of graph i uses the seed
// seeds[i] + a (o.seed + a when seeds is null; 0 draws entropy), exactly as
// a separate kway_device call with those options would, and the graph keeps
// its strictly best cut.  Writes d_part[g.n] and objval[ngraphs] (host).
// The attempts run batch_attempts_per_call at a time over a union holding
// that many copies of g; max_pack > 0 caps the copies (a caller with a
// memory budget of its own passes 1).
void kway_batch_device(Context& cx, const DevGraph& g, int32_t ngraphs, const int32_t* gvtx, const int32_t* nparts,
                       const float* ubvec, const Options& o, const uint64_t* seeds, int32_t* d_part, int32_t* objval,
                       KwayStats* st, int max_pack = 0);
// One multilevel bisection (partition/recursive.rs:138-187) of every graph
// of the union g: graph i (vertices gvtx[i]..gvtx[i+1], at least 3) into two
// parts with the target fractions tpw2[2*ncon*i ...] and the seed seeds[i];
// greedy or 2-way FM refinement per o.  d_part receives 0/1 per vertex.
void bisect_batch_device(Context& cx, const DevGraph& g, int32_t ngraphs, const int32_t* gvtx, const float* tpw2,
                         const float* ubvec, const Options& o, const uint64_t* seeds, int32_t* d_part);
// fm_refine (refine/fm.rs:12-29) of one graph with two parts: d_fmax holds
// floor(target * ub) per part and constraint; part is refined in place.
void fm2_refine_device(Context& cx, const DevGraph& g, const long long* d_fmax, int niter, int32_t* part);
// RMETIS_PartGraphRecursive (partition/recursive.rs:11-57) of g: part on
// the device, the edge cut returned.
int32_t recursive_device(Context& cx, const DevGraph& g, int32_t nparts, const float* tpwgts, const float* ubvec,
                         const Options& o, int32_t* d_part);
// RMETIS_NodeND (partition/nd.rs:16-44) of g (ncon 1): perm and iperm on
// the device.
void node_nd_device(Context& cx, const DevGraph& g, const Options& o, int32_t* d_perm, int32_t* d_iperm);

// Attempts per pass of kway_batch_device for a batch of n vertices, m
// entries, the given padded tile count and k*ncon total: bounded by the
// int32 ranges of the union, 60 % of the free device memory (plus the
// reusable bytes the caller's arena holds) and max_pack, then balanced over
// the fewest passes.
int32_t batch_attempts_per_call(int64_t n, int64_t m, int64_t tiles, int64_t kc_total, int32_t ngraphs, int ncuts,
                                int max_pack, size_t reusable);

// Test toggles (no counterpart in the reference): bit 0 recomputes the
// part-weight table after every projection instead of carrying the coarser
// level's table over (the two must give identical partitions).
constexpr int32_t kDebugNoPwCarry = 1;
void set_debug_flags(int32_t flags);
int32_t debug_flags();

// grow.rs:24-32: k distinct seed vertices of a graph of n, drawn from rng
// (throws Error(-4) when n < k: the reference's draw would not end).
std::vector<int32_t> grow_seeds(int32_t n, int32_t k, Rng& rng);

// One graph to grow (initial/grow.rs:12-103): its level g, first vertex v0
// and first entry e0 there, n vertices, m entries, k parts, targets
// tpw[k*ncon], its seeds, and out[n] (device, graph-local) for the parts.
struct GrowJob {
    DevGraph g;
    int64_t v0 = 0, e0 = 0;
    int32_t n = 0, m = 0, k = 0;
    std::vector<double> tpw;
    std::vector<int32_t> seeds;
    int32_t* out = nullptr;
    bool norep = true;    // no row repeats an entry (a contracted level)
};
// Grows every job's parts on the device, all in one launch (one workgroup
// per graph), with scratch from the context's arena.
void grow_on_device(Context& cx, const std::vector<GrowJob>& jobs);

}  // namespace rk
