// coarsen.hip -- matching and coarse-graph contraction on the GPU.
//
// Matching (coarsen/shem.rs:13-56, coarsen/rm.rs:13-43).  The reference's
// matchings are sequential greedy scans.  On the GPU the matching is the
// greedy matching in descending order of a strict total order on edges
// (EdgeKey below, ties broken by the endpoint ids), computed with the
// suitor algorithm (Manne and Halappanavar, IPDPS 2014): every vertex tries
// to become the suitor of its best neighbour whose current suitor it beats,
// a displaced suitor searches again, and mutual suitors are matched.  The
// result does not depend on the order of the attempts, so it is symmetric,
// deterministic for a given seed and independent of scheduling; it needs no
// rounds and one launch per level.
//
// Contraction (graph/mod.rs:284-384), bit-exact for a given symmetric matching:
//  * coarse ids ascend with the pair's smaller vertex (288-297): exclusive
//    scan of leader flags match[v] >= v;
//  * coarse vertex weights are summed, and stay absent when the fine graph
//    has none (300-314);
//  * coarse adjacency of {a < b}: first occurrence over adj(a) then adj(b)
//    in CSR order, internal edges dropped, weights of merged entries summed
//    (323-341).  Coarse vertices with degree sum <= kContractBig are merged
//    in per-wave LDS hash tables (k_contract_tile); larger ones use a block per vertex and
//    a hash table in global memory that records each coarse neighbour's
//    first position and weight sum (k_contract_big).
#include "common.hpp"
#include "kernels.hpp"
#include "tuning.hpp"

namespace rk {

namespace {

__device__ inline uint32_t mix32(uint32_t x) {
    // "lowbias32" integer finaliser (C. Wellons, hash-prospector).
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Pairs whose weight would exceed a cap (1.5 x a part's share of the
// coarsest level's vertices, see match_weight_cap) are not matched: without
// it a component can shrink to a few vertices that each outweigh a part
// while isolated vertices, which never match, keep the level above the
// coarsening limit; growing a part from such a vertex takes the whole
// component and no boundary is left to repair the balance.
//
// Strict total order on edges, symmetric in (v, u).  Heavy-edge matching:
// (smaller rank of the endpoints' smaller degree first, heavier edge, seeded
// hash) -- an edge is ranked by roughly when the reference's degree-ordered
// scan (shem.rs:23) would first reach it and by the weight its visiting
// endpoint maximises (shem.rs:36-44).  Degrees up to 4 rank exactly (grid
// corners, sides and interiors differ), larger ones by binary magnitude:
// exact large degrees chain the greedy matching's decisions along long paths
// of slowly rising degree (the matching then takes many dependent steps),
// while magnitudes keep almost all of its size and let weights decide more
// often.  Random matching: the seeded hash alone.
struct EdgeKey {
    uint64_t hi;
    uint32_t lo;
    __device__ bool gt(const EdgeKey& o) const { return hi > o.hi || (hi == o.hi && lo > o.lo); }
    __device__ bool eq(const EdgeKey& o) const { return hi == o.hi && lo == o.lo; }
};

// Rank of a degree d >= 1 in the edge order: d itself up to 4, then one
// rank per binary magnitude (5-7 share 4, 8-15 are 5, 16-31 are 6, ...).
__device__ inline uint32_t degree_rank(int32_t d) {
    d = max(d, 1);
    return (uint32_t)(min(d, 4) + max(0, 29 - (int)__clz(d)));
}

// Key of edge (v, u) of weight w whose endpoints' smaller degree rank is
// mrank.
__device__ inline EdgeKey edge_key_r(int32_t v, int32_t u, int32_t w, uint32_t mrank, uint64_t seed, int shem) {
    const uint32_t lo = (uint32_t)min(v, u), hi = (uint32_t)max(v, u);
    const uint32_t h = mix32(lo ^ mix32(hi ^ (uint32_t)seed) ^ (uint32_t)(seed >> 32));
    if (!shem) return EdgeKey{0ull, h};
    const uint32_t dpri = 0xFFFFFFFFu - mrank;
    return EdgeKey{((uint64_t)dpri << 32) | (uint64_t)((uint32_t)w ^ 0x80000000u), h};
}

// Strict total order on edges: the key, then the endpoint ids (graph-local,
// smaller pair first).  For the edges of one vertex this prefers, among
// equal keys, the smaller neighbour.
__device__ inline bool edge_before(const EdgeKey& k1, int32_t a1, int32_t b1, const EdgeKey& k2, int32_t a2,
                                   int32_t b2) {
    if (k1.gt(k2)) return true;
    if (!k1.eq(k2)) return false;
    const int32_t l1 = min(a1, b1), l2 = min(a2, b2);
    return l1 < l2 || (l1 == l2 && max(a1, b1) < max(a2, b2));
}

// suitor[u] is u's word: who courts u, and enough of that edge's key to
// compare against it without reloading the suitor's degree, plus u's own
// degree rank so that a vertex scanning u needs no access to u's row.
// Compact form (edge weight 0 <= w < 2^16): bit 31 of the high word set,
// the suitor id below it (all ones: nobody), and the low word
//   u's rank << 22 | rank of the pair's smaller degree << 16 | w.
// Wide form (w outside that range; bit 31 of the high word clear): the
// suitor id and w, the degrees then come from xadj.  The word of a pair is a
// function of the pair alone.
constexpr uint64_t kNoSuitor = 0xFFFFFFFF00000000ull;

__device__ inline uint64_t no_suitor_word(uint32_t urank) { return kNoSuitor | ((uint64_t)urank << 22); }

__device__ inline uint64_t suitor_word(int32_t s, int32_t w, uint32_t mrank, uint32_t urank) {
    if ((uint32_t)w < 0x10000u)
        return ((uint64_t)((uint32_t)s | 0x80000000u) << 32) | ((uint64_t)urank << 22) | ((uint64_t)mrank << 16) |
               (uint32_t)w;
    return ((uint64_t)(uint32_t)s << 32) | (uint32_t)w;
}

__device__ inline bool word_compact(uint64_t word) { return (word >> 63) != 0; }

__device__ inline int32_t suitor_id(uint64_t word) {
    const uint32_t hi = (uint32_t)(word >> 32);
    return hi == 0xFFFFFFFFu ? -1 : (int32_t)(hi & 0x7FFFFFFFu);
}

// Degree rank of u, the owner of word (from the word, or its row when wide).
__device__ inline uint32_t owner_rank(uint64_t word, int32_t u, const int32_t* xadj) {
    if (word_compact(word)) return (uint32_t)(word >> 22) & 63u;
    return degree_rank(xadj[u + 1] - xadj[u]);
}

// Key of the edge (s, u) recorded by word (s = suitor_id(word) >= 0; urank =
// u's degree rank); ids in the keys are graph-local (minus base).
__device__ inline EdgeKey suitor_key(uint64_t word, int32_t s, int32_t u, uint32_t urank, const int32_t* xadj,
                                     uint64_t seed, int shem, int32_t base) {
    const uint32_t lo = (uint32_t)word;
    if (word_compact(word)) return edge_key_r(s - base, u - base, (int32_t)(lo & 0xFFFFu), (lo >> 16) & 63u, seed, shem);
    const uint32_t rs = shem ? degree_rank(xadj[s + 1] - xadj[s]) : 0u;
    return edge_key_r(s - base, u - base, (int32_t)lo, min(rs, urank), seed, shem);
}

struct MatchArgs {
    int32_t n, shem;
    uint64_t seed;
    const int32_t* xadj;
    const int32_t* adjncy;
    const int32_t* adjwgt;
    int32_t* match;
    uint64_t* suitor;
    // weight cap (null: none): a pair is matched only when, for every
    // constraint c, vwgt[u][c] + vwgt[v][c] <= cap[c] (cap[g * ncon + c] for
    // graph g of a batched level)
    const int32_t* vwgt;
    const long long* cap;
    int32_t ncon;
    // batched levels (null for a single graph): graph of each vertex, and per
    // graph its first vertex, matching seed and whether it is matched here
    const int32_t* gseg;
    const int32_t* gv0;
    const uint64_t* gseed;
    const uint8_t* gon;
};

__device__ inline uint64_t ld_agent64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline bool pair_fits(const MatchArgs& a, int32_t x, int32_t y, const long long* cap) {
    for (int c = 0; c < a.ncon; ++c)
        if ((long long)a.vwgt[(int64_t)x * a.ncon + c] + a.vwgt[(int64_t)y * a.ncon + c] > cap[c]) return false;
    return true;
}

__global__ __launch_bounds__(kBlock) void k_suitor_init(int32_t n, const int32_t* __restrict__ xadj,
                                                        uint64_t* __restrict__ suitor) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
        suitor[v] = no_suitor_word(degree_rank(xadj[v + 1] - xadj[v]));
}

// State of a group's chain: the current searcher (-1: none), the neighbour
// of its last failed claim and the word that CAS returned, and whether its
// scans read the words at agent scope.  A searcher's first scan reads them
// through the caches (plain loads may return older words, which only make a
// neighbour look easier to claim: the claim's CAS then fails); after a
// failed claim its scans read them at agent scope, as coherent as the CAS,
// until it holds a vertex and the displaced suitor starts a first scan of
// its own.
struct Chain {
    int32_t cur = -1;
    int32_t fu = -1;
    uint64_t fword = 0;
    bool fresh = false;
};

// One step of a group's chain (see k_suitor): the searcher's scan and claim.
// CAPPED applies the weight cap (cap non-null); the kernel inlines both so
// that graphs whose cap cannot bind run the uncapped code.  Each lane takes
// B neighbours per round (G * B per group): their ids, then their weights,
// then their suitor words are all loaded before the first compare, at
// indices clamped into the row (the current vertex has edges), so that one
// round costs two dependent memory round trips whatever B is.
template <int G, int B, bool CAPPED>
__device__ __forceinline__ void suitor_step(const MatchArgs& a, int gl, uint64_t seed, int32_t base,
                                            const long long* cap, Chain& ch) {
    const int32_t cur = ch.cur;
    const int32_t fu = ch.fu;
    const uint64_t fword = ch.fword;
    const bool fresh = ch.fresh;
    const int32_t beg = a.xadj[cur], end = a.xadj[cur + 1];
    const uint32_t rc = degree_rank(end - beg);
    EdgeKey best{0ull, 0u};
    int32_t bu = -1, bw = 0;
    uint32_t br = 0;          // the best neighbour's degree rank
    uint64_t bseen = 0;
    // The best claimable neighbour, first over every neighbour: when
    // that pair fits the weight cap it is also the best among the
    // fitting ones; only otherwise is the scan repeated with the cap
    // applied to each neighbour (the weights are read only then).
    for (bool filt = false;;) {
        best = EdgeKey{0ull, 0u};
        bu = -1;
        bw = 0;
        br = 0;
        bseen = 0;
        for (int32_t j0 = beg + gl; j0 < end; j0 += G * B) {
            int32_t uu[B], ww[B];
            uint64_t ss[B];
#pragma unroll
            for (int i = 0; i < B; ++i) uu[i] = a.adjncy[min(j0 + i * G, end - 1)];
#pragma unroll
            for (int i = 0; i < B; ++i) ww[i] = a.adjwgt ? a.adjwgt[min(j0 + i * G, end - 1)] : 1;
#pragma unroll
            for (int i = 0; i < B; ++i) ss[i] = fresh ? ld_agent64(&a.suitor[uu[i]]) : a.suitor[uu[i]];
#pragma unroll
            for (int i = 0; i < B; ++i) {
                if (j0 + i * G >= end) break;
                const int32_t u = uu[i];
                if (CAPPED && filt && !pair_fits(a, cur, u, cap)) continue;   // the pair would be too heavy
                const int32_t w = ww[i];
                uint64_t sv = ss[i];
                const uint32_t ru = owner_rank(sv, u, a.xadj);
                if (u == fu && sv != fword) {
                    // keep the newer of the loaded word and the CAS result
                    const int32_t s1 = suitor_id(sv);
                    const int32_t s2 = suitor_id(fword);
                    if (s1 < 0) sv = fword;
                    else if (s2 >= 0) {
                        const EdgeKey k1 = suitor_key(sv, s1, u, ru, a.xadj, seed, a.shem, base);
                        const EdgeKey k2 = suitor_key(fword, s2, u, ru, a.xadj, seed, a.shem, base);
                        if (edge_before(k2, s2 - base, u - base, k1, s1 - base, u - base)) sv = fword;
                    }
                }
                const int32_t sid = suitor_id(sv);
                if (sid == cur) continue;
                const EdgeKey k = edge_key_r(cur - base, u - base, w, min(rc, ru), seed, a.shem);
                if (sid >= 0) {
                    const EdgeKey ks = suitor_key(sv, sid, u, ru, a.xadj, seed, a.shem, base);
                    if (!edge_before(k, cur - base, u - base, ks, sid - base, u - base)) continue;
                }
                if (bu < 0 || edge_before(k, cur - base, u - base, best, cur - base, bu - base)) {
                    best = k;
                    bu = u;
                    bw = w;
                    br = ru;
                    bseen = sv;
                }
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            const EdgeKey ob{__shfl_xor(best.hi, o, G), __shfl_xor(best.lo, o, G)};
            const int32_t ou = __shfl_xor(bu, o, G);
            const int32_t ow = __shfl_xor(bw, o, G);
            const uint32_t orr = __shfl_xor(br, o, G);
            const uint64_t os = __shfl_xor(bseen, o, G);
            if (ou >= 0 && (bu < 0 || edge_before(ob, cur - base, ou - base, best, cur - base, bu - base))) {
                best = ob;
                bu = ou;
                bw = ow;
                br = orr;
                bseen = os;
            }
        }
        if (CAPPED && !filt && cap && bu >= 0 && !pair_fits(a, cur, bu, cap)) {   // group-uniform
            filt = true;
            continue;
        }
        break;
    }
    if (bu < 0) {                                           // nothing to claim: cur stays unmatched
        ch.cur = -1;
        return;
    }
    // Claim bu.  When the word changed meanwhile but its new suitor is
    // still beaten, claim again at once: bu stays the best choice
    // (the other neighbours can only have become harder to claim).
    uint64_t old = 0;
    if (gl == 0) {
        const unsigned long long mine = suitor_word(cur, bw, min(rc, br), br);
        for (;;) {
            old = atomicCAS((unsigned long long*)&a.suitor[bu], (unsigned long long)bseen, mine);
            if (old == bseen) break;
            const int32_t sid = suitor_id(old);
            if (sid == cur) break;
            const EdgeKey ks = suitor_key(old, sid, bu, br, a.xadj, seed, a.shem, base);
            if (!edge_before(best, cur - base, bu - base, ks, sid - base, bu - base)) break;
            bseen = old;                                        // still beaten: claim against it
        }
    }
    old = __shfl(old, 0, G);
    bseen = __shfl(bseen, 0, G);
    if (old == bseen) {
        ch.cur = suitor_id(bseen);                             // the displaced suitor searches again
        ch.fu = -1;
        ch.fresh = false;
    } else {
        ch.fu = bu;
        ch.fword = old;
        ch.fresh = true;
    }
}

// G lanes per searching vertex (a group), B neighbours per lane and scan
// round, at least W waves per SIMD.  A vertex v of the level starts a chain: the
// current searcher scans its neighbours for the best one (edge order) whose
// suitor it beats, and claims it with a 64-bit CAS on the suitor word it
// compared against; on success the displaced suitor (if any) becomes the
// searcher, on failure the scan is repeated.  Keys of a claimed word only
// grow, so a stale read can only make a neighbour look easier to claim (the
// CAS then fails and returns the current word, which is used for that
// neighbour in the next scan): every retry sees a newer word, and the chain
// ends.
//
// Each loop iteration is one step of every group's chain.  A wave owns the
// vertex slots lo + w*GPW + i + r*S (i < GPW groups per wave, rounds r; w
// the wave's index and S the groups among the waves walking the part of the
// level from lo, below), dealt in that order: a group whose chain ended takes the
// wave's next slot at once (a ballot ranks the groups asking), so the groups
// do not wait for the longest chain of the wave, while the waves still walk
// the level together round by round (neighbouring vertices' words stay in
// the caches).
template <int G, int B, bool CAPPED, int W = 1>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W))) void k_suitor(MatchArgs a) {
    constexpr int GPW = 64 / G;
    const int gl = group_lane<G>();
    const int lane = __lane_id();
    // Blocks are dealt to the XCDs round-robin; when the grid is a multiple
    // of their count, each XCD's waves walk their own contiguous eighth of
    // the level (neighbours' rows and words, mostly in the same eighth, then
    // meet in that XCD's L2), else all waves walk the whole level.
    int64_t S, w0, hi = a.n;
    if (gridDim.x % kXcds == 0) {
        const int64_t reg = ((int64_t)a.n + kXcds - 1) / kXcds;
        const int64_t lo = min((int64_t)(blockIdx.x % kXcds) * reg, (int64_t)a.n);
        hi = min(lo + reg, (int64_t)a.n);
        S = (int64_t)(gridDim.x / kXcds) * (kBlock / G);
        w0 = lo + ((int64_t)(blockIdx.x / kXcds) * kBlock + threadIdx.x) / 64 * GPW;
    } else {
        S = (int64_t)gridDim.x * (kBlock / G);
        w0 = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / 64 * GPW;
    }
    int64_t slot = 0;             // wave-uniform: the wave's next slot
    Chain ch;
    bool more = true;             // group-uniform: slots may remain for the group
    // keys use graph-local ids and the graph's seed, so a graph of a batch
    // is matched exactly as it is alone
    uint64_t seed = a.seed;
    int32_t base = 0;
    const long long* cap = nullptr;
    for (;;) {
        // deal slots to the groups without a chain until each has one or the
        // level is done (a slot of a vertex without work is passed over)
        for (;;) {
            const bool ask = more && ch.cur < 0;
            const uint64_t askers = __ballot(ask && gl == 0);
            if (askers == 0) break;
            const int rank = __popcll(askers & ((1ull << lane) - 1ull));   // valid on the group's lane 0
            const int64_t t = slot + __shfl(rank, lane & ~(G - 1), 64);
            slot += __popcll(askers);
            // The slot's vertex, taken unless it is past the wave's part of
            // the level (every later slot is past it too), has no edges, or belongs to a graph
            // not matched here.  Every value is written on every asking lane
            // (selects, no divergent assignments).
            const int64_t v = w0 + (t % GPW) + (t / GPW) * S;
            const bool inr = ask && v < hi;
            const int64_t vv = inr ? v : 0;
            // the row bounds and the graph are loaded together
            const int32_t x0 = a.xadj[vv], x1 = a.xadj[vv + 1];
            const int32_t gi = a.gseg ? a.gseg[vv] : 0;
            bool take = inr && x1 != x0;
            if (a.gseg) take = take && a.gon[gi] != 0;
            if (ask) more = inr;
            if (a.gseg) {
                const uint64_t gsd = take ? a.gseed[gi] : 0;
                const int32_t gb = take ? a.gv0[gi] : 0;
                seed = take ? gsd : seed;
                base = take ? gb : base;
            }
            if constexpr (CAPPED) {
                const long long* c = a.cap && a.gseg ? a.cap + (int64_t)gi * a.ncon : a.cap;
                const bool bind = take && c && c[0] != LLONG_MAX;   // else the cap cannot bind here
                cap = take ? (bind ? c : nullptr) : cap;
            }
            ch.cur = take ? (int32_t)vv : ch.cur;
            ch.fu = take ? -1 : ch.fu;
            ch.fresh = take ? false : ch.fresh;
        }
        if (ch.cur < 0) break;                                 // only when no slot is left
        if (CAPPED && cap) suitor_step<G, B, true>(a, gl, seed, base, cap, ch);
        else suitor_step<G, B, false>(a, gl, seed, base, nullptr, ch);
    }
}

// Mutual suitors are matched; every other vertex is matched to itself.
__global__ __launch_bounds__(kBlock) void k_suitor_pairs(int32_t n, const uint64_t* __restrict__ suitor,
                                                         int32_t* __restrict__ match) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
        const int32_t s = suitor_id(suitor[v]);
        match[v] = (s >= 0 && suitor_id(suitor[s]) == (int32_t)v) ? s : (int32_t)v;
    }
}

// Batched levels: vertices of graphs not matched at this level (gseg/gon)
// lead no coarse vertex and get cmap -1; they leave the hierarchy here.
// Also clears the contraction's accumulators (degsum, cdeg; counters 2..9)
// in place of separate fills.
__global__ __launch_bounds__(kBlock) void k_leader(int32_t n, const int32_t* __restrict__ match,
                                                   int32_t* __restrict__ leader, const int32_t* __restrict__ gseg,
                                                   const uint8_t* __restrict__ gon, int32_t* __restrict__ degsum,
                                                   int32_t* __restrict__ cdeg, int32_t* __restrict__ counters) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n) {
        leader[v] = (match[v] >= (int32_t)v && (!gseg || gon[gseg[v]])) ? 1 : 0;
        degsum[v] = 0;
        cdeg[v] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x >= 2 && threadIdx.x < 10) counters[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kBlock) void k_cmap(int32_t n, int32_t ncon, const int32_t* __restrict__ xadj,
                                                 const int32_t* __restrict__ vwgt, const int32_t* __restrict__ vsize,
                                                 const int32_t* __restrict__ match, const int32_t* __restrict__ cid,
                                                 int32_t* __restrict__ cmap, int32_t* __restrict__ members,
                                                 int32_t* __restrict__ degsum, int32_t* __restrict__ cvwgt,
                                                 int32_t* __restrict__ cvsize, const int32_t* __restrict__ gseg,
                                                 const uint8_t* __restrict__ gon, int32_t* __restrict__ cgseg) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= n) return;
    if (gseg && !gon[gseg[v]]) {
        cmap[v] = -1;
        return;
    }
    const int32_t u = match[v];
    const int32_t a = min((int32_t)v, u);
    const int32_t cv = cid[a];
    cmap[v] = cv;
    if (u >= (int32_t)v) {
        const bool pair = u != (int32_t)v;
        members[2 * (int64_t)cv] = (int32_t)v;
        members[2 * (int64_t)cv + 1] = pair ? u : -1;
        degsum[cv] = (xadj[v + 1] - xadj[v]) + (pair ? xadj[u + 1] - xadj[u] : 0);
        if (cvwgt)
            for (int c = 0; c < ncon; ++c)
                cvwgt[(int64_t)cv * ncon + c] =
                    wadd(vwgt[v * ncon + c], pair ? vwgt[(int64_t)u * ncon + c] : 0);
        if (cvsize) cvsize[cv] = wadd(vsize[v], pair ? vsize[u] : 0);
        if (cgseg) cgseg[cv] = gseg[v];
    }
}

// Position t of the merged adjacency sequence adj(a) ++ adj(b) -> CSR index.
__device__ inline int32_t seq_edge(const int32_t* xadj, int32_t a, int32_t b, int32_t dega, int32_t t) {
    return t < dega ? xadj[a] + t : xadj[b] + (t - dega);
}


constexpr int kCWaves = 2;               // waves (independent tiles) per contraction block
constexpr int kCCap = 256;               // merged-sequence positions a wave merges per batch

__device__ inline int wave_excl_scan_c(int x, int* total) {
    const int lane = __lane_id();
    int inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    *total = __shfl(inc, 63, 64);
    return inc - x;
}

// Every wave merges tiles of 64 consecutive coarse vertices (lane = coarse
// vertex) on its own, with its own LDS area and no block barriers; waves are
// persistent over XCD-contiguous tile sequences (wave_tiles).  The merged
// sequences adj(a) ++ adj(b) of the tile's ordinary coarse vertices are
// taken in batches of whole vertices of at most CAP positions: each lane
// marks its own positions in a byte array of owners, then CAP / 64
// independent coalesced loads per lane fetch adjncy/adjwgt, and the cmap[]
// gathers follow.  The whole wave merges the batch at once in an LDS hash
// table: each coarse vertex of the batch owns a sub-table of twice its
// degree sum, where an entry claims its coarse neighbour's slot (CAS), adds
// its weight and keeps the smallest position (the first occurrence);
// internal entries are dropped.  The first occurrences, ranked in position
// order by one ballot per round, go out in sequence order at the vertex's
// tmp offset, so that the work of a batch is its entry count, whatever the
// degrees of its vertices.
template <int CAP, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_contract_tile(const int32_t* __restrict__ cn_dev,
                                                              const int32_t* __restrict__ xadj,
                                                              const int32_t* __restrict__ adjncy,
                                                              const int32_t* __restrict__ adjwgt,
                                                              const int32_t* __restrict__ cmap,
                                                              const int32_t* __restrict__ members,
                                                              const int32_t* __restrict__ degsum,
                                                              const int32_t* __restrict__ tmp_off,
                                                              int32_t* __restrict__ tmp_adj,
                                                              int32_t* __restrict__ tmp_w, int32_t* __restrict__ cdeg,
                                                              int32_t* __restrict__ big_list,
                                                              int32_t* __restrict__ counters, int32_t big_min) {
    constexpr int R = CAP / 64;               // batch positions per lane
    constexpr int T = 2 * CAP;                // hash slots
    __shared__ int32_t s_key_all[WAVES][T];   // coarse neighbour (-1: empty)
    __shared__ int32_t s_sum_all[WAVES][T];   // weight sum
    __shared__ int32_t s_pos_all[WAVES][T];   // first batch position
    __shared__ uint8_t s_own_all[WAVES][CAP];
    __shared__ int32_t s_loff_all[WAVES][65], s_xa_all[WAVES][64], s_xb_all[WAVES][64], s_da_all[WAVES][64],
        s_toff_all[WAVES][64], s_qb_all[WAVES][64], s_qe_all[WAVES][64];
    const int wv = threadIdx.x >> 6, lane = __lane_id();
    int32_t* s_key = s_key_all[wv];
    int32_t* s_sum = s_sum_all[wv];
    int32_t* s_pos = s_pos_all[wv];
    uint8_t* s_own = s_own_all[wv];
    int32_t *s_loff = s_loff_all[wv], *s_xa = s_xa_all[wv], *s_xb = s_xb_all[wv], *s_da = s_da_all[wv];
    int32_t *s_toff = s_toff_all[wv], *s_qb = s_qb_all[wv], *s_qe = s_qe_all[wv];
    const uint64_t below = (1ull << lane) - 1ull;
    const int32_t cn = *cn_dev;
    const int ntiles = (cn + 63) / 64;
    const WaveTiles wt = wave_tiles(ntiles, WAVES);
    bool not_one = false;
    for (int it = 0; it < wt.count; ++it) {
        const int tile = wt.base + it * wt.stride;
        const int32_t tbase = tile * 64;
        const int64_t cv = (int64_t)tbase + lane;
        const bool own = cv < cn;
        int32_t ds = 0, xa = 0, xb = 0, da = 0, toff = 0;
        if (own) {
            const int32_t a = members[2 * cv], b = members[2 * cv + 1];
            ds = degsum[cv];
            toff = tmp_off[cv];
            xa = xadj[a];
            da = xadj[a + 1] - xa;
            xb = b >= 0 ? xadj[b] : 0;
        }
        const bool big = own && ds > big_min;
        const uint64_t bigm = __ballot(big);
        if (bigm) {                                   // wave-aggregated append to the big list
            int base = 0;
            if (lane == 0) base = atomicAdd(&counters[2], __popcll(bigm));
            base = __shfl(base, 0, 64);
            if (big) big_list[base + __popcll(bigm & below)] = (int32_t)cv;
        }
        const int sds = (own && !big) ? ds : 0;
        int total;
        const int loff = wave_excl_scan_c(sds, &total);
        s_loff[lane] = loff;
        s_xa[lane] = xa;
        s_xb[lane] = xb;
        s_da[lane] = da;
        s_toff[lane] = toff;
        if (lane == 63) s_loff[64] = total;
        for (int vs = 0; vs < 64;) {
            const int P0 = __shfl(loff, vs, 64);
            const uint64_t fit = __ballot(lane >= vs && loff + sds - P0 <= CAP);
            const int ve = 64 - __builtin_clzll(fit);               // fit is lanes [vs, ve), ve > vs
            const int np = __shfl(loff + sds, ve - 1, 64) - P0;     // batch positions
            if (lane >= vs && lane < ve)
                for (int e = 0; e < sds; ++e) s_own[loff - P0 + e] = (uint8_t)lane;
            for (int i = lane; i < 2 * np; i += 64) {
                s_key[i] = -1;
                s_sum[i] = 0;
                s_pos[i] = INT32_MAX;
            }
            wave_sync();
            int32_t uu[R], ww[R], oo[R];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int p = lane + i * 64;
                uu[i] = 0;
                ww[i] = 1;
                oo[i] = 0;
                if (p < np) {
                    const int o = s_own[p];
                    oo[i] = o;
                    const int32_t tl = P0 + p - s_loff[o];
                    const int32_t j = tl < s_da[o] ? s_xa[o] + tl : s_xb[o] + (tl - s_da[o]);
                    uu[i] = adjncy[j];
                    if (adjwgt) ww[i] = adjwgt[j];
                }
            }
            // all gathers in flight before the first LDS access (positions
            // past the batch read cmap[0])
            int32_t cc[R];
#pragma unroll
            for (int i = 0; i < R; ++i) cc[i] = cmap[uu[i]];
            // insert: the slot of (owner, coarse neighbour) by linear probing
            // in the owner's sub-table, then its weight and first position
            int32_t sl[R];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int p = lane + i * 64;
                sl[i] = -1;
                if (p < np && cc[i] != tbase + oo[i]) {                 // internal edge (graph/mod.rs:330-332)
                    const int o = oo[i];
                    const int r0 = 2 * (s_loff[o] - P0);
                    const int r1 = 2 * (s_loff[o + 1] - P0);
                    int k = r0 + (int)(((uint64_t)mix32((uint32_t)cc[i]) * (uint32_t)(r1 - r0)) >> 32);
                    for (;;) {
                        const int32_t prev = atomicCAS(&s_key[k], -1, cc[i]);
                        if (prev == -1 || prev == cc[i]) break;
                        k = (k + 1 == r1) ? r0 : k + 1;
                    }
                    atomicAdd(&s_sum[k], ww[i]);                          // merge (graph/mod.rs:335-339)
                    atomicMin(&s_pos[k], p);
                    sl[i] = k;
                }
            }
            wave_sync();
            // rank the first occurrences in position order; the lanes holding
            // a vertex's first and last positions record the ranks around it
            int carry = 0;
            int32_t qq[R], sm[R];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int p = lane + i * 64;
                bool f = false;
                sm[i] = 0;
                if (sl[i] >= 0) {
                    f = s_pos[sl[i]] == p;
                    sm[i] = s_sum[sl[i]];
                }
                const uint64_t fm = __ballot(f);
                const int q = carry + __popcll(fm & below);
                carry += __popcll(fm);
                qq[i] = f ? q : -1;
                if (p < np) {
                    const int o = oo[i];
                    if (P0 + p == s_loff[o]) s_qb[o] = q;
                    if (P0 + p + 1 == s_loff[o + 1]) s_qe[o] = q + (f ? 1 : 0);
                }
            }
            wave_sync();
            if (lane >= vs && lane < ve && own && !big) cdeg[cv] = sds > 0 ? s_qe[lane] - s_qb[lane] : 0;
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (qq[i] >= 0) {
                    const int o = oo[i];
                    const int32_t at = s_toff[o] + (qq[i] - s_qb[o]);
                    tmp_adj[at] = cc[i];
                    tmp_w[at] = sm[i];
                    if (sm[i] != 1) not_one = true;
                }
            wave_sync();
            vs = ve;
        }
    }
    if (__ballot(not_one) && lane == 0) atomicOr(&counters[4], 1);
}

__device__ inline int32_t ld_agent(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block per large coarse vertex; open-addressing table of 2*ds+1 entries in
// global memory, carved from a pool with an atomic bump pointer (the table
// position does not affect the result).
__global__ __launch_bounds__(kBlock) void k_contract_big(const int32_t* __restrict__ xadj,
                                                         const int32_t* __restrict__ adjncy,
                                                         const int32_t* __restrict__ adjwgt,
                                                         const int32_t* __restrict__ cmap,
                                                         const int32_t* __restrict__ members,
                                                         const int32_t* __restrict__ degsum,
                                                         const int32_t* __restrict__ tmp_off,
                                                         int32_t* __restrict__ tmp_adj, int32_t* __restrict__ tmp_w,
                                                         int32_t* __restrict__ cdeg, const int32_t* __restrict__ big_list,
                                                         int32_t* __restrict__ counters, int32_t* __restrict__ hkeys,
                                                         int32_t* __restrict__ hfirst, int32_t* __restrict__ hsum) {
    __shared__ int64_t s_base;
    __shared__ int s_tmp[kBlock / 64];
    const int32_t nbig = counters[2];
    bool not_one = false;
    for (int bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
        const int32_t cv = big_list[bi];
        const int32_t ds = degsum[cv];
        const int32_t a = members[2 * (int64_t)cv], b = members[2 * (int64_t)cv + 1];
        const int32_t dega = xadj[a + 1] - xadj[a];
        const int64_t H = 2 * (int64_t)ds + 1;
        if (threadIdx.x == 0) s_base = (int64_t)atomicAdd((unsigned long long*)&counters[6], (unsigned long long)H);
        __syncthreads();
        int32_t* K = hkeys + s_base;
        int32_t* F = hfirst + s_base;
        int32_t* W = hsum + s_base;
        for (int64_t i = threadIdx.x; i < H; i += kBlock) {
            K[i] = -1;
            F[i] = INT32_MAX;
            W[i] = 0;
        }
        __threadfence();
        __syncthreads();
        for (int32_t t = threadIdx.x; t < ds; t += kBlock) {
            const int32_t j = seq_edge(xadj, a, b, dega, t);
            const int32_t cu = cmap[adjncy[j]];
            if (cu == cv) continue;
            const int32_t w = adjwgt ? adjwgt[j] : 1;
            int64_t s = mix32((uint32_t)cu) % (uint64_t)H;
            for (;;) {
                const int32_t prev = atomicCAS(&K[s], -1, cu);
                if (prev == -1 || prev == cu) {
                    atomicMin(&F[s], t);
                    atomicAdd(&W[s], w);
                    break;
                }
                s = (s + 1 == H) ? 0 : s + 1;
            }
        }
        __threadfence();
        __syncthreads();
        const int32_t off = tmp_off[cv];
        int32_t carry = 0;
        for (int32_t t0 = 0; t0 < ds; t0 += kBlock) {
            const int32_t t = t0 + threadIdx.x;
            bool first = false;
            int32_t cu = -1, sum = 0;
            if (t < ds) {
                const int32_t j = seq_edge(xadj, a, b, dega, t);
                cu = cmap[adjncy[j]];
                if (cu != cv) {
                    int64_t s = mix32((uint32_t)cu) % (uint64_t)H;
                    while (ld_agent(&K[s]) != cu) s = (s + 1 == H) ? 0 : s + 1;
                    first = ld_agent(&F[s]) == t;
                    sum = ld_agent(&W[s]);
                }
            }
            int tot;
            const int pos = block_flag_scan(first, s_tmp, &tot);
            if (first) {
                tmp_adj[off + carry + pos] = cu;
                tmp_w[off + carry + pos] = sum;
                if (sum != 1) not_one = true;
            }
            carry += tot;
        }
        if (threadIdx.x == 0) cdeg[cv] = carry;
        __syncthreads();
    }
    if (__ballot(not_one) && __lane_id() == 0) atomicOr(&counters[4], 1);
}

// The merged lists from the degree-sum layout of tmp to the coarse CSR.  A
// wave takes 64 consecutive coarse vertices, whose coarse ranges are one
// contiguous span: lane q of a round copies span position q, found by a
// binary search over the 64 range starts in LDS (the last vertex starting
// at or before q owns it), so that the stores of a round are one coalesced
// run whatever the degrees.
__global__ __launch_bounds__(kBlock) void k_contract_copy(int32_t n_max, const int32_t* __restrict__ cn_dev,
                                                          const int32_t* __restrict__ tmp_off,
                                                          const int32_t* __restrict__ tmp_adj,
                                                          const int32_t* __restrict__ tmp_w,
                                                          const int32_t* __restrict__ cxadj,
                                                          int32_t* __restrict__ cadj, int32_t* __restrict__ cw,
                                                          int32_t* __restrict__ counters) {
    constexpr int WPB = kBlock / 64;
    __shared__ int32_t s_dst_all[WPB][64], s_src_all[WPB][64];
    const int wv = threadIdx.x >> 6, lane = __lane_id();
    int32_t* s_dst = s_dst_all[wv];
    int32_t* s_src = s_src_all[wv];
    const int32_t cn = *cn_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[1] = cxadj[n_max];   // cm, read with the other counters
    const int64_t ntiles = ((int64_t)cn + 63) / 64;
    for (int64_t t = (int64_t)blockIdx.x * WPB + wv; t < ntiles; t += (int64_t)gridDim.x * WPB) {
        const int64_t v0 = t * 64;
        const int last = (int)min<int64_t>(64, cn - v0);            // vertices of this tile
        const int64_t cv = v0 + (lane < last ? lane : last - 1);
        const int32_t d0 = cxadj[v0];
        const int32_t span = cxadj[v0 + last] - d0;
        s_dst[lane] = cxadj[cv] - d0;
        s_src[lane] = tmp_off[cv];
        wave_sync();
        for (int32_t q = lane; q < span; q += 64) {
            int lo = 0, hi = last - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_dst[mid] <= q) lo = mid;
                else hi = mid - 1;
            }
            const int32_t from = s_src[lo] + (q - s_dst[lo]);
            cadj[d0 + q] = tmp_adj[from];
            cw[d0 + q] = tmp_w[from];
        }
        wave_sync();
    }
}

inline int grid_cap(int64_t items, int per_block, int cap) {
    int64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

}  // namespace

// Blocks of kBlock threads of `kern` resident on one CU (the occupancy
// query, cached per slot): persistent grids are sized to exactly what runs
// at once, since their waves own fixed shares of the work.
int resident_blocks(int slot, void (*kern)(MatchArgs)) {
    static std::atomic<int> cache[64] = {};
    int nb = cache[slot].load(std::memory_order_relaxed);
    if (nb <= 0) {
        RK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, kBlock, 0));
        nb = nb > 0 ? nb : 1;
        cache[slot].store(nb, std::memory_order_relaxed);
    }
    return nb;
}

// The suitor kernel's shape for a level: G lanes per chain and B neighbours
// per lane and round, chosen per level size and mean degree by the tuning
// table.
// Fewer lanes per chain keep more chains in flight (a step is a few
// dependent memory round trips and a CAS, so a level's time is its steps
// over the chains in flight); G * B covers a mean row in one or two rounds.
void match_graph(const DevGraph& g, int ctype_shem, uint64_t seed, CoarsenBufs& b, hipStream_t s,
                 const GraphSegs* gs, const long long* cap) {
    if (g.n <= 0) return;
    const MatchArgs a{g.n, ctype_shem, seed, g.xadj, g.adjncy, g.adjwgt, b.match, b.suitor, g.vwgt,
                      g.vwgt ? cap : nullptr, g.ncon, gs ? gs->gseg : nullptr, gs ? gs->gv0 : nullptr,
                      gs ? gs->gseed : nullptr, gs ? gs->gon : nullptr};
    const int grid = grid_cap(g.n, kBlock, 4096);
    hipLaunchKernelGGL(k_suitor_init, dim3(grid), dim3(kBlock), 0, s, g.n, g.xadj, b.suitor);
    const int var = tune(level_bucket(g.n, g.m));
    const bool capped = a.cap != nullptr;   // the capped variant only where a pair can reach the cap
    switch (var) {
#define RK_SUITOR(IDX, G_, B_, W_)                                                                            \
    case IDX: {                                                                                           \
        const auto kern = capped ? k_suitor<G_, B_, true, W_> : k_suitor<G_, B_, false, W_>;              \
        const int sgrid = grid_cap(g.n, kBlock / G_, device_cus() * resident_blocks(2 * IDX + capped, kern)); \
        hipLaunchKernelGGL(kern, dim3(sgrid), dim3(kBlock), 0, s, a);                                     \
        break;                                                                                            \
    }
        RK_SUITOR(0, 1, 4, 1)
        RK_SUITOR(1, 2, 4, 1)
        RK_SUITOR(2, 4, 2, 1)
        RK_SUITOR(3, 8, 1, 8)
#undef RK_SUITOR
        default: break;
    }
    hipLaunchKernelGGL(k_suitor_pairs, dim3(grid), dim3(kBlock), 0, s, g.n, b.suitor, b.match);
    RK_HIP(hipGetLastError());
}

ContractResult contract(const DevGraph& g, const int32_t* match, int32_t* cmap, DevGraph& cg, bool aggregate_vsize,
                        CoarsenBufs& b, CoarseAlloc& al, int32_t* heavy_list, hipStream_t s, const GraphSegs* gs) {
    // Coarse sizes stay on the device until the single synchronisation at the
    // end; arrays are allocated with the upper bounds cn <= n, cm <= m and
    // every kernel reads cn = counters[0].
    const int32_t n = g.n;
    const int nb = (n + kBlock - 1) / kBlock;
    int32_t* cn_dev = b.counters + 0;
    cg.ncon = g.ncon;
    cg.vwgt = g.vwgt ? al.alloc_i32((int64_t)n * g.ncon) : nullptr;
    cg.vsize = (aggregate_vsize && g.vsize) ? al.alloc_i32(n) : nullptr;
    cg.xadj = al.alloc_i32((int64_t)n + 1);
    cg.adjncy = al.alloc_i32(g.m > 0 ? g.m : 1);
    cg.adjwgt = al.alloc_i32(g.m > 0 ? g.m : 1);
    const int32_t* gseg = gs ? gs->gseg : nullptr;
    const uint8_t* gon = gs ? gs->gon : nullptr;
    hipLaunchKernelGGL(k_leader, dim3(nb), dim3(kBlock), 0, s, n, match, b.leader, gseg, gon, b.degsum, b.cdeg,
                       b.counters);
    exclusive_scan_i32(b.leader, b.cid, n, cn_dev, b.scan_tmp, s);      // coarse ids (graph/mod.rs:288-297)
    hipLaunchKernelGGL(k_cmap, dim3(nb), dim3(kBlock), 0, s, n, g.ncon, g.xadj, g.vwgt, g.vsize, match, b.cid, cmap,
                       b.members, b.degsum, cg.vwgt, cg.vsize, gseg, gon, gs ? gs->cgseg : nullptr);
    exclusive_scan_i32(b.degsum, b.tmp_off, n, nullptr, b.scan_tmp, s);
    {
        // 256-position batches in blocks of 2 waves (78 VGPRs, ~8 KB of LDS
        // per wave): on the 1,024-graph batch's levels 1.26x faster than
        // 512 x 4 and 1.02-1.09x faster than 256 x 4, 128 x 8 or 384 x 4
        static_assert(kContractBig <= kCCap, "a tile batch holds every ordinary coarse vertex");
        const auto kern = k_contract_tile<kCCap, kCWaves>;
        static std::atomic<int> occ{0};
        int nb = occ.load(std::memory_order_relaxed);
        if (nb <= 0) {
            RK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 64 * kCWaves, 0));
            nb = nb > 0 ? nb : 1;
            occ.store(nb, std::memory_order_relaxed);
        }
        const int blocks = (n + 64 * kCWaves - 1) / (64 * kCWaves);
        const int resident = device_cus() * nb;
        hipLaunchKernelGGL(kern, dim3(blocks < resident ? blocks : resident), dim3(64 * kCWaves), 0, s, cn_dev,
                           g.xadj, g.adjncy, g.adjwgt, cmap, b.members, b.degsum, b.tmp_off, b.tmp_adj, b.tmp_w,
                           b.cdeg, b.big_list, b.counters, kContractBig);
    }
    hipLaunchKernelGGL(k_contract_big, dim3(1024), dim3(kBlock), 0, s, g.xadj, g.adjncy, g.adjwgt, cmap,
                       b.members, b.degsum, b.tmp_off, b.tmp_adj, b.tmp_w, b.cdeg, b.big_list, b.counters,
                       b.hash_keys, b.hash_first, b.hash_sum);
    exclusive_scan_i32(b.cdeg, cg.xadj, n, cg.xadj + n, b.scan_tmp, s);   // cxadj[cn..n] = cm
    hipLaunchKernelGGL(k_contract_copy, dim3(grid_cap(((int64_t)n + 63) / 64, kBlock / 64, 65536)), dim3(kBlock), 0, s, n, cn_dev,
                       b.tmp_off, b.tmp_adj, b.tmp_w, cg.xadj, cg.adjncy, cg.adjwgt, b.counters);
    if (heavy_list) heavy_list_build(n, cn_dev, cg.xadj, heavy_list, b.counters + 8, s, true);
    // one read of counters[0..8] (cn, cm, ..., the weight flag, nheavy)
    int32_t* hc = b.host_small;
    RK_HIP(hipMemcpyAsync(hc, b.counters, 9 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipStreamSynchronize(s));
    RK_HIP(hipGetLastError());
    cg.n = hc[0];
    cg.m = hc[1];
    ContractResult r;
    r.cn = hc[0];
    r.cm = hc[1];
    r.adjwgt_not_one = hc[4] != 0;
    r.nheavy = heavy_list ? hc[8] : 0;
    return r;
}

}  // namespace rk
