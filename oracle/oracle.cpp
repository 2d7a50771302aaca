/*
 * oracle.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A loop-for-loop CPU restatement of the javagg/rmetis k-way greedy path.
 * It keeps the reference's own loop structure (per-(vertex, part) gain
 * rescans, linear-search edge merge) so that it doubles as the CPU baseline.
 * Compiled with -O3 -ffp-contract=off so the f64 threshold products round as
 * the reference's separate multiplies do.
 *
 * Parity status (how each function is pinned; tests/test_oracle.py):
 *   edge_cut, partition_weights, total_weight, boundary_vertices,
 *   compute_move_gain     -- the reference's own unit values (graph/mod.rs:455-485),
 *                            closed forms on banded grids, brute force on tiny graphs.
 *   greedy pass / refine  -- greedy.rs:229-251 and tests/parallel.rs:91-132
 *                            expectations, hand-derived cycle4 results, and an
 *                            independent brute-force transcription on tiny graphs.
 *   build_coarse_graph    -- graph/mod.rs:499-514, hand-derived path and grid
 *                            contractions, weight-conservation invariants.
 *   matchings             -- symmetry m[m[v]] == v (shem.rs:96-99, rm.rs:56-64),
 *                            maximality, heavy-edge preference on the shem.rs star.
 *   RNG (SmallRng)        -- published SplitMix64 and xoshiro256++ vectors;
 *                            gen_range by its closed form for power-of-two
 *                            ranges (rand 0.8.5's zone (range << lz) - 1 is
 *                            2^63 - 1 there: a draw is kept iff bit 63-j is
 *                            clear, the result is its top j bits) and by exact
 *                            rational arithmetic for range 3.
 *   grow_partition        -- hand-derived results: k = n gives the seed draw
 *                            order; a path grown from both ends splits at
 *                            ceil(n/2); a popped vertex claims every free
 *                            neighbour (star); leftovers go, in ascending order,
 *                            to the first part lightest on constraint 0.
 *   part_graph_kway       -- an independent Python transcription of the driver
 *                            (kway.rs:20-141) over the pieces pinned above:
 *                            RNG consumption order, f32 stop rule, projection,
 *                            targets per level, strictly-better attempt.  The
 *                            Rust crate itself cannot be run here (no cargo,
 *                            rand 0.8.5 not vendored), so the composed values
 *                            rest on the restated rand algorithms, not on
 *                            outputs of the crate.
 */
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <random>
#include <set>
#include <utility>
#include <vector>

namespace {

inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* Owned CSR graph mirroring graph/mod.rs:13-34 (empty vector = None). */
struct OGraph {
    int32_t n = 0, ncon = 1;
    std::vector<int32_t> xadj, adjncy, vwgt, adjwgt, vsize;
    bool has_vwgt = false, has_adjwgt = false, has_vsize = false;
    int32_t vw(int32_t v, int32_t c) const { return has_vwgt ? vwgt[(size_t)v * ncon + c] : 1; }  // graph/mod.rs:91-96
    int32_t ew(int64_t j) const { return has_adjwgt ? adjwgt[j] : 1; }                             // graph/mod.rs:111-116
    const int32_t* vwp() const { return has_vwgt ? vwgt.data() : nullptr; }
    const int32_t* ewp() const { return has_adjwgt ? adjwgt.data() : nullptr; }
};

int32_t edge_cut_impl(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                      const int32_t* adjwgt, const int32_t* part) {
    // graph/mod.rs:123-137
    int32_t cut = 0;
    for (int32_t v = 0; v < n; ++v) {
        int32_t pv = part[v];
        for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
            int32_t u = adjncy[j];
            if (part[u] != pv) cut = wadd(cut, adjwgt ? adjwgt[j] : 1);
        }
    }
    return cut / 2;
}

std::vector<int32_t> boundary_impl(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                                   const int32_t* part) {
    // graph/mod.rs:164-171
    std::vector<int32_t> b;
    for (int32_t v = 0; v < n; ++v) {
        int32_t pv = part[v];
        for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j)
            if (part[adjncy[j]] != pv) { b.push_back(v); break; }
    }
    return b;
}

int32_t move_gain(const int32_t* xadj, const int32_t* adjncy, const int32_t* adjwgt,
                  int32_t v, const int32_t* part, int32_t from, int32_t to) {
    // greedy.rs:197-218
    int32_t gain = 0;
    for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
        int32_t u = adjncy[j];
        int32_t w = adjwgt ? adjwgt[j] : 1;
        int32_t pu = part[u];
        if (pu == from) gain = wsub(gain, w);
        else if (pu == to) gain = wadd(gain, w);
    }
    return gain;
}

/* One pass of greedy.rs:36-193 with nprocs == 1.  mode 0: the balance-repair
 * block (60-123); mode 1: one iteration of the cut-optimisation loop (127-182).
 * The moves are applied to part at the end, exactly like the all_gather +
 * apply loops (110-119, 166-181). */
int32_t greedy_pass_impl(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                         const int32_t* vwgt, const int32_t* adjwgt, int32_t* part,
                         int32_t nparts, const double* tpw, const double* ub, int mode,
                         int32_t* pw, std::vector<int32_t>* moves, bool* any_move) {
    std::vector<int32_t> boundary = boundary_impl(n, xadj, adjncy, part);
    std::vector<int32_t> local_moves;
    std::vector<int32_t> vw(ncon);
    for (int32_t v : boundary) {
        int32_t pv = part[v];
        for (int32_t c = 0; c < ncon; ++c) vw[c] = vwgt ? vwgt[(size_t)v * ncon + c] : 1;
        int32_t best_gain, best_p;
        if (mode == 0) {
            bool source_overweight = false;                       // greedy.rs:69-74
            for (int32_t c = 0; c < ncon; ++c)
                if ((double)pw[(size_t)pv * ncon + c] > tpw[(size_t)pv * ncon + c] * ub[c]) {
                    source_overweight = true;
                    break;
                }
            if (!source_overweight) continue;
            best_gain = INT32_MIN;                                // greedy.rs:78-79
            best_p = -1;
        } else {
            best_gain = 0;                                        // greedy.rs:135-136
            best_p = pv;
        }
        for (int32_t p = 0; p < nparts; ++p) {                    // greedy.rs:81 / 138
            if (p == pv) continue;
            bool feasible = true;                                 // greedy.rs:85-88 / 142-145
            for (int32_t c = 0; c < ncon; ++c) {
                int32_t new_w = wadd(pw[(size_t)p * ncon + c], vw[c]);
                if (!((double)new_w <= tpw[(size_t)p * ncon + c] * ub[c])) { feasible = false; break; }
            }
            if (!feasible) continue;
            int32_t gain = move_gain(xadj, adjncy, adjwgt, v, part, pv, p);
            if (gain > best_gain) { best_gain = gain; best_p = p; }
        }
        bool moved = (mode == 0) ? (best_p != -1) : (best_p != pv);
        if (moved) {                                              // greedy.rs:99-106 / 156-163
            local_moves.push_back(v);
            local_moves.push_back(best_p);
            for (int32_t c = 0; c < ncon; ++c) {
                pw[(size_t)pv * ncon + c] = wsub(pw[(size_t)pv * ncon + c], vw[c]);
                pw[(size_t)best_p * ncon + c] = wadd(pw[(size_t)best_p * ncon + c], vw[c]);
            }
        }
    }
    bool any = false;
    for (size_t i = 0; i + 1 < local_moves.size(); i += 2) {
        int32_t v = local_moves[i], np = local_moves[i + 1];
        if (mode == 0) { part[v] = np; any = true; }             // greedy.rs:111-118
        else if (part[v] != np) { part[v] = np; any = true; }    // greedy.rs:170-179
    }
    if (any_move) *any_move = any;
    if (moves) *moves = local_moves;
    return (int32_t)(local_moves.size() / 2);
}

void pweights(int32_t n, int32_t ncon, const int32_t* vwgt, const int32_t* part,
              int32_t nparts, int32_t* pw) {
    // graph/mod.rs:140-149
    std::fill(pw, pw + (size_t)nparts * ncon, 0);
    for (int32_t v = 0; v < n; ++v) {
        int32_t p = part[v];
        for (int32_t c = 0; c < ncon; ++c)
            pw[(size_t)p * ncon + c] = wadd(pw[(size_t)p * ncon + c], vwgt ? vwgt[(size_t)v * ncon + c] : 1);
    }
}

int32_t greedy_refine_impl(const OGraph& g, int32_t* part, int32_t nparts, const double* tpw,
                           const double* ub, int32_t niter, int64_t* refine_edges) {
    // greedy.rs:36-193 with SingleComm (nprocs = 1, rank = 0)
    std::vector<int32_t> pw((size_t)nparts * g.ncon);
    pweights(g.n, g.ncon, g.vwp(), part, nparts, pw.data());                 // greedy.rs:50
    const int64_t m = g.xadj[g.n];
    greedy_pass_impl(g.n, g.ncon, g.xadj.data(), g.adjncy.data(), g.vwp(), g.ewp(), part, nparts,
                     tpw, ub, 0, pw.data(), nullptr, nullptr);
    if (refine_edges) *refine_edges += m;
    for (int32_t it = 0; it < niter; ++it) {                                 // greedy.rs:126
        bool any = false;
        greedy_pass_impl(g.n, g.ncon, g.xadj.data(), g.adjncy.data(), g.vwp(), g.ewp(), part,
                         nparts, tpw, ub, 1, pw.data(), nullptr, &any);
        if (refine_edges) *refine_edges += m;
        if (!any) break;                                                     // greedy.rs:183-185
    }
    return edge_cut_impl(g.n, g.xadj.data(), g.adjncy.data(), g.ewp(), part); // greedy.rs:192
}

// ---- refine/fm_kway.rs:12-158 (k-way FM) --------------------------------
//
// The reference keeps its queue in priority-queue 1.4.0's PriorityQueue
// (max-heap over i32 gains); the order in which it pops vertices of EQUAL
// gain is an implementation detail of that absent crate.  Here ties pop the
// smallest vertex id first.  Everything else follows fm_kway.rs line by line.

// fm_kway.rs:47-72: (best part, gain) over parts with a nonzero edge-weight
// sum, gain > 0 strictly, ties to the lower part; false when none.
bool fm_best_move(const OGraph& g, const int32_t* part, int32_t nparts, int32_t v, std::vector<int32_t>& edge_to,
                  int32_t* best_p, int32_t* best_g) {
    const int32_t pv = part[v];
    std::fill(edge_to.begin(), edge_to.begin() + nparts, 0);
    for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
        const int32_t u = g.adjncy[j];
        edge_to[part[u]] = wadd(edge_to[part[u]], g.ew(j));                 // 56-60
    }
    int32_t bp = -1, bg = 0;
    for (int32_t p = 0; p < nparts; ++p) {                                   // 62-70
        if (p == pv) continue;
        if (edge_to[p] == 0) continue;
        const int32_t gain = wsub(edge_to[p], edge_to[pv]);
        if (gain > bg) { bg = gain; bp = p; }
    }
    if (bp < 0) return false;
    *best_p = bp;
    *best_g = bg;
    return true;
}

// fm_kway.rs:33-158: one pass; returns its boolean (true when improved).
bool fm_pass_impl(const OGraph& g, int32_t* part, int32_t nparts, std::vector<int32_t>& pw, const double* tpw,
                  const double* ub, int32_t* nmoves_out) {
    const int32_t n = g.n, ncon = g.ncon;
    std::vector<int32_t> edge_to(nparts);
    std::vector<int32_t> bm_p(n, -1), bm_g(n, 0);       // best_moves (45): target, gain; -1 = None
    std::set<std::pair<int64_t, int32_t>> pq;            // (-gain, v): max gain first, then smallest v
    std::vector<int32_t> key(n, 0);                      // queued gain (0 = not queued)
    auto pq_set = [&](int32_t v, int32_t gain) {
        if (key[v] != 0) pq.erase({-(int64_t)key[v], v});
        key[v] = gain;
        pq.insert({-(int64_t)gain, v});
    };
    auto pq_remove = [&](int32_t v) {
        if (key[v] != 0) pq.erase({-(int64_t)key[v], v});
        key[v] = 0;
    };
    for (int32_t v = 0; v < n; ++v) {                                        // 76-85
        const int32_t pv = part[v];
        bool on_boundary = false;
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
            if (part[g.adjncy[j]] != pv) { on_boundary = true; break; }
        if (!on_boundary) continue;
        int32_t tp, gain;
        if (fm_best_move(g, part, nparts, v, edge_to, &tp, &gain)) {
            bm_p[v] = tp;
            bm_g[v] = gain;
            pq_set(v, gain);
        }
    }
    std::vector<char> locked(n, 0);
    std::vector<int32_t> mv_v, mv_from;
    int32_t cut = edge_cut_impl(n, g.xadj.data(), g.adjncy.data(), g.ewp(), part);   // 89
    int32_t best_cut = cut;
    size_t best_len = 0;
    std::vector<int32_t> vw(ncon);
    while (!pq.empty()) {                                                    // 93
        const int32_t v = pq.begin()->second;
        const int32_t gain = (int32_t)(-pq.begin()->first);
        pq.erase(pq.begin());
        key[v] = 0;
        if (locked[v] || gain <= 0) break;                                   // 94
        if (bm_p[v] < 0) continue;                                           // 96-99
        const int32_t to = bm_p[v], from = part[v];
        for (int32_t c = 0; c < ncon; ++c) vw[c] = g.vw(v, c);
        bool feasible = true;                                                // 105-109
        for (int32_t c = 0; c < ncon; ++c) {
            const int32_t new_w = wadd(pw[(size_t)to * ncon + c], vw[c]);
            if (!((double)new_w <= tpw[(size_t)to * ncon + c] * ub[c])) { feasible = false; break; }
        }
        if (!feasible) continue;
        part[v] = to;                                                        // 112-124
        for (int32_t c = 0; c < ncon; ++c) {
            pw[(size_t)from * ncon + c] = wsub(pw[(size_t)from * ncon + c], vw[c]);
            pw[(size_t)to * ncon + c] = wadd(pw[(size_t)to * ncon + c], vw[c]);
        }
        locked[v] = 1;
        mv_v.push_back(v);
        mv_from.push_back(from);
        cut = wsub(cut, gain);
        if (cut < best_cut) { best_cut = cut; best_len = mv_v.size(); }
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {                // 127-143
            const int32_t u = g.adjncy[j];
            if (locked[u]) continue;
            int32_t tp, ug;
            if (fm_best_move(g, part, nparts, u, edge_to, &tp, &ug)) {
                bm_p[u] = tp;
                bm_g[u] = ug;
                pq_set(u, ug);
            } else {
                pq_remove(u);
                bm_p[u] = -1;
            }
        }
    }
    for (size_t i = mv_v.size(); i > best_len; --i) {                        // 147-155 rollback
        const int32_t v = mv_v[i - 1], from = mv_from[i - 1], cur = part[v];
        for (int32_t c = 0; c < ncon; ++c) {
            pw[(size_t)cur * ncon + c] = wsub(pw[(size_t)cur * ncon + c], g.vw(v, c));
            pw[(size_t)from * ncon + c] = wadd(pw[(size_t)from * ncon + c], g.vw(v, c));
        }
        part[v] = from;
    }
    if (nmoves_out) *nmoves_out = (int32_t)best_len;
    const int32_t after = edge_cut_impl(n, g.xadj.data(), g.adjncy.data(), g.ewp(), part);
    return best_len > 0 && best_cut < wadd(after, 1);                        // 157
}

// fm_kway.rs:12-30
int32_t fm_refine_impl(const OGraph& g, int32_t* part, int32_t nparts, const double* tpw, const double* ub,
                       int32_t niter, int64_t* refine_edges) {
    std::vector<int32_t> pw((size_t)nparts * g.ncon);
    pweights(g.n, g.ncon, g.vwp(), part, nparts, pw.data());                 // 21
    for (int32_t it = 0; it < niter; ++it) {                                 // 23-27
        if (refine_edges) *refine_edges += g.xadj[g.n];
        if (!fm_pass_impl(g, part, nparts, pw, tpw, ub, nullptr)) break;
    }
    return edge_cut_impl(g.n, g.xadj.data(), g.adjncy.data(), g.ewp(), part);   // 29
}

// partition/kway.rs:145-162 (refine_level): Greedy or, through
// refine/mod.rs:26-38, k-way FM.
void refine_level_impl(const OGraph& g, int32_t* part, int32_t nparts, const double* tpw, const double* ub,
                       int32_t niter, int32_t rtype_greedy, int64_t* refine_edges) {
    if (rtype_greedy) greedy_refine_impl(g, part, nparts, tpw, ub, niter, refine_edges);
    else fm_refine_impl(g, part, nparts, tpw, ub, niter, refine_edges);
}

std::vector<int32_t> shem_impl(const OGraph& g, orc_rng* rng) {
    // coarsen/shem.rs:13-56
    const int32_t n = g.n;
    std::vector<int32_t> matching(n);
    for (int32_t v = 0; v < n; ++v) matching[v] = v;
    std::vector<char> visited(n, 0);
    std::vector<int32_t> perm(n);
    for (int32_t v = 0; v < n; ++v) perm[v] = v;
    std::vector<uint32_t> tb(n);
    for (int32_t v = 0; v < n; ++v) tb[v] = orc_rng_next_u32(rng) >> 16;   // shem.rs:22
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { // shem.rs:23 (stable)
        int32_t da = g.xadj[a + 1] - g.xadj[a], db = g.xadj[b + 1] - g.xadj[b];
        if (da != db) return da < db;
        return tb[a] < tb[b];
    });
    for (int32_t v : perm) {                                                 // shem.rs:25-54
        if (visited[v]) continue;
        visited[v] = 1;
        int32_t best_u = -1;
        int32_t best_w = INT32_MIN;
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
            int32_t u = g.adjncy[j];
            if (visited[u]) continue;
            int32_t w = g.ew(j);
            if (w > best_w) { best_w = w; best_u = u; }
        }
        if (best_u != -1) {
            matching[v] = best_u;
            matching[best_u] = v;
            visited[best_u] = 1;
        }
    }
    return matching;
}

std::vector<int32_t> rm_impl(const OGraph& g, orc_rng* rng) {
    // coarsen/rm.rs:13-43
    const int32_t n = g.n;
    std::vector<int32_t> matching(n);
    for (int32_t v = 0; v < n; ++v) matching[v] = v;
    std::vector<char> visited(n, 0);
    std::vector<int32_t> perm(n);
    for (int32_t v = 0; v < n; ++v) perm[v] = v;
    for (int64_t i = (int64_t)n - 1; i >= 1; --i) {                          // rm.rs:20-24
        uint64_t j = orc_rng_range_incl(rng, 0, (uint64_t)i);
        std::swap(perm[i], perm[j]);
    }
    for (int32_t v : perm) {
        if (visited[v]) continue;
        visited[v] = 1;
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
            int32_t u = g.adjncy[j];
            if (!visited[u]) {
                matching[v] = u;
                matching[u] = v;
                visited[u] = 1;
                break;
            }
        }
    }
    return matching;
}

/* graph/mod.rs:284-384 */
OGraph coarse_impl(const OGraph& g, const std::vector<int32_t>& matching, std::vector<int32_t>& cmap) {
    const int32_t n = g.n, ncon = g.ncon;
    cmap.assign(n, -1);
    int32_t cn = 0;
    for (int32_t v = 0; v < n; ++v) {                                        // 288-297
        if (cmap[v] == -1) {
            cmap[v] = cn;
            int32_t u = matching[v];
            if (u != v && cmap[u] == -1) cmap[u] = cn;
            ++cn;
        }
    }
    OGraph c;
    c.n = cn;
    c.ncon = ncon;
    c.has_vwgt = g.has_vwgt;
    c.has_vsize = g.has_vsize;
    if (c.has_vwgt) c.vwgt.assign((size_t)cn * ncon, 0);
    if (c.has_vsize) c.vsize.assign(cn, 0);
    for (int32_t v = 0; v < n; ++v) {                                        // 303-314
        int32_t cv = cmap[v];
        if (c.has_vwgt)
            for (int32_t k = 0; k < ncon; ++k)
                c.vwgt[(size_t)cv * ncon + k] = wadd(c.vwgt[(size_t)cv * ncon + k], g.vw(v, k));
        if (c.has_vsize) c.vsize[cv] = wadd(c.vsize[cv], g.vsize[v]);
    }
    std::vector<std::vector<std::pair<int32_t, int32_t>>> lists(cn);         // 321-341
    for (int32_t v = 0; v < n; ++v) {
        int32_t cv = cmap[v];
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
            int32_t cu = cmap[g.adjncy[j]];
            if (cu == cv) continue;
            int32_t ew = g.ew(j);
            auto& L = lists[cv];
            auto it = std::find_if(L.begin(), L.end(), [&](const std::pair<int32_t, int32_t>& e) { return e.first == cu; });
            if (it != L.end()) it->second = wadd(it->second, ew);
            else L.push_back({cu, ew});
        }
    }
    c.xadj.assign(cn + 1, 0);                                                // 344-364
    for (int32_t cv = 0; cv < cn; ++cv) c.xadj[cv + 1] = c.xadj[cv] + (int32_t)lists[cv].size();
    c.adjncy.reserve(c.xadj[cn]);
    std::vector<int32_t> w;
    w.reserve(c.xadj[cn]);
    bool any_not_one = false;
    for (int32_t cv = 0; cv < cn; ++cv)
        for (auto& e : lists[cv]) {
            c.adjncy.push_back(e.first);
            w.push_back(e.second);
            if (e.second != 1) any_not_one = true;
        }
    c.has_adjwgt = g.has_adjwgt || any_not_one;                              // 366-369
    if (c.has_adjwgt) c.adjwgt = std::move(w);
    return c;
}

std::vector<int32_t> total_weight_impl(const OGraph& g) {
    // graph/mod.rs:152-160
    std::vector<int32_t> tw(g.ncon, 0);
    for (int32_t v = 0; v < g.n; ++v)
        for (int32_t c = 0; c < g.ncon; ++c) tw[c] = wadd(tw[c], g.vw(v, c));
    return tw;
}

std::vector<double> targets_impl(const OGraph& g, int32_t nparts, const float* tpwgts) {
    // graph/mod.rs:392-410, flattened as kway.rs:67-69
    std::vector<int32_t> total = total_weight_impl(g);
    std::vector<double> t((size_t)nparts * g.ncon);
    for (int32_t p = 0; p < nparts; ++p)
        for (int32_t c = 0; c < g.ncon; ++c) {
            double f = tpwgts ? (double)tpwgts[(size_t)p * g.ncon + c] : 1.0 / (double)nparts;
            t[(size_t)p * g.ncon + c] = f * (double)total[c];
        }
    return t;
}

int grow_impl(const OGraph& g, int32_t nparts, const double* tpw, orc_rng* rng, int32_t* part) {
    // initial/grow.rs:12-103
    const int32_t n = g.n, ncon = g.ncon;
    if (n < nparts) return -1;  // grow.rs:24-32 would never terminate
    std::fill(part, part + n, -1);
    std::vector<int64_t> pwt((size_t)nparts * ncon, 0);
    std::vector<int32_t> seeds;
    std::vector<char> used(n, 0);
    while ((int32_t)seeds.size() < nparts) {                                 // 24-32
        uint64_t v = orc_rng_range_incl(rng, 0, (uint64_t)n - 1);
        if (!used[v]) { used[v] = 1; seeds.push_back((int32_t)v); }
    }
    std::vector<std::deque<int32_t>> queues(nparts);
    for (int32_t p = 0; p < nparts; ++p) queues[p].push_back(seeds[p]);
    for (int32_t p = 0; p < nparts; ++p) {                                   // 42-47
        int32_t s = seeds[p];
        part[s] = p;
        for (int32_t c = 0; c < ncon; ++c) pwt[(size_t)p * ncon + c] += g.vw(s, c);
    }
    int32_t active = nparts;
    while (active > 0) {                                                     // 50-86
        active = 0;
        for (int32_t p = 0; p < nparts; ++p) {
            bool at_target = true;
            for (int32_t c = 0; c < ncon; ++c)
                if (!((double)pwt[(size_t)p * ncon + c] >= tpw[(size_t)p * ncon + c])) { at_target = false; break; }
            if (at_target) continue;
            bool grew = false;
            while (!queues[p].empty()) {
                int32_t v = queues[p].front();
                queues[p].pop_front();
                for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
                    int32_t u = g.adjncy[j];
                    if (part[u] == -1) {
                        part[u] = p;
                        for (int32_t c = 0; c < ncon; ++c) pwt[(size_t)p * ncon + c] += g.vw(u, c);
                        queues[p].push_back(u);
                        grew = true;
                    }
                }
                if (grew) break;
            }
            if (grew) ++active;
        }
    }
    for (int32_t v = 0; v < n; ++v) {                                        // 89-100
        if (part[v] == -1) {
            int32_t best = 0;
            for (int32_t p = 1; p < nparts; ++p)
                if (pwt[(size_t)p * ncon] < pwt[(size_t)best * ncon]) best = p;  // first minimum
            part[v] = best;
            for (int32_t c = 0; c < ncon; ++c) pwt[(size_t)best * ncon + c] += g.vw(v, c);
        }
    }
    return 0;
}

// ---- refine/fm.rs:12-145 (2-way FM) -------------------------------------
//
// The reference's queue is priority-queue 1.4.0's PriorityQueue (absent
// here); among equal gains this restatement pops the smallest vertex first,
// as in the k-way FM above.
bool fm2_pass_impl(const OGraph& g, int32_t* part, std::vector<int32_t>& pw, const double* tpw, const double* ub) {
    const int32_t n = g.n, ncon = g.ncon;
    const int32_t* ew = g.ewp();
    std::vector<int32_t> gains(n);
    for (int32_t v = 0; v < n; ++v)                                          // fm.rs:44-49
        gains[v] = move_gain(g.xadj.data(), g.adjncy.data(), ew, v, part, part[v], 1 - part[v]);
    auto boundary = [&](int32_t v) {
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
            if (part[g.adjncy[j]] != part[v]) return true;
        return false;
    };
    std::set<std::pair<int32_t, int32_t>> pq;                                // (-gain, v)
    std::vector<int32_t> key(n, 0);
    std::vector<char> inq(n, 0);
    for (int32_t v = 0; v < n; ++v)                                          // fm.rs:52-57
        if (boundary(v)) {
            pq.insert({-gains[v], v});
            key[v] = gains[v];
            inq[v] = 1;
        }
    std::vector<char> locked(n, 0);
    std::vector<std::pair<int32_t, int32_t>> moves;                          // (v, from)
    int32_t cut = edge_cut_impl(n, g.xadj.data(), g.adjncy.data(), ew, part);
    int32_t best_cut = cut;
    size_t best_len = 0;
    while (!pq.empty()) {                                                    // fm.rs:66-121
        const auto top = *pq.begin();
        pq.erase(pq.begin());
        const int32_t v = top.second, gain = -top.first;
        inq[v] = 0;
        if (locked[v]) continue;
        const int32_t from = part[v], to = 1 - from;
        bool feasible = true;                                                // fm.rs:77-83
        for (int32_t c = 0; c < ncon; ++c)
            if (!((double)(pw[(size_t)to * ncon + c] + g.vw(v, c)) <= tpw[(size_t)to * ncon + c] * ub[c])) feasible = false;
        if (!feasible) continue;
        part[v] = to;                                                        // fm.rs:86-99
        for (int32_t c = 0; c < ncon; ++c) {
            pw[(size_t)from * ncon + c] -= g.vw(v, c);
            pw[(size_t)to * ncon + c] += g.vw(v, c);
        }
        locked[v] = 1;
        moves.push_back({v, from});
        cut -= gain;
        if (cut < best_cut) {
            best_cut = cut;
            best_len = moves.size();
        }
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {                // fm.rs:101-120
            const int32_t u = g.adjncy[j];
            if (locked[u]) continue;
            gains[u] = move_gain(g.xadj.data(), g.adjncy.data(), ew, u, part, part[u], 1 - part[u]);
            if (inq[u]) {
                pq.erase({-key[u], u});
                pq.insert({-gains[u], u});
                key[u] = gains[u];
            } else if (boundary(u)) {
                pq.insert({-gains[u], u});
                key[u] = gains[u];
                inq[u] = 1;
            }
        }
    }
    const bool improved = best_len > 0 && best_cut < edge_cut_impl(n, g.xadj.data(), g.adjncy.data(), ew, part);
    for (size_t i = moves.size(); i > best_len; --i) {                       // fm.rs:127-134
        const int32_t v = moves[i - 1].first, from = moves[i - 1].second, cur = part[v];
        for (int32_t c = 0; c < ncon; ++c) {
            pw[(size_t)cur * ncon + c] -= g.vw(v, c);
            pw[(size_t)from * ncon + c] += g.vw(v, c);
        }
        part[v] = from;
    }
    return improved;
}

// fm.rs:12-28
int32_t fm2_refine_impl(const OGraph& g, int32_t* part, const double* tpw, const double* ub, int32_t niter) {
    std::vector<int32_t> pw(2 * (size_t)g.ncon);
    pweights(g.n, g.ncon, g.vwp(), part, 2, pw.data());
    for (int32_t it = 0; it < niter; ++it)
        if (!fm2_pass_impl(g, part, pw, tpw, ub)) break;
    return edge_cut_impl(g.n, g.xadj.data(), g.adjncy.data(), g.ewp(), part);
}

// graph/mod.rs:209-270: the vertices of part `target` (ascending) with the
// edges between them, in CSR order; new_to_old maps the new ids back.
OGraph extract_impl(const OGraph& g, const int32_t* part, int32_t target, std::vector<int32_t>& new_to_old) {
    new_to_old.clear();
    for (int32_t v = 0; v < g.n; ++v)
        if (part[v] == target) new_to_old.push_back(v);
    std::vector<int32_t> o2n(g.n, -1);
    for (size_t i = 0; i < new_to_old.size(); ++i) o2n[new_to_old[i]] = (int32_t)i;
    OGraph s;
    s.n = (int32_t)new_to_old.size();
    s.ncon = g.ncon;
    s.has_vwgt = g.has_vwgt;
    s.has_adjwgt = g.has_adjwgt;
    s.has_vsize = g.has_vsize;
    s.xadj.assign((size_t)s.n + 1, 0);
    for (int32_t i = 0; i < s.n; ++i) {
        const int32_t v = new_to_old[i];
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
            if (part[g.adjncy[j]] == target) {
                s.adjncy.push_back(o2n[g.adjncy[j]]);
                if (g.has_adjwgt) s.adjwgt.push_back(g.adjwgt[j]);
            }
        s.xadj[i + 1] = (int32_t)s.adjncy.size();
        if (g.has_vwgt)
            for (int32_t c = 0; c < g.ncon; ++c) s.vwgt.push_back(g.vwgt[(size_t)v * g.ncon + c]);
        if (g.has_vsize) s.vsize.push_back(g.vsize[v]);
    }
    return s;
}

struct OOptions {
    int32_t ctype_shem = 1, iptype = 0, rtype_greedy = 0, niter = 10, ufactor = 30;
    uint64_t seed = 1;
};

// partition/recursive.rs:140-187 (multilevel_bisect); tpw2 = the two parts'
// fractions [2 * ncon] (f32), ub as f64.
std::vector<int32_t> bisect_impl(const OGraph& g, const float* tpw2, const double* ub, const OOptions& o,
                                 orc_rng* rng) {
    if (g.n <= 2) {                                                          // recursive.rs:147-150
        std::vector<int32_t> p(g.n);
        for (int32_t v = 0; v < g.n; ++v) p[v] = v;
        return p;
    }
    std::vector<OGraph> lg;                                                  // coarsen/mod.rs:36-71, nparts 2
    std::vector<std::vector<int32_t>> cmaps;
    const OGraph* cur = &g;
    for (;;) {
        std::vector<int32_t> mt = o.ctype_shem ? shem_impl(*cur, rng) : rm_impl(*cur, rng);
        std::vector<int32_t> cmap;
        OGraph cg = coarse_impl(*cur, mt, cmap);
        const float ratio = (float)cg.n / (float)cur->n;
        const int32_t cgn = cg.n;
        lg.push_back(std::move(cg));
        cmaps.push_back(std::move(cmap));
        if (cgn <= 40 || ratio > 0.95f) break;
        cur = &lg.back();
    }
    const OGraph& coarsest = lg.back();
    std::vector<double> ct = targets_impl(coarsest, 2, tpw2);                // recursive.rs:157-159
    std::vector<int32_t> part(coarsest.n);
    if (o.iptype == 1) {                                                     // initial/mod.rs:17-20
        for (int32_t v = 0; v < coarsest.n; ++v) part[v] = (int32_t)orc_rng_range_incl(rng, 0, 1);
    } else {
        grow_impl(coarsest, 2, ct.data(), rng, part.data());
    }
    auto refine2 = [&](const OGraph& lgph, int32_t* pt, const double* t) {   // refine/mod.rs:11-23
        if (o.rtype_greedy) greedy_refine_impl(lgph, pt, 2, t, ub, o.niter, nullptr);
        else fm2_refine_impl(lgph, pt, t, ub, o.niter);
    };
    refine2(coarsest, part.data(), ct.data());
    for (int64_t i = (int64_t)lg.size() - 1; i >= 0; --i) {                  // recursive.rs:169-184
        const OGraph& fine = i == 0 ? g : lg[i - 1];
        std::vector<int32_t> fp(fine.n);
        for (int32_t v = 0; v < fine.n; ++v) fp[v] = part[cmaps[i][v]];
        std::vector<double> lt = targets_impl(fine, 2, tpw2);
        refine2(fine, fp.data(), lt.data());
        part = std::move(fp);
    }
    return part;
}

// recursive.rs:60-137 (rb_recurse); tp = the fractions of the nparts parts.
void rb_recurse_impl(const OGraph& g, int32_t nparts, const std::vector<float>& tp, const double* ub,
                     const OOptions& o, orc_rng* rng, int32_t* part) {
    if (nparts == 1) {                                                       // recursive.rs:70-76
        std::fill(part, part + g.n, 0);
        return;
    }
    if (g.n <= nparts) {                                                     // recursive.rs:79-86
        for (int32_t v = 0; v < g.n; ++v) part[v] = std::min(v, nparts - 1);
        return;
    }
    const int32_t k1 = nparts / 2, k2 = nparts - k1, ncon = g.ncon;          // recursive.rs:89-104
    std::vector<float> lw(ncon, 0.f), rw(ncon, 0.f);
    for (int32_t c = 0; c < ncon; ++c) {
        for (int32_t p = 0; p < k1; ++p) lw[c] += tp[(size_t)p * ncon + c];
        for (int32_t p = k1; p < nparts; ++p) rw[c] += tp[(size_t)p * ncon + c];
    }
    const float total = lw[0] + rw[0];
    std::vector<float> bt(2 * (size_t)ncon);
    for (int32_t c = 0; c < ncon; ++c) {
        bt[c] = lw[c] / total;
        bt[ncon + c] = rw[c] / total;
    }
    std::vector<int32_t> bp = bisect_impl(g, bt.data(), ub, o, rng);         // recursive.rs:107
    std::vector<int32_t> n2o0, n2o1;                                          // recursive.rs:110-111
    OGraph g0 = extract_impl(g, bp.data(), 0, n2o0);
    OGraph g1 = extract_impl(g, bp.data(), 1, n2o1);
    const float s0 = lw[0], s1 = rw[0];                                      // recursive.rs:114-117
    std::vector<float> t0(tp.begin(), tp.begin() + (size_t)k1 * ncon), t1(tp.begin() + (size_t)k1 * ncon, tp.end());
    for (float& x : t0) x /= s0;
    for (float& x : t1) x /= s1;
    std::vector<int32_t> sp0(g0.n), sp1(g1.n);
    rb_recurse_impl(g0, k1, t0, ub, o, rng, sp0.data());                     // recursive.rs:123-124
    rb_recurse_impl(g1, k2, t1, ub, o, rng, sp1.data());
    for (size_t i = 0; i < n2o0.size(); ++i) part[n2o0[i]] = sp0[i];        // recursive.rs:127-132
    for (size_t i = 0; i < n2o1.size(); ++i) part[n2o1[i]] = sp1[i] + k1;
}

// recursive.rs:11-56
int32_t part_recursive_impl(const OGraph& g, int32_t nparts, const float* tpwgts, const float* ubvec,
                            const OOptions& o, int32_t* part) {
    if (nparts < 2 || nparts > g.n) return -2;                               // recursive.rs:18-26
    orc_rng rng;
    orc_rng_seed_from_u64(&rng, o.seed);                                     // recursive.rs:28-36
    std::vector<float> tp = tpwgts ? std::vector<float>(tpwgts, tpwgts + (size_t)nparts * g.ncon)
                                   : std::vector<float>((size_t)nparts * g.ncon, 1.0f / (float)nparts);
    std::vector<double> ub(g.ncon);
    for (int32_t c = 0; c < g.ncon; ++c) ub[c] = ubvec ? (double)ubvec[c] : 1.0 + (double)o.ufactor / 1000.0;
    rb_recurse_impl(g, nparts, tp, ub.data(), o, &rng, part);
    return 1;
}

// separator/mod.rs:10-61
void separator_impl(const OGraph& g, const OOptions& o, std::vector<int32_t>& left, std::vector<int32_t>& right,
                    std::vector<int32_t>& sep) {
    const float tpw[2] = {0.5f, 0.5f}, ubv[1] = {1.05f};
    std::vector<int32_t> part(g.n);
    // partition_recursive(graph, 2, Some(&tpwgts), Some(&ubvec), ..): the
    // two-entry arrays are read as [nparts * ncon] and [ncon] (ncon is 1 here)
    if (part_recursive_impl(g, 2, tpw, ubv, o, part.data()) != 1)
        for (int32_t v = 0; v < g.n; ++v) part[v] = v < g.n / 2 ? 0 : 1;    // separator/mod.rs:22-27
    std::vector<int32_t> pw(2 * (size_t)g.ncon);
    pweights(g.n, g.ncon, g.vwp(), part.data(), 2, pw.data());
    const int32_t smaller = pw[0] <= pw[(size_t)g.ncon] ? 0 : 1;
    std::vector<char> in_sep(g.n, 0);
    for (int32_t v = 0; v < g.n; ++v) {
        if (part[v] != smaller) continue;
        for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
            if (part[g.adjncy[j]] != part[v]) {
                in_sep[v] = 1;
                break;
            }
    }
    left.clear();
    right.clear();
    sep.clear();
    for (int32_t v = 0; v < g.n; ++v) {
        if (in_sep[v]) sep.push_back(v);
        else if (part[v] == smaller) left.push_back(v);
        else right.push_back(v);
    }
}

// nd.rs:48-120 (nd_recurse)
void nd_recurse_impl(const OGraph& g, const std::vector<int32_t>& orig, const OOptions& o, int32_t* perm,
                     int32_t& counter) {
    if (g.n <= 100) {                                                        // nd.rs:59-66
        for (int32_t v : orig) perm[v] = counter++;
        return;
    }
    std::vector<int32_t> left, right, sep;
    separator_impl(g, o, left, right, sep);
    if ((int64_t)sep.size() >= g.n / 2) {                                    // nd.rs:77-83
        for (int32_t v : orig) perm[v] = counter++;
        return;
    }
    std::vector<int32_t> side(g.n, 2);                                        // nd.rs:86-90
    for (int32_t v : left) side[v] = 0;
    for (int32_t v : right) side[v] = 1;
    for (int32_t t = 0; t < 2; ++t) {                                        // nd.rs:92-102
        if ((t == 0 ? left : right).empty()) continue;
        std::vector<int32_t> n2o;
        OGraph sg = extract_impl(g, side.data(), t, n2o);
        std::vector<int32_t> so(n2o.size());
        for (size_t i = 0; i < n2o.size(); ++i) so[i] = orig[n2o[i]];
        nd_recurse_impl(sg, so, o, perm, counter);
    }
    for (int32_t v : sep) perm[orig[v]] = counter++;                         // nd.rs:105-108
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

/* ---- rand 0.8.5 SmallRng = Xoshiro256PlusPlus (rand/src/rngs/xoshiro256plusplus.rs);
 *      seed_from_u64 fills the 32-byte seed with SplitMix64 outputs. ---- */
uint64_t orc_splitmix64_next(uint64_t* state) {
    *state += 0x9e3779b97f4a7c15ULL;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
void orc_rng_from_words(orc_rng* r, const uint64_t w[4]) {
    if (!w[0] && !w[1] && !w[2] && !w[3]) { orc_rng_seed_from_u64(r, 0); return; }
    for (int i = 0; i < 4; ++i) r->s[i] = w[i];
}
void orc_rng_seed_from_u64(orc_rng* r, uint64_t seed) {
    uint64_t st = seed, w[4];
    for (int i = 0; i < 4; ++i) w[i] = orc_splitmix64_next(&st);
    orc_rng_from_words(r, w);
}
uint64_t orc_rng_next_u64(orc_rng* r) {
    uint64_t* s = r->s;
    uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}
uint32_t orc_rng_next_u32(orc_rng* r) { return (uint32_t)(orc_rng_next_u64(r) >> 32); }
/* rand 0.8.5 UniformInt<usize>::sample_single_inclusive (widening multiply,
 * rejection zone (range << lz) - 1). gen_range(a..b) == range_incl(a, b-1). */
uint64_t orc_rng_range_incl(orc_rng* r, uint64_t low, uint64_t high) {
    uint64_t range = high - low + 1;
    if (range == 0) return orc_rng_next_u64(r);
    int lz = __builtin_clzll(range);
    uint64_t zone = (range << lz) - 1;
    for (;;) {
        uint64_t v = orc_rng_next_u64(r);
        unsigned __int128 prod = (unsigned __int128)v * range;
        uint64_t hi = (uint64_t)(prod >> 64), lo = (uint64_t)prod;
        if (lo <= zone) return low + hi;
    }
}

int32_t orc_edge_cut(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                     const int32_t* adjwgt, const int32_t* part) {
    return edge_cut_impl(n, xadj, adjncy, adjwgt, part);
}

void orc_partition_weights(int32_t n, int32_t ncon, const int32_t* vwgt, const int32_t* part,
                           int32_t nparts, int32_t* pw) {
    pweights(n, ncon, vwgt, part, nparts, pw);
}

void orc_total_weight(int32_t n, int32_t ncon, const int32_t* vwgt, int32_t* tw) {
    for (int32_t c = 0; c < ncon; ++c) tw[c] = 0;
    for (int32_t v = 0; v < n; ++v)
        for (int32_t c = 0; c < ncon; ++c) tw[c] = wadd(tw[c], vwgt ? vwgt[(size_t)v * ncon + c] : 1);
}

int32_t orc_boundary_vertices(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                              const int32_t* part, int32_t* out) {
    std::vector<int32_t> b = boundary_impl(n, xadj, adjncy, part);
    if (out) std::copy(b.begin(), b.end(), out);
    return (int32_t)b.size();
}

void orc_compute_target_weights(int32_t n, int32_t ncon, const int32_t* vwgt, int32_t nparts,
                                const float* tpwgts, double* out) {
    OGraph g;
    g.n = n;
    g.ncon = ncon;
    if (vwgt) { g.has_vwgt = true; g.vwgt.assign(vwgt, vwgt + (size_t)n * ncon); }
    std::vector<double> t = targets_impl(g, nparts, tpwgts);
    std::copy(t.begin(), t.end(), out);
}

int32_t orc_validate_csr(int32_t n, int32_t ncon, const int32_t* xadj, int64_t xadj_len,
                         const int32_t* adjncy, int64_t adjncy_len, const int32_t* vwgt,
                         int64_t vwgt_len, const int32_t* adjwgt, int64_t adjwgt_len) {
    // graph/validate.rs:16-142; returns 1 (ok) or -2 (MetisError::Input)
    if (n <= 0 || ncon <= 0) return -2;
    if (xadj_len != (int64_t)n + 1) return -2;
    if (xadj[0] != 0) return -2;
    for (int32_t i = 0; i < n; ++i)
        if (xadj[i] > xadj[i + 1]) return -2;
    int64_t ne = xadj[n];
    if (adjncy_len != ne) return -2;
    for (int32_t v = 0; v < n; ++v)
        for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
            int32_t u = adjncy[j];
            if (u < 0 || u >= n) return -2;
            if (u == v) return -2;
        }
    std::vector<int64_t> cnt(n, 0);
    for (int64_t j = 0; j < ne; ++j) cnt[adjncy[j]]++;
    for (int32_t v = 0; v < n; ++v)
        if (cnt[v] != (int64_t)(xadj[v + 1] - xadj[v])) return -2;
    if (adjwgt && adjwgt_len != ne) return -2;
    if (vwgt) {
        if (vwgt_len != (int64_t)n * ncon) return -2;
        for (int64_t i = 0; i < vwgt_len; ++i)
            if (vwgt[i] <= 0) return -2;
    }
    return 1;
}

int32_t orc_compute_move_gain(const int32_t* xadj, const int32_t* adjncy, const int32_t* adjwgt,
                              int32_t v, const int32_t* part, int32_t from, int32_t to) {
    return move_gain(xadj, adjncy, adjwgt, v, part, from, to);
}

int32_t orc_greedy_pass(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                        const int32_t* vwgt, const int32_t* adjwgt, int32_t* part,
                        int32_t nparts, const double* tpwgts, const double* ubvec,
                        int32_t mode, int32_t* pw, int32_t* moves_out) {
    std::vector<int32_t> moves;
    int32_t nm = greedy_pass_impl(n, ncon, xadj, adjncy, vwgt, adjwgt, part, nparts, tpwgts, ubvec,
                                  mode, pw, &moves, nullptr);
    if (moves_out) std::copy(moves.begin(), moves.end(), moves_out);
    return nm;
}

int32_t orc_greedy_refine(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                          const int32_t* vwgt, const int32_t* adjwgt, int32_t* part,
                          int32_t nparts, const double* tpwgts, const double* ubvec,
                          int32_t niter) {
    OGraph g;
    g.n = n;
    g.ncon = ncon;
    g.xadj.assign(xadj, xadj + n + 1);
    g.adjncy.assign(adjncy, adjncy + xadj[n]);
    if (vwgt) { g.has_vwgt = true; g.vwgt.assign(vwgt, vwgt + (size_t)n * ncon); }
    if (adjwgt) { g.has_adjwgt = true; g.adjwgt.assign(adjwgt, adjwgt + xadj[n]); }
    return greedy_refine_impl(g, part, nparts, tpwgts, ubvec, niter, nullptr);
}

int32_t orc_fm_kway_refine(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                           const int32_t* vwgt, const int32_t* adjwgt, int32_t* part, int32_t nparts,
                           const double* tpwgts, const double* ubvec, int32_t niter) {
    OGraph g;
    g.n = n;
    g.ncon = ncon;
    g.xadj.assign(xadj, xadj + n + 1);
    g.adjncy.assign(adjncy, adjncy + xadj[n]);
    if (vwgt) { g.has_vwgt = true; g.vwgt.assign(vwgt, vwgt + (size_t)n * ncon); }
    if (adjwgt) { g.has_adjwgt = true; g.adjwgt.assign(adjwgt, adjwgt + xadj[n]); }
    return fm_refine_impl(g, part, nparts, tpwgts, ubvec, niter, nullptr);
}

int32_t orc_fm_kway_pass(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                         const int32_t* vwgt, const int32_t* adjwgt, int32_t* part, int32_t nparts,
                         const double* tpwgts, const double* ubvec, int32_t* pw, int32_t* nmoves) {
    OGraph g;
    g.n = n;
    g.ncon = ncon;
    g.xadj.assign(xadj, xadj + n + 1);
    g.adjncy.assign(adjncy, adjncy + xadj[n]);
    if (vwgt) { g.has_vwgt = true; g.vwgt.assign(vwgt, vwgt + (size_t)n * ncon); }
    if (adjwgt) { g.has_adjwgt = true; g.adjwgt.assign(adjwgt, adjwgt + xadj[n]); }
    std::vector<int32_t> w(pw, pw + (size_t)nparts * ncon);
    const bool improved = fm_pass_impl(g, part, nparts, w, tpwgts, ubvec, nmoves);
    std::copy(w.begin(), w.end(), pw);
    return improved ? 1 : 0;
}

static OGraph make_graph(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                         const int32_t* vwgt, const int32_t* adjwgt, const int32_t* vsize) {
    OGraph g;
    g.n = n;
    g.ncon = ncon;
    g.xadj.assign(xadj, xadj + n + 1);
    g.adjncy.assign(adjncy, adjncy + xadj[n]);
    if (vwgt) { g.has_vwgt = true; g.vwgt.assign(vwgt, vwgt + (size_t)n * ncon); }
    if (adjwgt) { g.has_adjwgt = true; g.adjwgt.assign(adjwgt, adjwgt + xadj[n]); }
    if (vsize) { g.has_vsize = true; g.vsize.assign(vsize, vsize + n); }
    return g;
}

void orc_shem_matching(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                       const int32_t* adjwgt, orc_rng* rng, int32_t* matching) {
    OGraph g = make_graph(n, 1, xadj, adjncy, nullptr, adjwgt, nullptr);
    std::vector<int32_t> m = shem_impl(g, rng);
    std::copy(m.begin(), m.end(), matching);
}

void orc_random_matching(int32_t n, const int32_t* xadj, const int32_t* adjncy, orc_rng* rng,
                         int32_t* matching) {
    OGraph g = make_graph(n, 1, xadj, adjncy, nullptr, nullptr, nullptr);
    std::vector<int32_t> m = rm_impl(g, rng);
    std::copy(m.begin(), m.end(), matching);
}

int32_t orc_build_coarse_graph(int32_t n, int32_t ncon, const int32_t* xadj,
                               const int32_t* adjncy, const int32_t* vwgt,
                               const int32_t* adjwgt, const int32_t* vsize,
                               const int32_t* matching, int32_t* cxadj, int32_t* cadjncy,
                               int32_t* cadjwgt, int32_t* cvwgt, int32_t* cvsize,
                               int32_t* cmap, int32_t* cm, int32_t* flags) {
    OGraph g = make_graph(n, ncon, xadj, adjncy, vwgt, adjwgt, vsize);
    std::vector<int32_t> mt(matching, matching + n), cm_v;
    OGraph c = coarse_impl(g, mt, cm_v);
    std::copy(c.xadj.begin(), c.xadj.end(), cxadj);
    std::copy(c.adjncy.begin(), c.adjncy.end(), cadjncy);
    if (c.has_adjwgt) std::copy(c.adjwgt.begin(), c.adjwgt.end(), cadjwgt);
    else std::fill(cadjwgt, cadjwgt + c.xadj[c.n], 1);
    if (c.has_vwgt && cvwgt) std::copy(c.vwgt.begin(), c.vwgt.end(), cvwgt);
    if (c.has_vsize && cvsize) std::copy(c.vsize.begin(), c.vsize.end(), cvsize);
    std::copy(cm_v.begin(), cm_v.end(), cmap);
    *cm = c.xadj[c.n];
    *flags = (c.has_vwgt ? 1 : 0) | (c.has_vsize ? 2 : 0) | (c.has_adjwgt ? 4 : 0);
    return c.n;
}

int32_t orc_grow_partition(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                           const int32_t* vwgt, int32_t nparts, const double* tpwgts,
                           orc_rng* rng, int32_t* part) {
    OGraph g = make_graph(n, ncon, xadj, adjncy, vwgt, nullptr, nullptr);
    return grow_impl(g, nparts, tpwgts, rng, part);
}

int32_t orc_part_graph_kway(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                            const int32_t* vwgt, const int32_t* vsize, const int32_t* adjwgt,
                            int32_t nparts, const float* tpwgts, const float* ubvec,
                            int32_t ctype_shem, int32_t iptype, int32_t rtype_greedy,
                            int32_t ncuts, int32_t niter, uint64_t seed, int32_t ufactor,
                            int32_t* objval, int32_t* part, orc_stats* stats) {
    // partition/kway.rs:20-141 (comm = None => SingleComm)
    if (nparts < 2 || nparts > n) return -2;                                 // kway.rs:31-38
    OGraph g = make_graph(n, ncon, xadj, adjncy, vwgt, adjwgt, vsize);
    std::vector<double> ub(ncon);                                            // kway.rs:44-49
    for (int32_t c = 0; c < ncon; ++c)
        ub[c] = ubvec ? (double)ubvec[c] : 1.0 + (double)ufactor / 1000.0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    bool have_best = false;
    int32_t best_obj = 0;
    std::vector<int32_t> best_part;
    const int32_t limit = std::max(20 * nparts, 20);                         // coarsen/mod.rs:43
    for (int32_t attempt = 0; attempt < ncuts; ++attempt) {                  // kway.rs:54
        uint64_t s = seed;
        if (seed == 0) s = std::random_device{}() ^ ((uint64_t)std::random_device{}() << 32);
        else s = seed + (uint64_t)attempt;                                   // kway.rs:55-61
        orc_rng rng;
        orc_rng_seed_from_u64(&rng, s);
        // coarsen/mod.rs:36-71
        double t0 = now_s();
        std::vector<OGraph> lg;
        std::vector<std::vector<int32_t>> cmaps;
        const OGraph* cur = &g;
        for (;;) {
            std::vector<int32_t> mt = ctype_shem ? shem_impl(*cur, &rng) : rm_impl(*cur, &rng);
            std::vector<int32_t> cmap;
            OGraph cg = coarse_impl(*cur, mt, cmap);
            if (stats) stats->coarsen_edges += cur->xadj[cur->n];
            float ratio = (float)cg.n / (float)cur->n;
            int32_t cgn = cg.n;
            lg.push_back(std::move(cg));
            cmaps.push_back(std::move(cmap));
            if (cgn <= limit || ratio > 0.95f) break;
            cur = &lg.back();
        }
        double t1 = now_s();
        const OGraph& coarsest = lg.back();
        std::vector<double> ct = targets_impl(coarsest, nparts, tpwgts);     // kway.rs:67-69
        std::vector<int32_t> cpart(coarsest.n);
        if (iptype == 1) {                                                   // initial/random.rs:8-10
            for (int32_t v = 0; v < coarsest.n; ++v)
                cpart[v] = (int32_t)orc_rng_range_incl(&rng, 0, (uint64_t)nparts - 1);
        } else if (grow_impl(coarsest, nparts, ct.data(), &rng, cpart.data()) != 0) {
            return -4;
        }
        double t2 = now_s();
        refine_level_impl(coarsest, cpart.data(), nparts, ct.data(), ub.data(), niter, rtype_greedy,
                          stats ? &stats->refine_edges : nullptr);           // kway.rs:73
        std::vector<int32_t> pcur = std::move(cpart);
        for (int64_t i = (int64_t)lg.size() - 1; i >= 0; --i) {              // kway.rs:76-92
            const OGraph& fine = (i == 0) ? g : lg[i - 1];
            std::vector<int32_t> fp(fine.n);
            for (int32_t v = 0; v < fine.n; ++v) fp[v] = pcur[cmaps[i][v]];
            std::vector<double> lt = targets_impl(fine, nparts, tpwgts);
            refine_level_impl(fine, fp.data(), nparts, lt.data(), ub.data(), niter, rtype_greedy,
                              stats ? &stats->refine_edges : nullptr);   // kway.rs:90
            pcur = std::move(fp);
        }
        double t3 = now_s();
        int32_t obj = edge_cut_impl(g.n, g.xadj.data(), g.adjncy.data(), g.ewp(), pcur.data());
        if (!have_best || obj < best_obj) {                                  // kway.rs:96-98
            have_best = true;
            best_obj = obj;
            best_part = std::move(pcur);
        }
        if (stats) {
            stats->t_coarsen_s += t1 - t0;
            stats->t_initial_s += t2 - t1;
            stats->t_refine_s += t3 - t2;
            stats->nlevels = (int32_t)lg.size();
            stats->coarsest_n = coarsest.n;
        }
    }
    std::copy(best_part.begin(), best_part.end(), part);
    if (objval) *objval = best_obj;
    return 1;
}

int32_t orc_fm2_refine(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy, const int32_t* vwgt,
                       const int32_t* adjwgt, int32_t* part, const double* tpw, const double* ub, int32_t niter) {
    OGraph g = make_graph(n, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
    return fm2_refine_impl(g, part, tpw, ub, niter);
}

int32_t orc_fm2_pass(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy, const int32_t* vwgt,
                     const int32_t* adjwgt, int32_t* part, const double* tpw, const double* ub) {
    OGraph g = make_graph(n, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
    std::vector<int32_t> pw(2 * (size_t)g.ncon);
    pweights(g.n, g.ncon, g.vwp(), part, 2, pw.data());
    return fm2_pass_impl(g, part, pw, tpw, ub) ? 1 : 0;
}

int32_t orc_part_graph_recursive(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                                 const int32_t* vwgt, const int32_t* adjwgt, int32_t nparts, const float* tpwgts,
                                 const float* ubvec, int32_t ctype_shem, int32_t iptype, int32_t rtype_greedy,
                                 int32_t niter, uint64_t seed, int32_t ufactor, int32_t* objval, int32_t* part) {
    OGraph g = make_graph(n, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
    OOptions o;
    o.ctype_shem = ctype_shem;
    o.iptype = iptype;
    o.rtype_greedy = rtype_greedy;
    o.niter = niter;
    o.ufactor = ufactor;
    o.seed = seed;
    const int32_t rc = part_recursive_impl(g, nparts, tpwgts, ubvec, o, part);
    if (rc == 1 && objval) *objval = edge_cut_impl(g.n, g.xadj.data(), g.adjncy.data(), g.ewp(), part);
    return rc;
}

int32_t orc_node_nd(int32_t n, const int32_t* xadj, const int32_t* adjncy, const int32_t* vwgt, int32_t ctype_shem,
                    int32_t iptype, int32_t rtype_greedy, int32_t niter, uint64_t seed, int32_t ufactor,
                    int32_t* perm, int32_t* iperm) {
    // nd.rs:17-46: ordering positions assigned bottom-up, then inverted
    OGraph g = make_graph(n, 1, xadj, adjncy, vwgt, nullptr, nullptr);
    OOptions o;
    o.ctype_shem = ctype_shem;
    o.iptype = iptype;
    o.rtype_greedy = rtype_greedy;
    o.niter = niter;
    o.ufactor = ufactor;
    o.seed = seed;
    std::vector<int32_t> orig(n);
    for (int32_t v = 0; v < n; ++v) orig[v] = v;
    int32_t counter = 0;
    nd_recurse_impl(g, orig, o, perm, counter);
    for (int32_t v = 0; v < n; ++v) iperm[perm[v]] = v;
    return 1;
}

}  // extern "C"
